"""Pins the CPU oracle (oracle/afqmc_oracle.py) against closed forms,
invariants, finite differences of the overlap and brute force.

The reference's own tests hold no numeric fixtures (they are closed-form and
invariant checks, tests/test_*.py in the reference); the ones that apply to the
restated functions are restated here, next to stronger independent checks.
"""
import math

import numpy as np
import pytest
import scipy.linalg

from oracle import afqmc_oracle as orc


def _rand_orthonormal(rng, n, k):
    a = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    return np.linalg.qr(a)[0]


def _small_ham(rng, n, g, sym=True, h0=0.7):
    a = rng.standard_normal((n, n))
    h1 = 0.5 * (a + a.T)
    b = rng.standard_normal((g, n, n))
    chol = 0.5 * (b + b.transpose(0, 2, 1)) if sym else 0.3 * b
    return h0, h1, chol


def _walkers(rng, kind, nw, n, o):
    if kind == "restricted":
        return np.stack([_rand_orthonormal(rng, n, o) for _ in range(nw)])
    return (np.stack([_rand_orthonormal(rng, n, o) for _ in range(nw)]),
            np.stack([_rand_orthonormal(rng, n, o) for _ in range(nw)]))


def _trial(rng, n, o, uhf):
    cu = _rand_orthonormal(rng, n, o)
    cd = _rand_orthonormal(rng, n, o) if uhf else cu
    return (cu, cd)


def _pick(w, i):
    return w[i] if not isinstance(w, tuple) else (w[0][i], w[1][i])


CASES = [("restricted", False), ("unrestricted", False), ("unrestricted", True)]


@pytest.mark.parametrize("kind,uhf", CASES)
@pytest.mark.parametrize("sym", [True, False])
def test_force_bias_is_log_overlap_derivative(kind, uhf, sym):
    rng = np.random.default_rng(11)
    n, o, g = 6, 2, 5
    h0, h1, chol = _small_ham(rng, n, g, sym=sym)
    trial = _trial(rng, n, o, uhf)
    w = _walkers(rng, kind, 3, n, o)
    mctx = orc.build_meas_ctx(h1, chol, trial)
    fb = orc.force_bias(w, trial, mctx, kind)
    for i in range(3):
        ref = orc.force_bias_fd(_pick(w, i), trial, chol, kind)
        np.testing.assert_allclose(fb[i], ref, rtol=1e-8, atol=1e-9)


def test_force_bias_gauge_invariant_and_u_equals_r():
    """fb(w A) == fb(w); fb_u((w, w)) == fb_r(w) (reference tests/test_trial_rhf.py:69-85)."""
    rng = np.random.default_rng(2)
    n, o, g = 6, 3, 3
    h0, h1, chol = _small_ham(rng, n, g)
    trial = _trial(rng, n, o, False)
    mctx = orc.build_meas_ctx(h1, chol, trial)
    w = _walkers(rng, "restricted", 2, n, o)
    a = rng.standard_normal((o, o)) + 1.5 * np.eye(o) + 0.3j * rng.standard_normal((o, o))
    fb = orc.force_bias(w, trial, mctx, "restricted")
    np.testing.assert_allclose(orc.force_bias(w @ a, trial, mctx, "restricted"), fb, rtol=1e-11)
    np.testing.assert_allclose(orc.force_bias((w, w), trial, mctx, "unrestricted"), fb, rtol=1e-12)


@pytest.mark.parametrize("kind,uhf", CASES)
def test_energy_matches_overlap_derivatives(kind, uhf):
    """Manual (Green's function) energy == overlap-derivative energy
    (reference tests/test_auto.py:101-133, here to 1e-7 instead of 5e-3)."""
    rng = np.random.default_rng(5)
    n, o, g = 5, 2, 7
    h0, h1, chol = _small_ham(rng, n, g)
    trial = _trial(rng, n, o, uhf)
    w = _walkers(rng, kind, 3, n, o)
    mctx = orc.build_meas_ctx(h1, chol, trial)
    e = orc.energy(w, h0, trial, mctx, kind)
    for i in range(3):
        ref = orc.energy_fd(_pick(w, i), h0, h1, chol, trial, kind)
        np.testing.assert_allclose(e[i], ref, rtol=1e-7, atol=1e-7)


def test_energy_is_h0_when_h1_and_chol_zero():
    """reference tests/test_trial_rhf.py:88-110."""
    rng = np.random.default_rng(3)
    n, o, g = 5, 2, 4
    trial = _trial(rng, n, o, False)
    h1 = np.zeros((n, n))
    chol = np.zeros((g, n, n))
    mctx = orc.build_meas_ctx(h1, chol, trial)
    w = trial[0] @ (np.eye(o) * (1.1 + 0.2j))
    np.testing.assert_allclose(orc.energy(w[None], 1.234, trial, mctx, "restricted"), 1.234)
    np.testing.assert_allclose(orc.energy((w[None], w[None]), 1.234, trial, mctx, "unrestricted"), 1.234)


def test_energy_hartree_fock_closed_form():
    """w = C (walker = trial): E = h0 + 2 sum_i h_ii + sum_g [2 (sum_i L_ii)^2 - sum_ij L_ij L_ji]
    in the MO basis -- the textbook closed-shell Hartree-Fock energy."""
    rng = np.random.default_rng(7)
    n, o, g = 8, 3, 6
    h0, h1, chol = _small_ham(rng, n, g)
    c = np.linalg.qr(rng.standard_normal((n, o)))[0]
    trial = (c, c)
    mctx = orc.build_meas_ctx(h1, chol, trial)
    hmo = c.T @ h1 @ c
    lmo = np.einsum("pi,gpq,qj->gij", c, chol, c)
    ref = h0 + 2 * np.trace(hmo) + np.sum(2 * np.einsum("gii->g", lmo) ** 2
                                          - np.einsum("gij,gji->g", lmo, lmo))
    e = orc.energy((c + 0j)[None], h0, trial, mctx, "restricted")
    np.testing.assert_allclose(e[0], ref, rtol=1e-12)
    # gauge invariance of the mixed estimator
    a = rng.standard_normal((o, o)) + 2 * np.eye(o) + 0.5j * rng.standard_normal((o, o))
    np.testing.assert_allclose(orc.energy((c @ a)[None], h0, trial, mctx, "restricted")[0], ref, rtol=1e-11)


def test_energy_zero_variance_noninteracting():
    """For L = 0 and a trial = exact ground state of h1, <T|H = E0 <T| so the
    mixed estimator equals h0 + 2 sum_{i<o} eps_i for ANY walker."""
    rng = np.random.default_rng(8)
    n, o = 7, 3
    a = rng.standard_normal((n, n))
    h1 = 0.5 * (a + a.T)
    eps, v = np.linalg.eigh(h1)
    c = v[:, :o]
    trial = (c, c)
    chol = np.zeros((2, n, n))
    mctx = orc.build_meas_ctx(h1, chol, trial)
    w = _walkers(rng, "restricted", 4, n, o)
    np.testing.assert_allclose(orc.energy(w, 0.3, trial, mctx, "restricted").real,
                               0.3 + 2 * eps[:o].sum(), rtol=1e-11)


def test_overlap_raw_log_and_gauge():
    rng = np.random.default_rng(4)
    n, o = 6, 3
    c = _rand_orthonormal(rng, n, o)
    a = rng.standard_normal((o, o)) + 1j * rng.standard_normal((o, o))
    w = (c @ a)[None]
    ov = orc.overlap(w, (c, c), "restricted")
    np.testing.assert_allclose(ov[0], np.linalg.det(a) ** 2, rtol=1e-12)
    lo = orc.overlap(w, (c, c), "restricted", mode="log")
    np.testing.assert_allclose(np.exp(lo), ov, rtol=1e-12)
    wu = _walkers(rng, "restricted", 2, n, o)
    wd = _walkers(rng, "restricted", 2, n, o)
    ovu = orc.overlap((wu, wd), (c, c), "unrestricted")
    ref = np.linalg.det(np.einsum("pi,wpj->wij", c.conj(), wu)) * np.linalg.det(np.einsum("pi,wpj->wij", c.conj(), wd))
    np.testing.assert_allclose(ovu, ref, rtol=1e-12)


def test_taylor_matches_expm_within_truncation_bound():
    rng = np.random.default_rng(9)
    n, o = 10, 3
    v = rng.standard_normal((n, n)) * 0.05
    v = v + v.T
    w = rng.standard_normal((n, o)) + 1j * rng.standard_normal((n, o))
    a = 1j * 0.3
    out = orc.taylor_expm_action(a, v, w, 6)
    ref = scipy.linalg.expm(a * v) @ w
    x = np.linalg.norm(a * v, 2)
    bound = x ** 7 / 5040 * np.exp(x) * np.linalg.norm(w, 2)
    assert np.linalg.norm(out - ref, 2) <= bound * 1.0001 + 1e-14
    # exact polynomial identity: 6 terms == sum_k (aV)^k/k!
    ref6 = sum(np.linalg.matrix_power(a * v, k) / math.factorial(k) for k in range(7)) @ w
    np.testing.assert_allclose(out, ref6, rtol=1e-12, atol=1e-14)


def test_prop_ctx_closed_forms():
    rng = np.random.default_rng(10)
    n, o, g = 6, 2, 4
    h0, h1, chol = _small_ham(rng, n, g)
    c = np.linalg.qr(rng.standard_normal((n, o)))[0]
    rdm1 = orc.get_rdm1((c, c))
    dt = 0.05
    p = orc.build_prop_ctx(h0, h1, chol, rdm1, dt)
    mf_ref = 1j * np.array([2 * np.trace(chol[i] @ c @ c.T) for i in range(g)])
    np.testing.assert_allclose(p.mf_shifts, mf_ref, rtol=1e-12)
    np.testing.assert_allclose(p.h0_prop, -h0 + 0.5 * np.sum(mf_ref.imag ** 2), rtol=1e-12)
    h1e = h1 - 0.5 * sum(chol[i] @ chol[i] for i in range(g)) + sum(mf_ref[i].imag * chol[i] for i in range(g))
    ev, vec = np.linalg.eigh(0.5 * (h1e + h1e.T))
    np.testing.assert_allclose(p.exp_h1_half, vec @ np.diag(np.exp(-0.5 * dt * ev)) @ vec.T, rtol=1e-10, atol=1e-13)


def test_trotter_without_two_body_is_exp_h1():
    rng = np.random.default_rng(12)
    n, o, g, nw = 6, 2, 3, 4
    a = rng.standard_normal((n, n))
    h1 = 0.5 * (a + a.T)
    chol = np.zeros((g, n, n))
    c = np.linalg.qr(rng.standard_normal((n, o)))[0]
    p = orc.build_prop_ctx(0.0, h1, chol, orc.get_rdm1((c, c)), 0.1)
    w = _walkers(rng, "restricted", nw, n, o)
    x = rng.standard_normal((nw, g)) + 0j
    out = orc.apply_trotter(w, x, p, chol, 6, "restricted")
    np.testing.assert_allclose(out, scipy.linalg.expm(-0.1 * h1) @ w, rtol=1e-11, atol=1e-13)


def test_weight_update_closed_form():
    """h1 = chol = 0, h0 = 1: weights -> exp(-dt), pop shift -> 0.1
    (reference tests/test_prop_afqmc_step.py:46-94)."""
    n, o, g, nw = 4, 2, 3, 8
    h1 = np.zeros((n, n))
    chol = np.zeros((g, n, n))
    c = np.eye(n)[:, :o]
    sd = orc.make_system_data(1.0, h1, chol, (c, c), "restricted", 0.2)
    prm = orc.params(dt=0.2, n_exp_terms=4, pop_control_damping=0.1)
    w = np.broadcast_to(c + 0j, (nw, n, o)).copy()
    st = orc.state(w, np.ones(nw), orc.overlap(w, sd.trial, "restricted"), 0.0, 0.0)
    out = orc.afqmc_step(st, np.random.default_rng(0).standard_normal((nw, g)), sd, prm)
    np.testing.assert_allclose(out.weights, np.exp(-0.2) * np.ones(nw), rtol=1e-14)
    np.testing.assert_allclose(out.pop_control_ene_shift, 0.1, rtol=1e-12)


def test_phaseless_cut_floor_cap_and_nan():
    p = orc.prop_ctx(dt=0.01, sqrt_dt=0.1, exp_h1_half=np.eye(2), mf_shifts=np.zeros(1) + 0j, h0_prop=0j)
    prm = orc.params(dt=0.01)
    ov_old = np.ones(5, dtype=complex)
    #            ok      cos<0          > cap        < floor      NaN
    ov_new = np.array([1.2, -1.0 + 0.1j, 500.0, 1e-5, np.nan + 0j])
    w, pop = orc.phaseless_update(ov_new, ov_old, np.zeros(5, complex), np.zeros(5, complex),
                                  np.ones(5), 0.0, 0.0, p, prm)
    assert w[0] == pytest.approx(1.2)
    assert w[1] == 0.0 and w[2] == 0.0 and w[3] == 0.0
    assert np.isnan(w[4]) and np.isnan(pop)


def test_qr_phase_convention():
    rng = np.random.default_rng(13)
    w = np.stack([rng.standard_normal((9, 4)) + 1j * rng.standard_normal((9, 4)) for _ in range(3)])
    q, det_r = orc.qr_phase_fixed(w)
    for i in range(3):
        np.testing.assert_allclose(q[i].conj().T @ q[i], np.eye(4), atol=1e-13)
        r = q[i].conj().T @ w[i]
        np.testing.assert_allclose(np.tril(r, -1), 0, atol=1e-12)
        d = np.diag(r)
        assert np.all(d.real > 0) and np.allclose(d.imag, 0, atol=1e-12)
        np.testing.assert_allclose(det_r[i], np.prod(d), rtol=1e-12)


def _comb_bruteforce(weights, zeta):
    nw = len(weights)
    cw = []
    s = 0.0
    for x in weights:
        s = s + abs(x)
        cw.append(s)
    tot = cw[-1]
    idx = []
    for k in range(nw):
        z = tot * (k + zeta) / nw
        j = 0
        while j < nw and cw[j] < z:
            j += 1
        idx.append(j)
    return np.array(idx)


def test_sr_indices_bruteforce_and_zero_weights():
    rng = np.random.default_rng(14)
    for nw in (1, 7, 64, 257):
        w = np.abs(rng.standard_normal(nw))
        w[rng.random(nw) < 0.3] = 0.0
        z = rng.random()
        np.testing.assert_array_equal(orc.sr_indices(w, z, nw), _comb_bruteforce(w, z))
    np.testing.assert_array_equal(orc.sr_indices(np.zeros(16), 0.4, 16), np.zeros(16))
    # a walker of weight 0 is never selected; the walker counts are floor/ceil of nw*w/sum
    w = np.array([0.0, 3.0, 0.0, 1.0])
    idx = orc.sr_indices(w, 0.5, 4)
    assert set(idx.tolist()) <= {1, 3}
    assert np.sum(idx == 1) == 3 and np.sum(idx == 3) == 1


def test_block_equals_cadenced_run():
    from ad_afqmc_prototype_amd import synthetic as syn
    s = syn.make_system(0, 8, 2, 16)
    sd = orc.make_system_data(s.h0, s.h1, s.chol, (s.mo_up, s.mo_dn), "restricted", 0.01)
    prm = orc.params(dt=0.01, n_prop_steps=4)
    st0 = orc.init_state(sd, 8)
    f = [syn.fields_for_step(0, t, 8, 16) for t in range(4)]
    st_a, obs = orc.afqmc_block(st0, f, 0.37, sd, prm)
    st_b, blocks = orc.run_cadenced(st0, sd, prm, n_steps=4, qr_every=4, sr_every=4,
                                    fields_fn=lambda t: f[t], zeta_fn=lambda k: 0.37)
    np.testing.assert_array_equal(st_a.walkers, st_b.walkers)
    assert obs["energy"] == blocks[0]["energy"]
    np.testing.assert_array_equal(obs["sr_indices"], blocks[0]["sr_indices"])


def test_log_overlap_mode_matches_raw_trajectory():
    from ad_afqmc_prototype_amd import synthetic as syn
    s = syn.make_system(0, 24, 6, 96)
    sd = orc.make_system_data(s.h0, s.h1, s.chol, (s.mo_up, s.mo_dn), "restricted", 0.01)
    prm = orc.params(dt=0.01)
    st_r = orc.init_state(sd, 16)
    st_l = orc.init_state(sd, 16, overlap_mode="log")
    for t in range(5):
        f = syn.fields_for_step(0, t, 16, 96)
        st_r = orc.afqmc_step(st_r, f, sd, prm)
        st_l = orc.afqmc_step(st_l, f, sd, prm, overlap_mode="log")
    np.testing.assert_allclose(st_l.weights, st_r.weights, rtol=1e-10)
    np.testing.assert_allclose(np.exp(st_l.overlaps), st_r.overlaps, rtol=1e-10)
    assert np.mean(st_r.weights > 0) > 0.8  # the small workload stays in the working regime
