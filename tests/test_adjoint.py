"""CPU checks of the differentiable restatement that the HIP adjoint is compared
with (oracle/adjoint_oracle.py): its energy estimate equals the numpy oracle's
run on the same fields and comb uniforms, and its autograd gradients equal
central finite differences of the numpy oracle (common random numbers)."""
import numpy as np
import pytest
import torch

from ad_afqmc_prototype_amd import synthetic as syn
from oracle import adjoint_oracle as ado
from oracle import afqmc_oracle as orc


def _tiny(seed=11, n=8, o=2, g=16, nw=16, n_blocks=2, n_prop=3, dt=0.01, scale=0.5):
    h0, h1, chol = syn.make_hamiltonian(seed, n, g, chol_scale=scale)
    h0 = 1.5
    c = np.linalg.eigh(h1)[1][:, :o].copy()
    rng = np.random.default_rng(seed + 100)
    fields = rng.standard_normal((n_blocks * n_prop, nw, g))
    zetas = rng.random(n_blocks)
    spec = ado.run_spec(h0=h0, trial=c, n_walkers=nw, dt=dt, n_blocks=n_blocks,
                        n_prop_steps=n_prop, fields=fields, zetas=zetas)
    return h0, h1, chol, c, spec


def numpy_run(h0, h1, chol, spec):
    """The same run on the numpy oracle (log-overlaps): (J, e_blocks, w_blocks, comb indices)."""
    c = spec.trial
    sd = orc.make_system_data(h0, h1, chol, (c, c), "restricted", spec.dt)
    prm = orc.params(dt=spec.dt, n_exp_terms=spec.n_exp_terms, n_prop_steps=spec.n_prop_steps,
                     weight_floor=spec.weight_floor, weight_cap=spec.weight_cap,
                     pop_control_damping=spec.damping, shift_ema=spec.shift_ema)
    st = orc.init_state(sd, spec.n_walkers, overlap_mode="log")
    eb, wb, idx = [], [], []
    t = spec.n_prop_steps
    for b in range(spec.n_blocks):
        st, obs = orc.afqmc_block(st, spec.fields[b * t:(b + 1) * t], float(spec.zetas[b]), sd, prm,
                                  overlap_mode="log")
        eb.append(obs["energy"])
        wb.append(obs["weight"])
        idx.append(obs["sr_indices"])
    eb, wb = np.array(eb), np.array(wb)
    return float(np.sum(wb * eb) / np.sum(wb)), eb, wb, idx


def test_restatement_matches_numpy_oracle():
    h0, h1, chol, c, spec = _tiny()
    j, eb, wb, idx = ado.run_energy(torch.tensor(h1), torch.tensor(chol), spec, return_blocks=True)
    j_np, eb_np, wb_np, idx_np = numpy_run(h0, h1, chol, spec)
    np.testing.assert_allclose(eb, eb_np, rtol=1e-11)
    np.testing.assert_allclose(wb, wb_np, rtol=1e-11)
    assert abs(float(j) - j_np) <= 1e-11 * abs(j_np)
    for a, b in zip(idx, idx_np):
        np.testing.assert_array_equal(a, b)
    # the run must exercise the weights (no walker cut, comb not trivial)
    assert np.all(wb > 0) and len(set(idx_np[0].tolist())) > 1


def _fd(fun, eps):
    return (fun(eps) - fun(-eps)) / (2.0 * eps)


def test_autograd_h1_matches_central_differences():
    h0, h1, chol, c, spec = _tiny()
    rng = np.random.default_rng(5)
    m = rng.standard_normal(h1.shape)
    v = 0.5 * (m + m.T)
    j, gh, _ = ado.gradients(h0, h1, chol, spec)
    d_ad = float(np.sum(gh * v))
    idx0 = numpy_run(h0, h1, chol, spec)[3]

    def jfun(eps):
        out = numpy_run(h0, h1 + eps * v, chol, spec)
        for a, b in zip(out[3], idx0):   # same comb picks: the FD sees no jump
            np.testing.assert_array_equal(a, b)
        return out[0]

    d_fd = _fd(jfun, 1e-6)
    assert abs(d_ad) > 1e-3
    assert abs(d_ad - d_fd) <= 1e-6 * abs(d_fd), (d_ad, d_fd)


def test_autograd_chol_matches_central_differences():
    h0, h1, chol, c, spec = _tiny()
    rng = np.random.default_rng(6)
    a = rng.standard_normal(chol.shape)
    d = 0.5 * (a + a.transpose(0, 2, 1)) * 0.05
    j, _, gl = ado.gradients(h0, h1, chol, spec)
    d_ad = float(np.sum(gl * d))
    idx0 = numpy_run(h0, h1, chol, spec)[3]

    def jfun(eps):
        out = numpy_run(h0, h1, chol + eps * d, spec)
        for x, y in zip(out[3], idx0):
            np.testing.assert_array_equal(x, y)
        return out[0]

    d_fd = _fd(jfun, 1e-5)
    assert abs(d_ad) > 1e-4
    assert abs(d_ad - d_fd) <= 1e-6 * abs(d_fd), (d_ad, d_fd)


@pytest.mark.parametrize("n,o", [(6, 2), (7, 3)])
def test_torch_qr_gauge_gradient(n, o):
    """The phase-fixed QR of the restatement is differentiated by torch's QR
    backward; check it against central differences of the numpy QR on a random
    complex matrix (the gauge diag(R) real >= 0 fixes Q uniquely)."""
    rng = np.random.default_rng(n)
    a = rng.standard_normal((1, n, o)) + 1j * rng.standard_normal((1, n, o))
    wr = rng.standard_normal((1, n, o)) + 1j * rng.standard_normal((1, n, o))
    da = rng.standard_normal((1, n, o)) + 1j * rng.standard_normal((1, n, o))

    def f_np(x):
        q = orc.qr_phase_fixed(x)[0]
        return float(np.sum((np.conj(wr) * q).real))

    at = torch.tensor(a, requires_grad=True)
    q = ado.qr_phase_fixed(at)
    val = torch.sum((torch.conj(torch.tensor(wr)) * q).real)
    (g,) = torch.autograd.grad(val, at)
    d_ad = float(np.sum((np.conj(g.numpy()) * da).real))
    d_fd = (f_np(a + 1e-6 * da) - f_np(a - 1e-6 * da)) / 2e-6
    assert abs(d_ad - d_fd) <= 1e-7 * max(1.0, abs(d_fd))


def test_adjoint_refuses_host_tensors():
    """The product adjoint runs on the HIP kernels only: host tensors are an error."""
    from ad_afqmc_prototype_amd.prop import adjoint as adj
    h0, h1, chol, c, spec = _tiny()
    run = adj.adjoint_run(trial=torch.tensor(c), h0=h0, dt=spec.dt, n_walkers=spec.n_walkers,
                          n_blocks=spec.n_blocks, n_prop_steps=spec.n_prop_steps,
                          fields=torch.tensor(spec.fields), zetas=torch.tensor(spec.zetas))
    with pytest.raises(ValueError):
        adj.run_energy(torch.tensor(h1), torch.tensor(chol), run)
