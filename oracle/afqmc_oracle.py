"""CPU FP64 restatement of the reference phaseless-AFQMC hot path.

TEST INFRASTRUCTURE ONLY.  This module is the oracle for the HIP path: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product package never imports it and never
falls back to it.

Everything is numpy complex128 / float64, batched over walkers with a leading
axis.  Walker kinds: "restricted" -> array (nw, n, o); "unrestricted" ->
tuple (wu, wd), each (nw, n, o).  Trial: a pair (c_up, c_dn) of (n, o) arrays;
an RHF trial passes the same array twice.

Each function cites the reference file:line it follows (paths relative to the
reference package ``ad_afqmc_prototype/``).  Deviations, all explicit keyword
switches that default to the reference behaviour:

* ``overlap_mode="log"`` carries log-overlaps (log|det| + i*phase) instead of the
  raw det^2 of meas/rhf.py:22-24; identical ratios wherever det^2 is finite.
* fields and comb uniforms are arguments (the reference draws them with
  jax.random inside prop/afqmc.py:86-88 and prop/blocks.py:81-82).
* the cumulative sum of the comb is a sequential left-to-right FP64 sum
  (numpy.cumsum order), which fixes the reconfiguration indices bit-exactly.

Parity of every function here is pinned in tests/test_oracle.py by closed forms,
invariants, finite differences of the overlap and brute force on tiny systems.
"""
from __future__ import annotations

from dataclasses import dataclass, replace
from typing import Any, Optional

import numpy as np
import scipy.linalg


# --------------------------------------------------------------------------
# trial / overlap / Green's function
# --------------------------------------------------------------------------

def is_unrestricted(w) -> bool:
    """walkers.py:80-81"""
    return isinstance(w, tuple) and len(w) == 2


def n_walkers(w) -> int:
    """walkers.py:84-85"""
    return w[0].shape[0] if is_unrestricted(w) else w.shape[0]


def overlap_matrix(c: np.ndarray, w: np.ndarray) -> np.ndarray:
    """m = C^H w, batched (trial/rhf.py:36, meas/rhf.py:23)."""
    return np.einsum("pi,wpj->wij", c.conj(), w)


def det(m: np.ndarray) -> np.ndarray:
    """trial/rhf.py:25-26 (LU determinant)."""
    return np.linalg.det(m)


def logdet(m: np.ndarray) -> np.ndarray:
    """complex log of det(m): log|det| + i*arg(det)."""
    sign, logabs = np.linalg.slogdet(m)
    return logabs + 1j * np.angle(sign)


def overlap(w, trial, kind: str, mode: str = "raw") -> np.ndarray:
    """overlap_r (trial/rhf.py:35-37, meas/rhf.py:22-24): det(C^H w)^2;
    overlap_u (trial/rhf.py:40-44, meas/rhf.py:53-57): det(Cu^H wu) det(Cd^H wd).
    mode="log" returns the complex log of the same quantity."""
    cu, cd = trial
    if kind == "restricted":
        m = overlap_matrix(cu, w)
        return 2.0 * logdet(m) if mode == "log" else det(m) ** 2
    if kind == "unrestricted":
        wu, wd = w
        mu = overlap_matrix(cu, wu)
        md = overlap_matrix(cd, wd)
        if mode == "log":
            return logdet(mu) + logdet(md)
        return det(mu) * det(md)
    if kind == "generalized":
        m = overlap_matrix_g(cu, w)
        return logdet(m) if mode == "log" else det(m)
    raise ValueError(f"unknown walker_kind: {kind}")


def overlap_matrix_g(c: np.ndarray, w: np.ndarray) -> np.ndarray:
    """overlap_g (trial/rhf.py:47-53, meas/rhf.py:108-114): generalized walkers
    (nw, 2n, 2o) against an RHF trial: m = [C^H w_top; C^H w_bot] (2o x 2o)."""
    n = c.shape[0]
    return np.concatenate([overlap_matrix(c, w[:, :n, :]), overlap_matrix(c, w[:, n:, :])], axis=1)


def half_green(w: np.ndarray, c: np.ndarray, kind: str = "restricted") -> np.ndarray:
    """_half_green_from_overlap_matrix (meas/rhf.py:15-19): solve(m^T, w^T) -> (nw, o, n);
    generalized walkers: (nw, 2o, 2n) from the stacked overlap matrix."""
    m = overlap_matrix_g(c, w) if kind == "generalized" else overlap_matrix(c, w)
    return np.linalg.solve(np.swapaxes(m, 1, 2), np.swapaxes(w, 1, 2))


# --------------------------------------------------------------------------
# measurement context and kernels
# --------------------------------------------------------------------------

@dataclass(frozen=True)
class meas_ctx:
    """rhf_meas_ctx (meas/rhf.py:133-138), one entry per spin sector."""
    rot_h1: tuple          # per spin (o, n)
    rot_chol: tuple        # per spin (G, o, n)
    rot_chol_flat: tuple   # per spin (G, o*n)


def build_meas_ctx(h1: np.ndarray, chol: np.ndarray, trial) -> meas_ctx:
    """build_meas_ctx (meas/rhf.py:141-150), for each spin's trial orbitals."""
    out_h1, out_rc, out_rf = [], [], []
    for c in trial:
        ch = c.conj().T
        rot_h1 = ch @ h1
        rot_chol = np.einsum("pi,gij->gpj", ch, chol, optimize=True)
        out_h1.append(rot_h1)
        out_rc.append(rot_chol)
        out_rf.append(rot_chol.reshape(rot_chol.shape[0], -1))
    return meas_ctx(tuple(out_h1), tuple(out_rc), tuple(out_rf))


def force_bias(w, trial, mctx: meas_ctx, kind: str) -> np.ndarray:
    """force_bias_kernel_r (meas/rhf.py:27-33): 2 * rot_chol_flat @ vec(G);
    force_bias_kernel_u (meas/rhf.py:60-72): rot_chol_flat @ vec(Gu + Gd),
    generalised to per-spin trial orbitals (UHF): Lu @ vec(Gu) + Ld @ vec(Gd)."""
    cu, cd = trial
    if kind == "restricted":
        g = half_green(w, cu)
        return 2.0 * (g.reshape(g.shape[0], -1) @ mctx.rot_chol_flat[0].T)
    if kind == "generalized":
        # force_bias_kernel_g (meas/rhf.py:117-130): rot_chol_flat @ vec(G_upup + G_dndn)
        n, o = cu.shape
        g = half_green(w, cu, kind)
        gs = g[:, :o, :n] + g[:, o:, n:]
        return gs.reshape(gs.shape[0], -1) @ mctx.rot_chol_flat[0].T
    wu, wd = w
    gu = half_green(wu, cu)
    gd = half_green(wd, cd)
    nw = gu.shape[0]
    return (gu.reshape(nw, -1) @ mctx.rot_chol_flat[0].T
            + gd.reshape(nw, -1) @ mctx.rot_chol_flat[1].T)


def _coulomb_exchange(g: np.ndarray, rot_chol: np.ndarray):
    """f_g = rot_chol_g @ G^T (o x o); c_g = tr f_g; exc = sum_g tr(f_g f_g)
    (meas/rhf.py:45-47).  Returns c (nw, G) and exc (nw,)."""
    nw = g.shape[0]
    c_all = np.empty((nw, rot_chol.shape[0]), dtype=np.complex128)
    exc = np.empty(nw, dtype=np.complex128)
    for i in range(nw):
        f = np.einsum("gij,kj->gik", rot_chol, g[i], optimize=True)
        c_all[i] = np.einsum("gii->g", f)
        exc[i] = np.einsum("gij,gji->", f, f)
    return c_all, exc


def energy(w, h0: float, trial, mctx: meas_ctx, kind: str) -> np.ndarray:
    """energy_kernel_r (meas/rhf.py:36-50) and energy_kernel_u (meas/rhf.py:75-105),
    the latter with per-spin half-rotated integrals (UHF generalisation)."""
    cu, cd = trial
    if kind == "generalized":
        raise KeyError("generalized walkers have no energy kernel (meas/rhf.py:169-174)")
    if kind == "restricted":
        g = half_green(w, cu)
        e1 = 2.0 * np.einsum("wij,ij->w", g, mctx.rot_h1[0])
        c, exc = _coulomb_exchange(g, mctx.rot_chol[0])
        e2 = 2.0 * np.sum(c * c, axis=1) - exc
        return h0 + e1 + e2
    wu, wd = w
    gu = half_green(wu, cu)
    gd = half_green(wd, cd)
    e1 = np.einsum("wij,ij->w", gu, mctx.rot_h1[0]) + np.einsum("wij,ij->w", gd, mctx.rot_h1[1])
    cu_, exu = _coulomb_exchange(gu, mctx.rot_chol[0])
    cd_, exd = _coulomb_exchange(gd, mctx.rot_chol[1])
    e2 = (np.sum(cu_ * cu_, axis=1) + np.sum(cd_ * cd_, axis=1)
          + 2.0 * np.sum(cu_ * cd_, axis=1) - exu - exd) / 2.0
    return h0 + e1 + e2


# --------------------------------------------------------------------------
# propagator
# --------------------------------------------------------------------------

@dataclass(frozen=True)
class prop_ctx:
    """chol_afqmc_ctx (prop/chol_afqmc_ops.py:15-42)."""
    dt: float
    sqrt_dt: float
    exp_h1_half: np.ndarray   # (n, n)
    mf_shifts: np.ndarray     # (G,) complex (purely imaginary)
    h0_prop: complex


def mf_shifts(chol: np.ndarray, rdm1: np.ndarray, basis: str = "restricted") -> np.ndarray:
    """_mf_shifts (prop/chol_afqmc_ops.py:53-63): 1j * tr(L_g dm) with dm the
    total density (restricted basis) or the given spin-orbital one
    (generalized basis)."""
    if basis == "restricted":
        dm = rdm1[0] + rdm1[1] if rdm1.ndim == 3 and rdm1.shape[0] == 2 else rdm1
    else:
        dm = rdm1
    return 1.0j * np.einsum("gij,ji->g", chol, dm, optimize=True)


def get_rdm1(trial) -> np.ndarray:
    """get_rdm1 (trial/rhf.py:29-32), per spin for a UHF trial."""
    cu, cd = trial
    return np.stack([cu @ cu.conj().T, cd @ cd.conj().T], axis=0)


def build_prop_ctx(h0: float, h1: np.ndarray, chol: np.ndarray, rdm1: np.ndarray,
                   dt: float, basis: str = "restricted") -> prop_ctx:
    """_build_prop_ctx (prop/chol_afqmc_ops.py:77-103); a generalized basis
    skips the v0 / v1 corrections (:86-94)."""
    mf = mf_shifts(chol, rdm1, basis)
    h0_prop = -h0 - 0.5 * np.sum(mf ** 2)
    h1_eff = h1
    if basis == "restricted":
        v0m = 0.5 * np.einsum("gik,gkj->ij", chol, chol, optimize=True)
        mf_r = (1.0j * mf).real
        v1m = np.einsum("g,gik->ik", mf_r, chol, optimize=True)
        h1_eff = h1 - v0m - v1m
    exp_h1_half = scipy.linalg.expm(-0.5 * dt * h1_eff)
    return prop_ctx(dt=float(dt), sqrt_dt=float(np.sqrt(dt)), exp_h1_half=exp_h1_half,
                    mf_shifts=mf, h0_prop=complex(h0_prop))


def taylor_expm_action(a, mat: np.ndarray, vecs: np.ndarray, n_terms: int) -> np.ndarray:
    """taylor_expm_action (prop/utils.py:20-31), batched over a leading axis."""
    out = vecs
    term = vecs
    for k in range(1, n_terms + 1):
        term = (a / k) * (mat @ term)
        out = out + term
    return out


def make_vhs(chol: np.ndarray, x: np.ndarray) -> np.ndarray:
    """_make_vhs_split_flat (prop/chol_afqmc_ops.py:70-74), batched: (nw, n, n)."""
    g, n = chol.shape[0], chol.shape[1]
    flat = chol.reshape(g, n * n)
    # contiguous copies of the real / imaginary planes: a strided operand
    # sends numpy's matmul down its slow non-BLAS loop
    v = (np.ascontiguousarray(x.real) @ flat) + 1j * (np.ascontiguousarray(x.imag) @ flat)
    return v.reshape(x.shape[0], n, n)


def apply_trotter(w, fields: np.ndarray, pctx: prop_ctx, chol: np.ndarray,
                  n_terms: int, kind: str):
    """_apply_trotter_r / _apply_trotter_u (prop/chol_afqmc_ops.py:173-196)."""
    vhs = make_vhs(chol, fields)
    a = 1.0j * pctx.sqrt_dt
    e = pctx.exp_h1_half
    if kind == "restricted":
        w1 = e @ w
        w2 = taylor_expm_action(a, vhs, w1, n_terms)
        return e @ w2
    if kind == "generalized":
        # _apply_trotter_g_from_restricted (prop/chol_afqmc_ops.py:118-124,157-170,200-212):
        # both spin blocks of the spin-orbital walker with the same E and VHS
        n = e.shape[0]
        halves = [e @ taylor_expm_action(a, vhs, e @ w[:, s * n:(s + 1) * n, :], n_terms) for s in (0, 1)]
        return np.concatenate(halves, axis=1)
    wu, wd = w
    return (e @ taylor_expm_action(a, vhs, e @ wu, n_terms),
            e @ taylor_expm_action(a, vhs, e @ wd, n_terms))


# --------------------------------------------------------------------------
# step / block
# --------------------------------------------------------------------------

@dataclass(frozen=True)
class params:
    """afqmc_params (prop/types.py:19-32)."""
    dt: float = 0.005
    n_chunks: int = 1
    n_exp_terms: int = 6
    pop_control_damping: float = 0.1
    weight_floor: float = 1.0e-3
    weight_cap: float = 100.0
    n_prop_steps: int = 50
    shift_ema: float = 0.1


@dataclass(frozen=True)
class state:
    """prop_state (prop/types.py:10-16) minus the PRNG key (fields are injected)."""
    walkers: Any
    weights: np.ndarray
    overlaps: np.ndarray          # raw, or log when overlap_mode == "log"
    pop_control_ene_shift: float
    e_estimate: float


@dataclass(frozen=True)
class system_data:
    h0: float
    h1: np.ndarray
    chol: np.ndarray
    trial: tuple
    kind: str
    mctx: meas_ctx
    pctx: prop_ctx


def make_system_data(h0, h1, chol, trial, kind, dt) -> system_data:
    trial = tuple(np.asarray(c) for c in trial)
    mctx = build_meas_ctx(h1, chol, trial)
    pctx = build_prop_ctx(h0, h1, chol, get_rdm1(trial), dt)
    return system_data(h0, h1, chol, trial, kind, mctx, pctx)


def natorbs(dm: np.ndarray, n_occ: int) -> np.ndarray:
    """_natorbs (walkers.py:13-16)."""
    dm = 0.5 * (dm + dm.conj().T)
    vecs = np.linalg.eigh(dm)[1][:, ::-1]
    return vecs[:, :n_occ]


def init_walkers(rdm1: np.ndarray, nocc: int, nw: int, kind: str):
    """init_walkers (walkers.py:19-77), restricted / unrestricted / generalized
    (from the spin blocks: [[up, 0], [0, dn]], walkers.py:27-49)."""
    if kind == "generalized":
        n = rdm1.shape[1]
        top = np.concatenate([natorbs(rdm1[0], nocc), np.zeros((n, nocc))], axis=1)
        bot = np.concatenate([np.zeros((n, nocc)), natorbs(rdm1[1], nocc)], axis=1)
        w0 = np.concatenate([top, bot], axis=0) + 0.0j
        return np.broadcast_to(w0, (nw,) + w0.shape).copy()
    if kind == "restricted":
        w0 = natorbs(rdm1[0] + rdm1[1], nocc) + 0.0j
        return np.broadcast_to(w0, (nw,) + w0.shape).copy()
    wu = natorbs(rdm1[0], nocc) + 0.0j
    wd = natorbs(rdm1[1], nocc) + 0.0j
    return (np.broadcast_to(wu, (nw,) + wu.shape).copy(),
            np.broadcast_to(wd, (nw,) + wd.shape).copy())


def init_state(sd: system_data, nw: int, overlap_mode: str = "raw", walkers=None,
               e_estimate: float | None = None) -> state:
    """init_prop_state (prop/afqmc.py:17-71) without the PRNG key; e_estimate
    given directly skips the energy (generalized walkers have none)."""
    nocc = sd.trial[0].shape[1]
    if walkers is None:
        walkers = init_walkers(get_rdm1(sd.trial), nocc, nw, sd.kind)
    ov = overlap(walkers, sd.trial, sd.kind, overlap_mode)
    if e_estimate is None:
        e_est = float(np.mean(np.real(energy(walkers, sd.h0, sd.trial, sd.mctx, sd.kind))))
    else:
        e_est = float(e_estimate)
    return state(walkers, np.ones(nw), ov, e_est, e_est)


def field_shift_terms(fields: np.ndarray, fb: np.ndarray, pctx: prop_ctx):
    """prop/afqmc.py:97-101."""
    mf = pctx.mf_shifts
    field_shifts = -pctx.sqrt_dt * (1.0j * fb - mf)
    shifted = fields - field_shifts
    shift_term = np.sum(shifted * mf, axis=1)
    fb_term = np.sum(fields * field_shifts - 0.5 * field_shifts * field_shifts, axis=1)
    return shifted, shift_term, fb_term


def importance(ov_new, ov_old, shift_term, fb_term, pop_shift, pctx: prop_ctx,
               overlap_mode: str = "raw"):
    """prop/afqmc.py:116-124: the complex importance function
    imp_fun = exp(exponent) * ratio with
    exponent = -sqrt_dt shift_term + fb_term + dt (pop_shift + h0_prop),
    ratio = ov_new / ov_old, and the phase theta = angle(exp(-sqrt_dt shift_term) * ratio).

    Returns (log_imp, mag, theta): log_imp the complex log of imp_fun (its
    imaginary part is defined modulo 2 pi), mag = |imp_fun|.  overlap_mode="raw"
    is the reference formula verbatim; "log" takes ov_* as log-overlaps and
    forms the same numbers as exp(Re(exponent + dlog)) and
    Im(-sqrt_dt shift_term + dlog), without the overflow of the raw product."""
    with np.errstate(all="ignore"):
        exponent = (-pctx.sqrt_dt * shift_term + fb_term
                    + pctx.dt * (pop_shift + pctx.h0_prop))
        if overlap_mode == "log":
            dlog = ov_new - ov_old
            log_imp = exponent + dlog
            mag = np.exp(np.real(log_imp))
            theta = np.imag(-pctx.sqrt_dt * shift_term + dlog)
        else:
            ratio = ov_new / ov_old
            imp = np.exp(exponent) * ratio
            log_imp = np.log(imp)
            mag = np.abs(imp)
            theta = np.angle(np.exp(-pctx.sqrt_dt * shift_term) * ratio)
    return log_imp, mag, theta


def phaseless_update(ov_new, ov_old, shift_term, fb_term, weights, pop_shift, e_estimate,
                     pctx: prop_ctx, prm: params, overlap_mode: str = "raw"):
    """prop/afqmc.py:116-138.  Returns (weights_new, pop_shift_new).

    The importance magnitude and phase come from importance(); then the
    phaseless projection |imp| cos(theta), the floor / cap cuts (NaN passes
    both, as NaN comparisons are false), the weight cap and the population
    shift from the clipped mean weight."""
    _, mag, theta = importance(ov_new, ov_old, shift_term, fb_term, pop_shift, pctx,
                               overlap_mode)
    with np.errstate(all="ignore"):
        imp_ph = mag * np.cos(theta)
        imp_ph = np.where(imp_ph < prm.weight_floor, 0.0, imp_ph)
        imp_ph = np.where(imp_ph > prm.weight_cap, 0.0, imp_ph)
        w_new = weights * imp_ph
        w_new = np.where(w_new > prm.weight_cap, 0.0, w_new)
        avg_w = np.maximum(np.mean(w_new), 1.0e-300)
        pop_new = e_estimate - prm.pop_control_damping * (np.log(avg_w) / pctx.dt)
    return w_new, float(np.real(pop_new))


def afqmc_step(st: state, fields: np.ndarray, sd: system_data, prm: params,
               overlap_mode: str = "raw", return_intermediates: bool = False):
    """afqmc_step (prop/afqmc.py:74-147) with injected fields."""
    fb = force_bias(st.walkers, sd.trial, sd.mctx, sd.kind)
    shifted, shift_term, fb_term = field_shift_terms(fields, fb, sd.pctx)
    walkers_new = apply_trotter(st.walkers, shifted, sd.pctx, sd.chol, prm.n_exp_terms, sd.kind)
    ov_new = overlap(walkers_new, sd.trial, sd.kind, overlap_mode)
    w_new, pop_new = phaseless_update(ov_new, st.overlaps, shift_term, fb_term, st.weights,
                                      st.pop_control_ene_shift, st.e_estimate, sd.pctx, prm,
                                      overlap_mode)
    out = state(walkers_new, w_new, ov_new, pop_new, st.e_estimate)
    if return_intermediates:
        return out, dict(force_bias=fb, shifted_fields=shifted, shift_term=shift_term,
                         fb_term=fb_term)
    return out


def qr_phase_fixed(mat: np.ndarray):
    """_qr (walkers.py:169-180), batched: reduced QR with diag(R) real >= 0."""
    q, r = np.linalg.qr(mat, mode="reduced")
    d = np.diagonal(r, axis1=1, axis2=2)
    abs_d = np.abs(d)
    phase = d / np.where(abs_d == 0, 1.0 + 0.0j, abs_d)
    q = q * np.conj(phase)[:, None, :]
    r = phase[:, :, None] * r
    det_r = np.prod(np.diagonal(r, axis1=1, axis2=2), axis=1)
    return q, det_r


def orthonormalize(w, kind: str):
    """orthonormalize / orthogonalize (walkers.py:183-211); generalized walkers
    are one (2n x 2o) matrix, as restricted ones."""
    if kind == "unrestricted":
        return (qr_phase_fixed(w[0])[0], qr_phase_fixed(w[1])[0])
    return qr_phase_fixed(w)[0]


def sr_indices(weights: np.ndarray, zeta: float, nw: int) -> np.ndarray:
    """_sr_indices (walkers.py:233-240): sequential cumsum, comb, searchsorted left.

    searchsorted orders NaN after every number (numpy and jax.numpy alike), so
    with a NaN weight every tooth (z = NaN) lands on the first NaN of the
    cumulative sum.  An index past the end (a tooth above the last cumulative
    sum by rounding) is clamped to nw - 1, as jax's gather clamps the
    out-of-range index of w[idx] (walkers.py:258-263)."""
    cw = np.cumsum(np.abs(weights))
    tot = cw[-1]
    z = tot * (np.arange(nw) + zeta) / nw
    idx = np.searchsorted(cw, z, side="left")
    return np.minimum(idx, nw - 1).astype(np.int32)


def stochastic_reconfiguration(w, weights: np.ndarray, zeta: float, kind: str):
    """stochastic_reconfiguration (walkers.py:243-265)."""
    nw = weights.shape[0]
    cw = np.cumsum(np.abs(weights))
    avg = cw[-1] / nw
    weights_new = np.full((nw,), avg, dtype=weights.dtype)
    idx = sr_indices(weights, zeta, nw)
    if kind == "unrestricted":
        return (w[0][idx], w[1][idx]), weights_new, idx
    return w[idx], weights_new, idx


def block_energy(e_samples: np.ndarray, weights: np.ndarray, e_ref: float, pop_shift: float,
                 dt: float, shift_ema: float):
    """prop/blocks.py:59-79: a sample farther than sqrt(2/dt) from e_ref is
    replaced by e_ref (NaN samples stay: the comparison is false); the block
    energy is the weighted mean, or e_ref when the weights sum to zero; the
    population shift moves by the EMA (1 - a) pop + a e_block.
    Returns (e_block, w_sum, e_clipped, pop_new)."""
    e = np.real(np.asarray(e_samples))
    thresh = np.sqrt(2.0 / dt)
    with np.errstate(invalid="ignore"):
        e = np.where(np.abs(e - e_ref) > thresh, e_ref, e)
    w_sum = np.sum(weights)
    w_sum_safe = 1.0 if w_sum == 0 else w_sum
    e_block = np.sum(weights * e) / w_sum_safe
    if w_sum == 0:
        e_block = e_ref
    pop = (1.0 - shift_ema) * pop_shift + shift_ema * e_block
    return float(e_block), float(w_sum), e, float(np.real(pop))


def measure_block(st: state, sd: system_data, prm: params, overlap_mode: str = "raw"):
    """prop/blocks.py:53-79: QR, overlaps, clipped energies, block energy, pop EMA."""
    walkers = orthonormalize(st.walkers, sd.kind)
    ov = overlap(walkers, sd.trial, sd.kind, overlap_mode)
    e = np.real(energy(walkers, sd.h0, sd.trial, sd.mctx, sd.kind))
    e_block, w_sum, e, pop = block_energy(e, st.weights, st.e_estimate,
                                          st.pop_control_ene_shift, prm.dt, prm.shift_ema)
    st = replace(st, walkers=walkers, overlaps=ov, pop_control_ene_shift=pop)
    return st, e_block, w_sum, e


def afqmc_block(st: state, fields_seq, zeta: float, sd: system_data, prm: params,
                overlap_mode: str = "raw"):
    """afqmc_block (prop/blocks.py:22-95) with injected fields and comb uniform."""
    for t in range(prm.n_prop_steps):
        st = afqmc_step(st, fields_seq[t], sd, prm, overlap_mode)
    st, e_block, w_sum, _ = measure_block(st, sd, prm, overlap_mode)
    w_sr, weights_sr, idx = stochastic_reconfiguration(st.walkers, st.weights, zeta, sd.kind)
    ov = overlap(w_sr, sd.trial, sd.kind, overlap_mode)
    st = replace(st, walkers=w_sr, weights=weights_sr, overlaps=ov)
    return st, {"energy": e_block, "weight": w_sum, "sr_indices": idx}


def run_cadenced(st: state, sd: system_data, prm: params, *, n_steps: int, qr_every: int,
                 sr_every: int, fields_fn, zeta_fn, overlap_mode: str = "raw"):
    """Independent cadences built from the same primitives as afqmc_block: a QR
    (plus overlap refresh, walkers.py:206-211 / blocks.py:53-57) every
    ``qr_every`` steps; every ``sr_every`` steps a measurement (blocks.py:59-79)
    followed by comb reconfiguration (blocks.py:81-92).  With
    qr_every == sr_every == n_prop_steps this is afqmc_block exactly."""
    blocks = []
    ev = 0
    for t in range(1, n_steps + 1):
        st = afqmc_step(st, fields_fn(t - 1), sd, prm, overlap_mode)
        if t % sr_every == 0:
            st, e_block, w_sum, _ = measure_block(st, sd, prm, overlap_mode)
            w_sr, weights_sr, idx = stochastic_reconfiguration(st.walkers, st.weights,
                                                               zeta_fn(ev), sd.kind)
            ov = overlap(w_sr, sd.trial, sd.kind, overlap_mode)
            st = replace(st, walkers=w_sr, weights=weights_sr, overlaps=ov)
            blocks.append({"energy": e_block, "weight": w_sum, "sr_indices": idx})
            ev += 1
        elif t % qr_every == 0:
            w = orthonormalize(st.walkers, sd.kind)
            st = replace(st, walkers=w, overlaps=overlap(w, sd.trial, sd.kind, overlap_mode))
    return st, blocks


# --------------------------------------------------------------------------
# overlap-derivative measurements (meas/auto.py) -- independent checks
# --------------------------------------------------------------------------

def _ov_single(w, trial, kind):
    if kind in ("restricted", "generalized"):
        return overlap(w[None], trial, kind)[0]
    return overlap((w[0][None], w[1][None]), trial, kind)[0]


def _apply_single(mat, w, kind):
    if kind == "restricted":
        return mat @ w
    if kind == "generalized":  # a spatial one-body operator acts on both spin blocks
        n = mat.shape[0]
        return np.concatenate([mat @ w[:n], mat @ w[n:]], axis=0)
    return (mat @ w[0], mat @ w[1])


def _axpy(w, a, v, kind):
    if kind in ("restricted", "generalized"):
        return w + a * v
    return (w[0] + a * v[0], w[1] + a * v[1])


def force_bias_fd(w, trial, chol, kind, eps: float = 1.0e-5) -> np.ndarray:
    """Force bias as d/dx_g log <T|(1 + x L_g)|w> at x=0 (the principle of
    meas/auto.py:60-83), by 4th-order central finite differences."""
    out = np.empty(chol.shape[0], dtype=np.complex128)
    for g in range(chol.shape[0]):
        lw = _apply_single(chol[g], w, kind)
        f = lambda x: np.log(_ov_single(_axpy(w, x, lw, kind), trial, kind))
        out[g] = (-f(2 * eps) + 8 * f(eps) - 8 * f(-eps) + f(-2 * eps)) / (12 * eps)
    return out


def energy_fd(w, h0, h1, chol, trial, kind, eps: float = 1.0e-3) -> complex:
    """E = h0 + [d/dx <T|(1+x h1')|w> + 1/2 sum_g d2/dx2 <T|(1+x L_g+x^2 L_g^2/2)|w>]/<T|w>
    with h1' = h1 - 1/2 sum_g L_g L_g^T (meas/auto.py:27-36, 132-176), by
    4th-order central differences (the polynomial in x has degree <= 4*nocc,
    so a small eps plus a 4th-order stencil is accurate to ~1e-9 relative)."""
    v0 = 0.5 * np.einsum("gik,gjk->ij", chol, chol)
    h1e = h1 - v0
    ov0 = _ov_single(w, trial, kind)
    hw = _apply_single(h1e, w, kind)
    f1 = lambda x: _ov_single(_axpy(w, x, hw, kind), trial, kind)
    d1 = (-f1(2 * eps) + 8 * f1(eps) - 8 * f1(-eps) + f1(-2 * eps)) / (12 * eps)
    d2 = 0.0
    for g in range(chol.shape[0]):
        lw = _apply_single(chol[g], w, kind)
        llw = _apply_single(chol[g], lw, kind)

        def f2(x):
            wx = _axpy(_axpy(w, x, lw, kind), 0.5 * x * x, llw, kind)
            return _ov_single(wx, trial, kind)

        d2 += (-f2(2 * eps) + 16 * f2(eps) - 30 * ov0 + 16 * f2(-eps) - f2(-2 * eps)) / (12 * eps * eps)
    return h0 + (d1 + 0.5 * d2) / ov0
