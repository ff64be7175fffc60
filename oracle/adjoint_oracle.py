"""Differentiable CPU restatement (torch, complex128) of one phaseless-AFQMC run
with restricted walkers and an RHF trial, for checking the HIP adjoint.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the package.

The function differentiated is the run's energy estimate

    J(h1, L) = sum_b W_b e_b / sum_b W_b

over its blocks: the ratio estimator of the reference's final analysis
(stat_utils.py:41-50, driver.py:165-170), with W_b the block weight and e_b the
block energy (prop/blocks.py:59-79).  The run holds everything the parameters
reach:

* set-up: mean-field shifts, h0_prop, h1_eff and exp(-dt/2 h1_eff)
  (prop/chol_afqmc_ops.py:59-103), half-rotated integrals (meas/rhf.py:141-150);
* the initial energy estimate (prop/afqmc.py:17-71), which also starts the
  population shift;
* n_blocks blocks of afqmc_block (prop/blocks.py:22-95): afqmc_step
  (prop/afqmc.py:74-147) n_prop_steps times, QR, clipped local energies, block
  energy and shift EMA, comb reconfiguration (walkers.py:233-265).

Fields and comb uniforms are inputs (common random numbers).  Discrete
decisions -- comb indices, the floor / cap cuts of the importance function, the
energy clip -- are piecewise constant in the parameters: autograd here
differentiates the branch taken, which is what jax.grad of the reference does.
Overlaps are carried as log-overlaps (log det^2), as the HIP path does.

Pinning: jax is absent, so jax.grad of the reference cannot be run.  This
restatement stands in for it and is pinned by tests/test_adjoint.py: its J
equals the numpy oracle's run (oracle/afqmc_oracle.py) on the same inputs, and
its autograd gradient equals central differences of the numpy oracle's J.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

c128 = torch.complex128
f64 = torch.float64


@dataclass(frozen=True)
class run_spec:
    h0: float
    trial: np.ndarray        # (n, o) real RHF orbitals
    n_walkers: int
    dt: float
    n_blocks: int
    n_prop_steps: int
    fields: np.ndarray       # (n_blocks * n_prop_steps, nw, naux)
    zetas: np.ndarray        # (n_blocks,)
    n_exp_terms: int = 6
    weight_floor: float = 1.0e-3
    weight_cap: float = 100.0
    damping: float = 0.1
    shift_ema: float = 0.1


def setup(h0: float, h1: torch.Tensor, chol: torch.Tensor, c: torch.Tensor, dt: float):
    """prop/chol_afqmc_ops.py:59-103 and meas/rhf.py:141-150 for an RHF trial:
    mf_g = i tr(L_g (P + P)), P = C C^T; h0_prop = -h0 - sum mf^2 / 2;
    h1_eff = h1 - sum L_g L_g / 2 - sum_g Re(i mf_g) L_g; E = expm(-dt/2 h1_eff);
    rot_g = C^T L_g; rot_h1 = C^T h1."""
    p = c @ c.T
    mf = 1.0j * torch.einsum("gij,ji->g", chol, 2.0 * p).to(c128)
    h0_prop = -h0 - 0.5 * torch.sum(mf * mf)
    h1_eff = (h1 - 0.5 * torch.einsum("gik,gkj->ij", chol, chol)
              - torch.einsum("g,gij->ij", (1.0j * mf).real, chol))
    e_half = torch.linalg.matrix_exp(-0.5 * dt * h1_eff)
    rot = torch.einsum("pi,gpj->gij", c, chol)
    rot_h1 = c.T @ h1
    return mf, h0_prop, e_half, rot, rot_h1


def log_overlap(w: torch.Tensor, c: torch.Tensor) -> torch.Tensor:
    """2 log det(C^T w) (meas/rhf.py:22-24 as a logarithm)."""
    sign, logabs = torch.linalg.slogdet(c.to(c128).T @ w)
    return 2.0 * (logabs + 1.0j * torch.angle(sign))


def greens(w: torch.Tensor, c: torch.Tensor) -> torch.Tensor:
    """G = (w (C^T w)^-1)^T, (nw, o, n) (meas/rhf.py:15-19)."""
    m = c.to(c128).T @ w
    return torch.linalg.solve(m.transpose(1, 2), w.transpose(1, 2))


def energy(g: torch.Tensor, h0: float, rot: torch.Tensor, rot_h1: torch.Tensor) -> torch.Tensor:
    """energy_kernel_r (meas/rhf.py:36-50): h0 + 2 sum G.rot_h1 + 2 sum c_g^2 - sum tr(f_g f_g)."""
    rc = rot.to(c128)
    f = torch.einsum("gap,wbp->wgab", rc, g)
    cg = torch.einsum("wgaa->wg", f)
    exch = torch.einsum("wgab,wgba->w", f, f)
    e1 = 2.0 * torch.einsum("wap,ap->w", g, rot_h1.to(c128))
    return h0 + e1 + 2.0 * torch.sum(cg * cg, dim=1) - exch


def qr_phase_fixed(w: torch.Tensor) -> torch.Tensor:
    """_qr (walkers.py:169-180): Q of the reduced QR with diag(R) real >= 0."""
    q, r = torch.linalg.qr(w, mode="reduced")
    d = torch.diagonal(r, dim1=1, dim2=2).detach()
    a = torch.abs(d)
    phase = d / torch.where(a == 0, torch.ones_like(a), a)
    return q * torch.conj(phase)[:, None, :]


def natorbs(c: np.ndarray, nw: int) -> torch.Tensor:
    """walkers.py:13-16 + init_walkers: top-o eigenvectors of the total rdm1."""
    dm = 2.0 * c @ c.T
    vecs = np.linalg.eigh(0.5 * (dm + dm.T))[1][:, ::-1][:, : c.shape[1]]
    w = torch.as_tensor(np.ascontiguousarray(vecs), dtype=c128)
    return w.unsqueeze(0).expand(nw, *w.shape).contiguous()


def run_energy(h1: torch.Tensor, chol: torch.Tensor, spec: run_spec, *, return_blocks: bool = False):
    """J(h1, chol) of the run described by ``spec`` (differentiable in h1, chol)."""
    c = torch.as_tensor(spec.trial, dtype=f64)
    dt, sd = spec.dt, spec.dt ** 0.5
    mf, h0_prop, e_half, rot, rot_h1 = setup(spec.h0, h1, chol, c, dt)
    ec = e_half.to(c128)
    lc = chol.to(c128)
    nw = spec.n_walkers
    a = 1.0j * sd

    w = natorbs(spec.trial, nw)
    e_est = torch.mean(energy(greens(w, c), spec.h0, rot, rot_h1).real)
    weights = torch.ones(nw, dtype=f64)
    lo = log_overlap(w, c)
    pop = e_est
    thresh = (2.0 / dt) ** 0.5
    e_blocks, w_blocks, idx_all = [], [], []
    t = 0
    for b in range(spec.n_blocks):
        for _ in range(spec.n_prop_steps):
            x = torch.as_tensor(spec.fields[t], dtype=f64)
            t += 1
            g = greens(w, c)
            fb = 2.0 * torch.einsum("gap,wap->wg", rot.to(c128), g)
            fs = -sd * (1.0j * fb - mf)
            xbar = x - fs
            shift = torch.sum(xbar * mf, dim=1)
            fbt = torch.sum(x * fs - 0.5 * fs * fs, dim=1)
            vhs = torch.einsum("wg,gpq->wpq", xbar, lc)
            y = ec @ w
            term, acc = y, y
            for k in range(1, spec.n_exp_terms + 1):
                term = (a / k) * (vhs @ term)
                acc = acc + term
            w_new = ec @ acc
            lo_new = log_overlap(w_new, c)
            z = -sd * shift + fbt + dt * (pop + h0_prop) + lo_new - lo
            ph = -sd * shift + lo_new - lo
            imp = torch.exp(z.real) * torch.cos(ph.imag)
            imp = torch.where(imp < spec.weight_floor, torch.zeros_like(imp), imp)
            imp = torch.where(imp > spec.weight_cap, torch.zeros_like(imp), imp)
            wn = weights * imp
            weights = torch.where(wn > spec.weight_cap, torch.zeros_like(wn), wn)
            avg = torch.clamp(torch.mean(weights), min=1.0e-300)
            pop = e_est - spec.damping * torch.log(avg) / dt
            w, lo = w_new, lo_new
        w = qr_phase_fixed(w)
        lo = log_overlap(w, c)
        e = energy(greens(w, c), spec.h0, rot, rot_h1).real
        e = torch.where(torch.abs(e - e_est) > thresh, e_est.expand_as(e), e)
        wsum = torch.sum(weights)
        e_b = torch.sum(weights * e) / wsum if float(wsum.detach()) != 0.0 else e_est
        pop = (1.0 - spec.shift_ema) * pop + spec.shift_ema * e_b
        cw = torch.cumsum(torch.abs(weights), dim=0)
        tot = cw[-1]
        z_comb = tot.detach() * (torch.arange(nw, dtype=f64) + float(spec.zetas[b])) / nw
        idx = torch.searchsorted(cw.detach(), z_comb, right=False).clamp_(max=nw - 1)
        w, lo = w[idx], lo[idx]
        weights = (tot / nw).expand(nw).clone()
        e_blocks.append(e_b)
        w_blocks.append(wsum)
        idx_all.append(idx.numpy().astype(np.int32))
    eb = torch.stack(e_blocks)
    wb = torch.stack(w_blocks)
    j = torch.sum(wb * eb) / torch.sum(wb)
    if return_blocks:
        return j, eb.detach().numpy(), wb.detach().numpy(), idx_all
    return j


def gradients(h0: float, h1: np.ndarray, chol: np.ndarray, spec: run_spec):
    """(J, dJ/dh1, dJ/dL) by torch.autograd through run_energy."""
    h1t = torch.tensor(h1, dtype=f64, requires_grad=True)
    lt = torch.tensor(chol, dtype=f64, requires_grad=True)
    j = run_energy(h1t, lt, spec)
    gh, gl = torch.autograd.grad(j, (h1t, lt))
    return float(j.detach()), gh.numpy(), gl.numpy()
