"""Hand-written adjoint of the phaseless-AFQMC run (propagate + measure) behind a
torch.autograd.Function.

The differentiated quantity is the run's energy estimate

    J = sum_b W_b e_b / sum_b W_b

over its blocks -- the ratio estimator of the reference's final analysis
(stat_utils.py:41-50, driver.py:165-170) -- as a function of the one-body
Hamiltonian h1 and the Cholesky vectors L, for fixed auxiliary fields and comb
uniforms (common random numbers).  d<E>/d lambda for h1 + lambda V is
sum(dJ/dh1 * V).

Forward: the run's blocks (prop/blocks.py:22-95) on the library's HIP kernels
(force bias, field shifts, VHS, Trotter, Green's functions / log-overlaps,
phaseless update, QR, local energy, block energy, comb), restricted walkers and
an RHF trial, the initial energy estimate included (prop/afqmc.py:17-71).  The
state at the start of every block is kept as a checkpoint, and the last block's
intermediates (the segment the backward pass starts with); nothing else.

Backward: block by block in reverse, the block is recomputed from its checkpoint
(one block = n_prop_steps steps is the checkpoint interval; the last block is
not recomputed) and its adjoint applied step by step in reverse.  Discrete decisions (comb indices, the floor /
cap cuts of the importance function, the energy clip) are held at the values the
forward run took, i.e. the derivative of the branch taken.  Complex adjoints
follow z_bar = dJ/dRe z + i dJ/dIm z.  Per step, in reverse
(prop/afqmc.py:74-147):

* phaseless update: s = exp(Re Z) cos(Im Y) with Z = -sqrt_dt shift + fb_term +
  dt (pop + h0_prop) + lo_new - lo and Y = -sqrt_dt shift + lo_new - lo, through
  the population-shift mean;
* log-overlap / Green's function G = (w m^-1)^T, m = C^T w:
  w_bar = [G_bar^T - C (conj(G) G_bar^T - ld_bar I)] m^-H;
* Trotter w' = E T(E w), T = sum_k (a VHS)^k / k!, a = i sqrt_dt: the adjoint
  recursion t_bar_{k-1} = u_bar + conj(a/k) VHS^H t_bar_k, with
  VHS_bar = sum_k conj(a/k) t_bar_k t_{k-1}^H and E_bar = Re(w_bar' u^H + y0_bar w^H);
* VHS = sum_g xbar_g L_g: xbar_bar_g = <L_g, VHS_bar>, L_bar_g += Re(conj(xbar_g) VHS_bar);
* field shifts (prop/afqmc.py:97-101) and force bias fb = 2 rot . vec(G)
  (meas/rhf.py:27-33).

and per block end (prop/blocks.py:53-92): comb (scatter-add of the gathered
walkers, d|w|), block energy and shift EMA, energy clip, local energy
(meas/rhf.py:36-50: dE/dG = 2 rot_h1 + 4 sum_g c_g rot_g - 2 sum_g f_g rot_g,
the last sum as G Y with the fixed Y[(b,q),(a,p)] = sum_g R[g,a,q] R[g,b,p]),
and the QR with diag(R) real (walkers.py:169-211):
A_bar = [(I - Q Q^H) Q_bar + Q N] R^-H, K = Q^H Q_bar,
N = tril(K - K^H, -1) + i Im diag(K).

The set-up (mean-field shifts, h0_prop, h1_eff, exp(-dt/2 h1_eff), half-rotated
integrals; prop/chol_afqmc_ops.py:59-103, meas/rhf.py:141-150) is one-time
torch work outside the Function, so autograd carries the adjoints of its
outputs back to h1 and L.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from .. import kernels
from .. import walkers as wk
from ..core.system import system
from .chol_afqmc_ops import chol_sym_rows

c128 = torch.complex128
f64 = torch.float64


@dataclass(frozen=True)
class adjoint_run:
    """One run: RHF trial (n, o) float64, fields (n_blocks * n_prop_steps, nw,
    naux) float64 and comb uniforms (n_blocks,) float64, all on the device."""

    trial: torch.Tensor
    h0: float
    dt: float
    n_walkers: int
    n_blocks: int
    n_prop_steps: int
    fields: torch.Tensor
    zetas: torch.Tensor
    n_exp_terms: int = 6
    weight_floor: float = 1.0e-3
    weight_cap: float = 100.0
    damping: float = 0.1
    shift_ema: float = 0.1


def device_random_numbers(seed: int, n_blocks: int, n_prop_steps: int, n_walkers: int, naux: int,
                          device):
    """The fields and comb uniforms a run with this seed draws, in the
    reference's order (prop/afqmc.py:86-88 per step, prop/blocks.py:81-82 per
    block), from the device Threefry stream."""
    key = kernels.rng_key(seed, device)
    fields = torch.empty((n_blocks * n_prop_steps, n_walkers, naux), dtype=f64, device=device)
    zetas = torch.empty((n_blocks,), dtype=f64, device=device)
    for b in range(n_blocks):
        for t in range(n_prop_steps):
            key, sub = kernels.rng_split(key)
            kernels.rng_normal(sub, (n_walkers, naux), out=fields[b * n_prop_steps + t])
        key, sub = kernels.rng_split(key)
        kernels.rng_uniform(sub, out=zetas[b])
    return fields, zetas


def setup_tensors(h1: torch.Tensor, chol: torch.Tensor, trial: torch.Tensor, h0: float, dt: float):
    """(exp(-dt/2 h1_eff), m, h0_prop, rot, rot_h1) with mf = i m, as
    differentiable torch functions of h1 and chol (prop/chol_afqmc_ops.py:59-103,
    meas/rhf.py:141-150, RHF trial: the total rdm1 is 2 C C^T)."""
    # (each contraction as one GEMM over the flattened Cholesky index, the same
    # general formula as the einsums of the reference, so autograd's adjoints
    # of chol are the reference's)
    g, n, _ = chol.shape
    p = trial @ trial.T
    m = chol.reshape(g, n * n) @ (2.0 * p).T.reshape(n * n)        # sum_ij L_g[i,j] (2P)[j,i]
    h0_prop = -h0 + 0.5 * torch.sum(m * m)
    rows = chol.permute(1, 0, 2).reshape(n, g * n)                  # [i, (g, k)] = L_g[i, k]
    ll = rows @ chol.reshape(g * n, n)                               # sum_g L_g L_g
    h1_eff = h1 - 0.5 * ll + (m @ chol.reshape(g, n * n)).view(n, n)
    e_half = torch.linalg.matrix_exp(-0.5 * dt * h1_eff)
    rot = (trial.T @ rows).view(trial.shape[1], g, n).permute(1, 0, 2)  # C^T L_g
    rot_h1 = trial.T @ h1
    return e_half, m, h0_prop, rot, rot_h1


class _layouts:
    """Device operands of the HIP kernels, built from the set-up tensors."""

    def __init__(self, e_half, m, h0_prop, chol, rot, rot_h1, run: adjoint_run):
        g, n, _ = chol.shape
        o = rot.shape[1]
        self.n, self.o, self.g, self.nw = n, o, g, run.n_walkers
        ldv = kernels.ldv_for(n)
        self.e_half = e_half.contiguous()
        self.ldv = ldv
        self.e_pad = torch.zeros((n, ldv), dtype=f64, device=chol.device)
        self.e_pad[:, :n] = e_half
        self.e_pad_t = torch.zeros((n, ldv), dtype=f64, device=chol.device)
        self.e_pad_t[:, :n] = e_half.T
        # the Trotter adjoint's workspace: forward terms and scaled adjoints
        tc = (o + 15) // 16 * 16
        self.trotter_ws = torch.empty((2 * (2 * run.n_exp_terms + 2) * run.n_walkers * n * tc,), dtype=f64,
                                      device=chol.device)
        self.e_half_bar = torch.empty((run.n_walkers, n, n), dtype=f64, device=chol.device)
        self.chol = chol.contiguous()
        self.chol_sym = chol_sym_rows(chol)                  # B operand of the symmetric VHS
        self.mf = (1.0j * m).to(c128).contiguous()
        self.h0_prop = float(h0_prop)
        self.rot = rot.contiguous()
        self.rot_flat = self.rot.reshape(g, o * n)
        self.rot_flat_t = self.rot_flat.T.contiguous()       # (o*n, naux): B operand of G_bar
        self.chol_flat = self.chol.reshape(g, n * n)          # B operand of xbar_bar
        opad = (o + 15) // 16 * 16
        self.rot_pad = torch.zeros((1, g, opad, ldv), dtype=f64, device=chol.device)
        self.rot_pad[0, :, :o, :n] = rot
        self.rot_h1 = rot_h1.contiguous()
        self.rot_h1_k = self.rot_h1.reshape(1, o, n)
        self.trial = run.trial.contiguous()
        self._y = None

    def energy_adj_y(self):
        """Y[(b,q),(a,p)] = sum_g R[g,a,q] R[g,b,p] of the energy adjoint, built
        on first use."""
        if self._y is None:
            r = self.rot.reshape(self.g, self.o * self.n)
            self._y = _regroup(r.T @ r, self.o, self.n).contiguous()
        return self._y


def _dgemm(A, B, C, alpha: float = 1.0, beta: float = 0.0):
    """C = alpha A B^T + beta C on the library's fp64 MFMA GEMM; when the 128 x
    128 tiles of C alone would leave the GPU mostly idle, K is cut into chunks
    that run side by side in one launch (afq_dgemm_nt_splitk) and are summed in
    order."""
    M, K = A.shape
    N = B.shape[0]
    tiles = -(-M // 128) * -(-N // 128)
    splits = 1
    while tiles * splits < 512 and K // (2 * splits) >= 64 and splits < 16:
        splits *= 2
    if splits == 1:
        return kernels.dgemm_nt(A, B, C, alpha=alpha, beta=beta)
    ws = torch.empty((splits * M * N,), dtype=f64, device=A.device)
    out = C if beta == 0.0 else torch.empty_like(C)
    kernels._call("afq_dgemm_nt_splitk", A, M, N, K, float(alpha), A.data_ptr(), K, B.data_ptr(), K,
                  out.data_ptr(), N, splits, ws.data_ptr(), kernels._stream(A))
    if beta != 0.0:
        C.mul_(beta).add_(out)
    return C


def _gc(planes, nw, o, n):
    """Green's-function planes (2, nw, o*n) as complex (nw, o, n)."""
    return torch.complex(planes[0], planes[1]).view(nw, o, n)


def _energies(lay: _layouts, run: adjoint_run, greens):
    return kernels.energy(greens, lay.nw, lay.n, lay.o, 1, lay.rot_pad, lay.rot_h1_k, run.h0)


def _block(lay: _layouts, run: adjoint_run, st, b: int, e_est: float, record: bool):
    """afqmc_block (prop/blocks.py:22-95) from state st = (walkers, weights,
    log-overlaps, pop shift, Green's planes); with ``record`` also the per-step
    and block-end intermediates the adjoint needs."""
    w, wt, lo, pop, gpl = st
    sd = run.dt ** 0.5
    steps = []
    for t in range(run.n_prop_steps):
        x = run.fields[b * run.n_prop_steps + t]
        fb = kernels.force_bias(gpl, lay.rot_flat, 2.0)
        xbar, shift, fbt = kernels.field_shifts(x, fb, lay.mf, sd)
        vhs = kernels.vhs_sym(xbar.view(2 * lay.nw, -1), lay.chol_sym, lay.n).view(2, lay.nw, lay.n, lay.ldv)
        w_new = kernels.trotter(w, vhs, lay.e_pad, sd, run.n_exp_terms)
        g_new, ld_new = kernels.overlap_greens(w_new, lay.trial)
        lo_new, wt_new, pop_new = kernels.phaseless_update(
            ld_new, lo, shift, fbt, wt, pop, e_estimate=e_est, h0_prop=lay.h0_prop, dt=run.dt,
            sqrt_dt=sd, weight_floor=run.weight_floor, weight_cap=run.weight_cap,
            damping=run.damping)
        if record:
            steps.append(dict(w=w, g=gpl, lo=lo, wt=wt, pop=pop, x=x, xbar=xbar, shift=shift,
                              fbt=fbt, vhs=vhs, w_new=w_new, g_new=g_new, ld_new=ld_new))
        w, gpl, lo, wt, pop = w_new, g_new, lo_new, wt_new, pop_new
    q = kernels.orthonormalize(w)
    gq, ldq = kernels.overlap_greens(q, lay.trial)
    eloc = _energies(lay, run, gq)
    be = kernels.block_energy(eloc, wt, pop, e_ref=e_est, thresh=(2.0 / run.dt) ** 0.5,
                              shift_ema=run.shift_ema)
    idx, wt_sr = kernels.comb_indices(wt, run.zetas[b])
    w_sr = kernels.gather_rows(q, idx)
    g_sr = kernels.gather_rows(gq, idx, nplanes=2)
    lo_sr = (2.0 * ldq[:, 0])[idx.long()]
    end = None
    if record:
        end = dict(w=w, wt=wt, q=q, gq=gq, eloc=eloc, e_b=be[0], w_b=be[1], idx=idx)
    return (w_sr, wt_sr, lo_sr, be[2].clone(), g_sr), be[0], be[1], idx, steps, end


def _initial_state(lay: _layouts, run: adjoint_run):
    trial = lay.trial
    p = trial @ trial.T
    w0 = wk.init_walkers(system(lay.n, (lay.o, lay.o), "restricted"), torch.stack([p, p]),
                         run.n_walkers)
    g0, ld0 = kernels.overlap_greens(w0, trial)
    e_est = torch.mean(_energies(lay, run, g0).real)
    wt = torch.ones((run.n_walkers,), dtype=f64, device=trial.device)
    return (w0, wt, 2.0 * ld0[:, 0], e_est.clone(), g0), e_est


def _forward(lay: _layouts, run: adjoint_run, keep_checkpoints: bool):
    st, e_est = _initial_state(lay, run)
    e_est_f = float(e_est)
    chk, eb, wb, idx_all = [], [], [], []
    last = None
    for b in range(run.n_blocks):
        if keep_checkpoints:
            chk.append(st)
        # the last segment's intermediates are kept, not recomputed
        rec = keep_checkpoints and b == run.n_blocks - 1
        st, e_b, w_b, idx, steps, end = _block(lay, run, st, b, e_est_f, record=rec)
        if rec:
            last = (steps, end)
        eb.append(e_b)
        wb.append(w_b)
        idx_all.append(idx)
    eb = torch.stack(eb)
    wb = torch.stack(wb)
    j = torch.sum(wb * eb) / torch.sum(wb)
    return dict(j=j, eb=eb, wb=wb, idx=idx_all, chk=chk, e_est=e_est_f, last=last)


# --------------------------------------------------------------------------
# adjoints of the pieces
# --------------------------------------------------------------------------

def _ct(x):
    return x.conj().transpose(-1, -2)


def _greens_adj(w, trial, g_bar=None, ld_bar=None, out=None):
    """w_bar of G = (w m^-1)^T (adjoint g_bar, (nw, o, n)) and log det m
    (adjoint ld_bar), m = C^T w:  w_bar = [G_bar^T - C (conj(G) G_bar^T -
    ld_bar I)] m^-H, on afq_greens_adjoint; added to ``out`` when given."""
    nw, n, o = w.shape
    acc = out is not None
    if out is None:
        out = torch.empty_like(w)
    gb = g_bar.contiguous() if g_bar is not None else None
    lb = ld_bar.contiguous() if ld_bar is not None else None
    kernels._call("afq_greens_adjoint", w, nw, n, o, w.data_ptr(), trial.data_ptr(),
                  gb.data_ptr() if gb is not None else None, lb.data_ptr() if lb is not None else None,
                  out.data_ptr(), int(acc), kernels._stream(w))
    return out


def _qr_adj(a, q, q_bar):
    """A_bar of Q = qr(A) (reduced, diag(R) real): [(I - Q Q^H) Q_bar + Q N] R^-H."""
    qh = _ct(q)
    r = torch.triu(qh @ a)
    k = qh @ q_bar
    nmat = torch.tril(k - _ct(k), -1) + 1.0j * torch.diag_embed(torch.diagonal(k, dim1=1, dim2=2).imag)
    x = q_bar - q @ k + q @ nmat
    return _ct(torch.linalg.solve_triangular(r, _ct(x), upper=True))


def _energy_adj(g_c, e_bar, lay: _layouts, acc):
    """G_bar of the local energies (meas/rhf.py:36-50) for real adjoints e_bar of
    Re E; accumulates rot_bar and rot_h1_bar.  The half-rotated integrals R are
    real, so every product runs on the real and imaginary planes of G.  With
    f_g = R_g G_w^T (o x o) and c_g = tr f_g:
      G_bar_w = e_bar_w conj(2 rot_h1 + 4 sum_g c_g R_g - 2 sum_g f_g R_g)
      rot_bar_g += Re sum_w e_bar_w (4 c_g G_w - 2 f_g^T G_w)
    The sums over g never form f_g: sum_g f_g R_g = G_w Y with the run's fixed
    Y[(b,q),(a,p)] = sum_g R[g,a,q] R[g,b,p] (the rows of R^T R regrouped), and
    sum_w e_w f_g^T G_w = R_g Z with Z[(b,q),(a,p)] = sum_w e_w G_w[a,q] G_w[b,p]
    (the rows of (e G)^T G regrouped), so each term is one GEMM over all
    walkers on (o n)-wide operands."""
    nw, o, n = g_c.shape
    g = lay.g
    on = o * n
    r = lay.rot.reshape(g, on)                          # rows g, columns (a, p)
    y = lay.energy_adj_y()
    de = []
    rot_acc = torch.zeros((g, on), dtype=f64, device=g_c.device)
    for k, gk in enumerate((g_c.real.reshape(nw, on), g_c.imag.reshape(nw, on))):
        sign = 1.0 if k == 0 else -1.0
        c = r @ gk.T                                    # (g, nw): c_g of each walker
        d = (gk @ y).mul_(-2.0)
        d.addmm_(c.T, r, alpha=4.0)
        de.append(d.view(nw, o, n))
        ge = gk * e_bar[:, None]
        z = _regroup(ge.T @ gk, o, n)                   # Z[(b,q),(a,p)]
        rot_acc.addmm_(r, z, alpha=-2.0 * sign)
        rot_acc.addmm_(c * e_bar[None, :], gk, alpha=4.0 * sign)
    out = torch.complex(e_bar[:, None, None] * (de[0] + 2.0 * lay.rot_h1), -e_bar[:, None, None] * de[1])
    acc["rot_h1"] += 2.0 * (e_bar @ g_c.real.reshape(nw, on)).view(o, n)
    acc["rot"] += rot_acc.view(g, o, n)
    return out


def _regroup(x, o: int, n: int):
    """X[(a,q),(b,p)] -> X[(b,q),(a,p)] for an (o n) x (o n) matrix."""
    return x.view(o, n, o, n).permute(2, 1, 0, 3).reshape(o * n, o * n)


def _trotter_adj(lay: _layouts, run: adjoint_run, w_in, vhs_planes, w_bar_out, acc):
    """(w_bar_in, VHS_bar planes (2, nw, n, n)) of w' = E T(E w) on the fused
    HIP adjoint (afq_trotter_adjoint: forward terms recomputed with VHS on chip,
    reverse recursion, VHS_bar = sum_k t_bar_k t_{k-1}^H); accumulates
    E_bar = Re(w'_bar u^H + t_bar_0 w^H), summed over walkers."""
    nw, n, o = w_in.shape
    w_bar_out = w_bar_out.contiguous()
    w_bar_in = torch.empty_like(w_in)
    vhs_bar = torch.empty((2, nw, n, n), dtype=f64, device=w_in.device)
    kernels._call("afq_trotter_adjoint", w_in, nw, n, o, lay.ldv, run.n_exp_terms, run.dt ** 0.5,
                  lay.e_pad.data_ptr(), lay.e_pad_t.data_ptr(), vhs_planes.data_ptr(), w_in.data_ptr(),
                  w_bar_out.data_ptr(), w_bar_in.data_ptr(), vhs_bar.data_ptr(), lay.e_half_bar.data_ptr(),
                  lay.trotter_ws.data_ptr(), kernels._stream(w_in))
    acc["e_half"] += lay.e_half_bar.sum(0)
    return w_bar_in, vhs_bar


def _step_adj(lay: _layouts, run: adjoint_run, tp, adj, e_est: float, acc):
    """Adjoint of one afqmc_step (prop/afqmc.py:74-147)."""
    w_bar, wt_bar, lo_bar, pop_bar = adj
    nw, n, o, g = lay.nw, lay.n, lay.o, lay.g
    dt, sd = run.dt, run.dt ** 0.5
    # phaseless update
    lo_new = 2.0 * tp["ld_new"][:, 0]
    z = -sd * tp["shift"] + tp["fbt"] + dt * (tp["pop"] + lay.h0_prop) + lo_new - tp["lo"]
    ph = -sd * tp["shift"] + lo_new - tp["lo"]
    mag = torch.exp(z.real)
    cs, sn = torch.cos(ph.imag), torch.sin(ph.imag)
    imp = mag * cs
    keep1 = ~(imp < run.weight_floor) & ~(imp > run.weight_cap)
    s = torch.where(keep1, imp, torch.zeros_like(imp))
    raw = tp["wt"] * s
    keep2 = ~(raw > run.weight_cap)
    avg = torch.where(keep2, raw, torch.zeros_like(raw)).mean()
    m_bar = torch.where(avg > 1.0e-300, -run.damping / (dt * avg.clamp(min=1.0e-300)) * pop_bar,
                        torch.zeros_like(avg))
    gw = torch.where(keep2, wt_bar + m_bar / nw, torch.zeros_like(raw))
    wt_bar_in = gw * s
    s_bar = torch.where(keep1, gw * tp["wt"], torch.zeros_like(raw))
    zr = s_bar * mag * cs
    yi = -s_bar * mag * sn
    amix = torch.complex(zr, yi)
    shift_bar = -sd * amix
    fbt_bar = torch.complex(zr, torch.zeros_like(zr))
    pop_bar_in = dt * zr.sum()
    acc["h0_prop"] += dt * zr.sum()
    acc["e_est"] += pop_bar
    lo_bar_in = -amix
    ld_bar_new = 2.0 * (lo_bar + amix)
    # (w_bar is this pass's own tensor: accumulated in place)
    w_bar = _greens_adj(tp["w_new"], lay.trial, ld_bar=ld_bar_new, out=w_bar.resolve_conj().contiguous())
    # Trotter
    w_bar_in, vhs_bar = _trotter_adj(lay, run, tp["w"], tp["vhs"], w_bar, acc)
    # VHS = sum_g xbar_g L_g: xbar_bar = VHS_bar . L (real and imaginary planes
    # as rows) on the fp64 MFMA GEMM, and L_bar += Re(xbar)^T Re(VHS_bar) +
    # Im(xbar)^T Im(VHS_bar) as one library GEMM over K = 2 nw
    v2 = vhs_bar.reshape(2 * nw, n * n)
    xbb = torch.empty((2 * nw, g), dtype=f64, device=v2.device)
    _dgemm(v2, lay.chol_flat, xbb)
    xbar_bar = torch.complex(xbb[:nw], xbb[nw:])
    xbar_c = torch.complex(tp["xbar"][0], tp["xbar"][1])
    cb = acc["chol"].view(g, n * n)
    cb.addmm_(tp["xbar"].reshape(2 * nw, g).T, v2)
    # field shifts: fs = -sqrt_dt (i fb - mf), xbar = x - fs
    mf = lay.mf
    xbar_bar = xbar_bar + mf.conj()[None, :] * shift_bar[:, None]
    mf_bar = (xbar_c.conj() * shift_bar[:, None]).sum(0)
    fs_bar = xbar_c.conj() * fbt_bar[:, None] - xbar_bar
    fb_bar = 1.0j * sd * fs_bar
    mf_bar = mf_bar + sd * fs_bar.sum(0)
    acc["m"] += mf_bar.imag
    # force bias fb = 2 rot . vec(G): G_bar = 2 fb_bar . rot, rot_bar += 2 Re(fb_bar^H G)
    fb2 = torch.stack([fb_bar.real, fb_bar.imag]).reshape(2 * nw, g)
    gb = torch.empty((2 * nw, o * n), dtype=f64, device=fb2.device)
    _dgemm(fb2, lay.rot_flat_t, gb, alpha=2.0)
    g_bar = torch.complex(gb[:nw], gb[nw:]).view(nw, o, n)
    rb = acc["rot"].view(g, o * n)
    rb.addmm_(fb2.T, tp["g"].reshape(2 * nw, o * n), alpha=2.0)
    w_bar_in = _greens_adj(tp["w"], lay.trial, g_bar=g_bar, out=w_bar_in)
    return (w_bar_in, wt_bar_in, lo_bar_in, pop_bar_in)


def _block_end_adj(lay: _layouts, run: adjoint_run, end, adj, e_b_bar, w_b_bar, e_est: float, acc):
    """Adjoint of QR, local energy, block energy / shift EMA and comb
    (prop/blocks.py:53-92)."""
    w_bar_sr, wt_bar_sr, lo_bar_sr, pop_bar_b = adj
    nw, n, o = lay.nw, lay.n, lay.o
    idx = end["idx"].long()
    q_bar = torch.zeros_like(end["q"]).index_add_(0, idx, w_bar_sr)
    lo_bar_q = torch.zeros((nw,), dtype=c128, device=idx.device).index_add_(0, idx, lo_bar_sr)
    wt = end["wt"]
    wt_bar = torch.sign(wt) * (wt_bar_sr.sum() / nw)
    alpha = run.shift_ema
    pop_bar = (1.0 - alpha) * pop_bar_b
    eb_bar = e_b_bar + alpha * pop_bar_b
    e_re = end["eloc"].real
    clipped = torch.abs(e_re - e_est) > (2.0 / run.dt) ** 0.5
    e_cl = torch.where(clipped, torch.full_like(e_re, e_est), e_re)
    w_b, e_b = end["w_b"], end["e_b"]
    if float(w_b) != 0.0:
        e_i_bar = wt * eb_bar / w_b
        wt_bar = wt_bar + (e_cl - e_b) * eb_bar / w_b + w_b_bar
    else:
        acc["e_est"] += eb_bar
        e_i_bar = torch.zeros_like(wt)
        wt_bar = wt_bar + w_b_bar
    acc["e_est"] += torch.where(clipped, e_i_bar, torch.zeros_like(e_i_bar)).sum()
    e_i_bar = torch.where(clipped, torch.zeros_like(e_i_bar), e_i_bar)
    gq_c = _gc(end["gq"], nw, o, n)
    gq_bar = _energy_adj(gq_c, e_i_bar, lay, acc)
    q_bar = _greens_adj(end["q"], lay.trial, g_bar=gq_bar, ld_bar=2.0 * lo_bar_q, out=q_bar)
    w_bar = _qr_adj(end["w"], end["q"], q_bar)
    return (w_bar, wt_bar, torch.zeros((nw,), dtype=c128, device=idx.device), pop_bar)


class _RunEnergy(torch.autograd.Function):
    """J(e_half, m, h0_prop, chol, rot, rot_h1) with a hand-written backward."""

    @staticmethod
    def forward(ctx, e_half, m, h0_prop, chol, rot, rot_h1, run):
        lay = _layouts(e_half.detach(), m.detach(), h0_prop.detach(), chol.detach(), rot.detach(),
                       rot_h1.detach(), run)
        fw = _forward(lay, run, keep_checkpoints=True)
        ctx.lay, ctx.run, ctx.fw = lay, run, fw
        return fw["j"].clone()

    @staticmethod
    def backward(ctx, gj):
        lay, run, fw = ctx.lay, ctx.run, ctx.fw
        dev = lay.chol.device
        nw, n, o, g = lay.nw, lay.n, lay.o, lay.g
        acc = dict(e_half=torch.zeros((n, n), dtype=f64, device=dev),
                   m=torch.zeros((g,), dtype=f64, device=dev),
                   h0_prop=torch.zeros((), dtype=f64, device=dev),
                   chol=torch.zeros((g, n, n), dtype=f64, device=dev),
                   rot=torch.zeros((g, o, n), dtype=f64, device=dev),
                   rot_h1=torch.zeros((o, n), dtype=f64, device=dev),
                   e_est=torch.zeros((), dtype=f64, device=dev))
        eb, wb, j = fw["eb"], fw["wb"], fw["j"]
        ntot = wb.sum()
        eb_bar = gj * wb / ntot
        wb_bar = gj * (eb - j) / ntot
        e_est = fw["e_est"]
        adj = (torch.zeros((nw, n, o), dtype=c128, device=dev), torch.zeros((nw,), dtype=f64, device=dev),
               torch.zeros((nw,), dtype=c128, device=dev), torch.zeros((), dtype=f64, device=dev))
        for b in range(run.n_blocks - 1, -1, -1):
            if b == run.n_blocks - 1:
                steps, end = fw["last"]
                fw["last"] = None
            else:
                _, _, _, _, steps, end = _block(lay, run, fw["chk"][b], b, e_est, record=True)
            adj = _block_end_adj(lay, run, end, adj, eb_bar[b], wb_bar[b], e_est, acc)
            for tp in reversed(steps):
                adj = _step_adj(lay, run, tp, adj, e_est, acc)
            del steps, end
        # initial state: pop shift = e_estimate = mean Re E(w0)
        acc["e_est"] += adj[3]
        g0 = _gc(fw["chk"][0][4], nw, o, n)
        _energy_adj(g0, (acc["e_est"] / nw).expand(nw), lay, acc)
        return acc["e_half"], acc["m"], acc["h0_prop"], acc["chol"], acc["rot"], acc["rot_h1"], None


def _check_run(h1, chol, run: adjoint_run):
    if h1.dtype != f64 or chol.dtype != f64:
        raise TypeError("h1 and chol must be float64")
    if h1.device.type != "cuda" or chol.device != h1.device:
        raise ValueError("h1 and chol must be HIP tensors on one device")
    g, n, n2 = chol.shape
    if n2 != n or tuple(h1.shape) != (n, n):
        raise ValueError(f"shapes h1 {tuple(h1.shape)} / chol {tuple(chol.shape)} do not match")
    if tuple(run.trial.shape[:1]) != (n,) or run.trial.dtype != f64:
        raise ValueError("trial must be float64 (n, o)")
    if n > 128 or run.trial.shape[1] > 32:
        raise ValueError(f"the Trotter adjoint kernel takes n <= 128 and o <= 32, got n={n} "
                         f"o={run.trial.shape[1]}")
    steps = run.n_blocks * run.n_prop_steps
    if tuple(run.fields.shape) != (steps, run.n_walkers, g) or run.fields.dtype != f64:
        raise ValueError(f"fields must be float64 {(steps, run.n_walkers, g)}")
    if tuple(run.zetas.shape) != (run.n_blocks,) or run.zetas.dtype != f64:
        raise ValueError(f"zetas must be float64 ({run.n_blocks},)")


def run_energy(h1: torch.Tensor, chol: torch.Tensor, run: adjoint_run) -> torch.Tensor:
    """The run's energy estimate J as a 0-d tensor, differentiable in h1 and chol
    (backward = the hand-written adjoint; the set-up goes through autograd)."""
    _check_run(h1, chol, run)
    e_half, m, h0_prop, rot, rot_h1 = setup_tensors(h1, chol, run.trial, run.h0, run.dt)
    return _RunEnergy.apply(e_half, m, h0_prop, chol, rot, rot_h1, run)


def run_blocks(h1: torch.Tensor, chol: torch.Tensor, run: adjoint_run):
    """Forward only: (J, block energies, block weights, comb indices per block)."""
    _check_run(h1, chol, run)
    with torch.no_grad():
        e_half, m, h0_prop, rot, rot_h1 = setup_tensors(h1, chol, run.trial, run.h0, run.dt)
        lay = _layouts(e_half, m, h0_prop, chol, rot, rot_h1, run)
        fw = _forward(lay, run, keep_checkpoints=False)
    return float(fw["j"]), fw["eb"], fw["wb"], fw["idx"]


def energy_gradients(h1: torch.Tensor, chol: torch.Tensor, run: adjoint_run, v: torch.Tensor | None = None):
    """J, dJ/dh1, dJ/dL and, for a one-body perturbation h1 + lambda V, dJ/dlambda."""
    h1v = h1.detach().clone().requires_grad_(True)
    lv = chol.detach().clone().requires_grad_(True)
    j = run_energy(h1v, lv, run)
    gh, gl = torch.autograd.grad(j, (h1v, lv))
    out = dict(energy=float(j.detach()), d_h1=gh, d_chol=gl)
    if v is not None:
        out["d_lambda"] = float(torch.sum(gh * v))
    return out
