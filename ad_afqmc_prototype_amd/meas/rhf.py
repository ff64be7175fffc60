"""Half-rotated measurements for RHF/UHF trials (meas/rhf.py:15-176 of the
reference), batched over walkers on the HIP kernels.

Force bias: fb = alpha * rot_chol_flat . vec(G) as one fp64 MFMA GEMM over the
whole walker batch (alpha = 2 for restricted walkers, meas/rhf.py:33;
spin sectors side by side along K for unrestricted walkers, meas/rhf.py:71-72).
Energy: meas/rhf.py:36-50 / 75-105.  Where n < naux the exchange
sum_g tr(f_g f_g) is the quadratic form vec(G)^T Y vec(G) with the fixed
Y = sum_g R_g (x) R_g regrouped (one GEMM over all walkers); otherwise the
(naux, o, o) intermediate is reduced inside the fused kernel.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import torch

from .. import kernels
from ..core.ops import k_energy, k_force_bias, meas_ops
from ..core.system import system
from ..ham.chol import ham_chol
from ..trial import rhf as trial_rhf


@dataclass(frozen=True)
class rhf_meas_ctx:
    """Half-rotated integrals per trial spin sector: rot_h1 (2, o, n),
    rot_chol (2, naux, o, n) (both sectors alias one another for an RHF trial),
    rot_chol_flat = rot_chol[0] viewed as (naux, o*n), plus device layouts
    built on first use (force-bias operand per walker kind, padded energy
    operand, exchange quadratic form).  exchange: "auto" measures the exchange
    as a quadratic form where that is cheaper (quadratic_exchange), "per_g" on
    the fused per-g kernels always."""

    rot_h1: torch.Tensor
    rot_chol: torch.Tensor
    rot_chol_flat: torch.Tensor
    precision: str = "fp64"
    exchange: str = "auto"
    _cache: dict = field(default_factory=dict, compare=False, repr=False)

    def fb_operand(self, nspin: int):
        """B operand of the force-bias GEMM: float64 (naux, k); float32 copy in
        fp32 mode; (hi, lo) bf16 planes (naux, k padded to 8) in bf16x3 mode.
        For complex trial orbitals a pair (real part, imaginary part) of them."""
        key = ("fb", nspin)
        if key not in self._cache:
            if nspin == 1:
                op = self.rot_chol_flat
            else:
                g = self.rot_chol.shape[1]
                op = torch.cat([self.rot_chol[0].reshape(g, -1),
                                self.rot_chol[1].reshape(g, -1)], dim=1).contiguous()

            def fmt(x):
                x = x.contiguous()
                if self.precision == "fp32":
                    return kernels.convert_f32(x)
                if self.precision == "bf16x3":
                    return kernels.split_bf16(x)
                return x
            self._cache[key] = (fmt(op.real), fmt(op.imag)) if op.is_complex() else fmt(op)
        return self._cache[key]

    @property
    def is_complex(self) -> bool:
        return self.rot_chol.is_complex()

    def exchange_operand(self, s: int):
        """L of spin sector s, (o n, o n): the lower triangle of
        Y[(a,q),(b,p)] = sum_g R[g,a,p] R[g,b,q] (R = rot_chol[s], real) with its
        diagonal halved, so Y = L + L^T.  The exchange sum_g tr(f_g f_g) of a
        walker with f_g = R_g G^T is the quadratic form vec(G)^T Y vec(G) =
        2 vec(G)^T L^T vec(G) (Y is symmetric), one GEMM Z = 2 G L^T whose
        column tiles stop at the diagonal (kernels.gemm_nt_lower).  float64;
        float32 in fp32 mode; (hi, lo) bf16 planes in bf16x3 mode."""
        key = ("exchange", s)
        if key not in self._cache:
            y = exchange_lower(self.rot_chol[s])
            if self.precision == "fp32":
                y = kernels.convert_f32(y)
            elif self.precision == "bf16x3":
                y = kernels.split_bf16(y)
            self._cache[key] = y
        return self._cache[key]

    def energy_operands(self, nspin: int):
        """(rot_chol_pad, rot_h1 (nspin, o, n)) for the energy kernels:
        rot_chol_pad is (nspin, naux, opad, ldr) float64 / float32 in fp64 /
        fp32 mode, and a (hi, lo) pair of bf16 planes (nspin * naux * opad,
        ldr32) with ldr32 = n rounded up to 32 in bf16x3 mode."""
        key = ("energy", nspin)
        if key not in self._cache:
            _, g, o, n = self.rot_chol.shape
            opad = (o + 15) // 16 * 16
            if self.precision == "bf16x3":
                ldr = (n + 31) // 32 * 32
                pad = torch.zeros((nspin, g, opad, ldr), dtype=torch.float64, device=self.rot_chol.device)
                pad[:, :, :o, :n] = self.rot_chol[:nspin]
                op = kernels.split_bf16(pad.reshape(-1, ldr))
                del pad
            else:
                ldr = kernels.ldv_for(n)
                dt = torch.float64 if self.precision == "fp64" else torch.float32
                op = torch.zeros((nspin, g, opad, ldr), dtype=dt, device=self.rot_chol.device)
                op[:, :, :o, :n] = self.rot_chol[:nspin]
            self._cache[key] = (op, self.rot_h1[:nspin].contiguous())
        return self._cache[key]


def exchange_lower(rot: torch.Tensor) -> torch.Tensor:
    """L (o n, o n) for the half-rotated Cholesky rot (naux, o, n), real: the
    lower triangle of Y[(a,q),(b,p)] = sum_g rot[g,a,p] rot[g,b,q] with its
    diagonal halved, so Y = L + L^T and the exchange vec(G)^T Y vec(G) =
    2 vec(G)^T L^T vec(G)."""
    g, o, n = rot.shape
    r = rot.reshape(g, o * n)
    y = torch.tril((r.T @ r).view(o, n, o, n).permute(0, 3, 2, 1).reshape(o * n, o * n))
    y.diagonal().mul_(0.5)
    return y


def build_meas_ctx(ham_data: ham_chol, trial_data, precision: str = "fp64",
                   exchange: str = "auto") -> rhf_meas_ctx:
    """meas/rhf.py:141-150 (per spin sector for a UHF trial)."""
    if ham_data.basis != "restricted":
        raise ValueError("RHF meas_ops currently assumes ham_chol.basis == 'restricted'.")
    cs = [trial_data.mo_up] if trial_data.restricted else [trial_data.mo_up, trial_data.mo_dn]
    # C^H h1 and C^H L_g (complex for complex trial orbitals)
    rot_h1 = [c.conj().T @ ham_data.h1.to(c.dtype) for c in cs]
    rot_chol = [torch.einsum("pi,gpj->gij", c.conj(), ham_data.chol.to(c.dtype)).contiguous() for c in cs]
    if trial_data.restricted:
        rh = torch.stack([rot_h1[0], rot_h1[0]])
        rc = rot_chol[0].unsqueeze(0).expand(2, *rot_chol[0].shape)
    else:
        rh = torch.stack(rot_h1)
        rc = torch.stack(rot_chol)
    flat = rc[0].reshape(rc.shape[1], -1)
    if exchange not in ("auto", "per_g"):
        raise ValueError(f"unknown exchange {exchange!r}")
    return rhf_meas_ctx(rot_h1=rh.contiguous(), rot_chol=rc, rot_chol_flat=flat, precision=precision,
                        exchange=exchange)


def _nspin(walkers) -> int:
    return 2 if isinstance(walkers, tuple) else 1


def deferred_view(meas_ctx: rhf_meas_ctx, trial_data, e_half: torch.Tensor):
    """(trial, meas ctx) for walkers held as E^-1 W (a step that left out its
    final one-body half E, symmetric; prop_state.deferred).  The overlap matrix
    C^T W = (E C)^T (E^-1 W), so the trial becomes E C: its log det and the
    rows G' = ((E^-1 W) m^-1)^T = G E^-1 are what the Green's function kernels
    compute from the held walkers.  Every contraction of G with a half-rotated
    operand X (o x n) is one with X E of G': sum_ij X_ij G_ij = sum_ij (X E)_ij
    G'_ij, so the context's rot_h1 and rot_chol become rot_h1 E and
    rot_chol_g E (force bias, energy).  Built once per (context, trial, E)."""
    key = ("deferred", id(trial_data), id(e_half))
    hit = meas_ctx._cache.get(key)
    if hit is not None and hit[0] is trial_data and hit[1] is e_half:
        return hit[2], hit[3]
    e = e_half.to(meas_ctx.rot_chol.dtype)
    if trial_data.restricted:
        trial_e = trial_rhf.rhf_trial(e @ trial_data.mo_coeff)
    else:
        trial_e = trial_rhf.uhf_trial(e @ trial_data.mo_up, e @ trial_data.mo_dn)
    rot_h1 = (meas_ctx.rot_h1 @ e).contiguous()
    rot_chol = (meas_ctx.rot_chol @ e).contiguous()
    flat = rot_chol[0].reshape(rot_chol.shape[1], -1)
    ctx_e = rhf_meas_ctx(rot_h1=rot_h1, rot_chol=rot_chol, rot_chol_flat=flat,
                         precision=meas_ctx.precision, exchange=meas_ctx.exchange)
    meas_ctx._cache[key] = (trial_data, e_half, trial_e, ctx_e)
    return trial_e, ctx_e


def force_bias_planes(walkers, meas_ctx: rhf_meas_ctx, trial_data, greens=None) -> torch.Tensor:
    """Force bias as split planes (2, nw, naux)."""
    nspin = _nspin(walkers)
    if greens is None:
        greens = trial_rhf.greens(walkers, trial_data, split=meas_ctx.precision == "bf16x3")
    if greens.greens is None:
        # only the bf16 planes: the bf16x3 force bias reads nothing else
        if meas_ctx.precision != "bf16x3" or greens.split is None:
            greens = trial_rhf.greens(walkers, trial_data, split=meas_ctx.precision == "bf16x3")
        else:
            w0 = walkers[0] if nspin == 2 else walkers
            nw, n, o = w0.shape
            return _force_bias_split(greens.split, nw, nspin * o * n, meas_ctx, nspin,
                                     2.0 if nspin == 1 else 1.0)
    return _force_bias_rows(greens.greens, meas_ctx, nspin, 2.0 if nspin == 1 else 1.0, greens.split)


def _force_bias_split(split: tuple, nw: int, k: int, meas_ctx: rhf_meas_ctx, nspin: int,
                      alpha: float) -> torch.Tensor:
    """bf16x3 force bias (2, nw, naux) from the bf16 planes (2 nw, pad8(k))
    of the Green's function rows."""
    op = meas_ctx.fb_operand(nspin)
    naux = meas_ctx.rot_chol.shape[1]
    gh, gl = split
    if meas_ctx.is_complex:
        return _combine_complex_fb([_gemm_planes_bf16x3(gh, gl, o_, nw, naux, alpha) for o_ in op])
    fb = torch.empty((2, nw, naux), dtype=torch.float64, device=gh.device)
    kernels.gemm_bf16x3_nt(gh, gl, op[0], op[1], fb.view(2 * nw, naux), alpha=alpha, tag="force_bias")
    return fb


def _gemm_planes_bf16x3(gh, gl, op, nw, naux, alpha):
    out = torch.empty((2, nw, naux), dtype=torch.float64, device=gh.device)
    kernels.gemm_bf16x3_nt(gh, gl, op[0], op[1], out.view(2 * nw, naux), alpha=alpha, tag="force_bias")
    return out


def _combine_complex_fb(parts):
    """Force-bias planes with complex half-rotated Cholesky R = Rr + i Ri from
    the products of the Green's planes (Gr, Gi) with Rr and with Ri:
    fb = (Gr Rr - Gi Ri) + i (Gr Ri + Gi Rr)."""
    pr, pi = parts
    return torch.stack([pr[0] - pi[1], pi[0] + pr[1]]).contiguous()


def force_bias_planes_g(walkers, meas_ctx: rhf_meas_ctx, trial_data, greens=None) -> torch.Tensor:
    """force_bias_kernel_g (meas/rhf.py:117-130) as split planes (2, nw, naux):
    rot_chol_flat @ vec(G_upup + G_dndn) from the (2o, 2n) Green's function."""
    if greens is None:
        greens = trial_rhf.greens_g(walkers, trial_data)
    nw, n2, o2 = walkers.shape
    n, o = n2 // 2, o2 // 2
    gv = greens.greens.view(2, nw, o2, n2)
    rows = (gv[:, :, :o, :n] + gv[:, :, o:, n:]).reshape(2, nw, o * n).contiguous()
    return _force_bias_rows(rows, meas_ctx, 1, 1.0)


def _force_bias_rows(g: torch.Tensor, meas_ctx: rhf_meas_ctx, nspin: int, alpha: float,
                     split: tuple | None = None) -> torch.Tensor:
    """fb planes (2, nw, naux) = alpha * rows of g (2, nw, k) . rot^T in the ctx's precision
    (bf16x3: ``split``, the bf16 planes of g's rows when the Green's function
    kernel made them, else split here)."""
    op = meas_ctx.fb_operand(nspin)
    if meas_ctx.is_complex:
        # two real products (real and imaginary part of C^H L) combined
        return _combine_complex_fb([_force_bias_real(g, o_, meas_ctx, alpha, split) for o_ in op])
    return _force_bias_real(g, op, meas_ctx, alpha, split)


def _force_bias_real(g: torch.Tensor, op, meas_ctx: rhf_meas_ctx, alpha: float,
                     split: tuple | None = None) -> torch.Tensor:
    if meas_ctx.precision == "fp64":
        return kernels.force_bias(g, op, alpha)
    _, nw, k = g.shape
    naux = meas_ctx.rot_chol.shape[1]
    fb = torch.empty((2, nw, naux), dtype=torch.float64, device=g.device)
    flat = g.reshape(2 * nw, k)
    if meas_ctx.precision == "fp32":
        kernels.sgemm_nt(kernels.convert_f32(flat), op, fb.view(2 * nw, naux), alpha=alpha,
                         tag="force_bias")
    else:
        gh, gl = split if split is not None else kernels.split_bf16(flat)
        kernels.gemm_bf16x3_nt(gh, gl, op[0], op[1], fb.view(2 * nw, naux), alpha=alpha,
                               tag="force_bias")
    return fb


def force_bias_kernel_r(walkers, _ham, meas_ctx: rhf_meas_ctx, trial_data, greens=None):
    """meas/rhf.py:27-33, batched: (nw, naux) complex."""
    fb = force_bias_planes(walkers, meas_ctx, trial_data, greens)
    return torch.complex(fb[0], fb[1])


def force_bias_kernel_g(walkers, _ham, meas_ctx: rhf_meas_ctx, trial_data, greens=None):
    """meas/rhf.py:117-130, batched: (nw, naux) complex."""
    fb = force_bias_planes_g(walkers, meas_ctx, trial_data, greens)
    return torch.complex(fb[0], fb[1])


def force_bias_kernel_u(walkers, _ham, meas_ctx: rhf_meas_ctx, trial_data, greens=None):
    """meas/rhf.py:60-72, batched: (nw, naux) complex."""
    fb = force_bias_planes(walkers, meas_ctx, trial_data, greens)
    return torch.complex(fb[0], fb[1])


# occupied orbitals per spin up to which the energy runs in the fused kernels
ENERGY_KERNEL_MAX_O = 32


def _energy_library(g: torch.Tensor, nw: int, n: int, o: int, nspin: int, meas_ctx: rhf_meas_ctx,
                    h0: float, budget_bytes: int = 1 << 29) -> torch.Tensor:
    """meas/rhf.py:36-50 (restricted) and 75-105 (unrestricted) on the device's
    library fp64 GEMMs, for shapes the fused energy kernels do not take
    (o > 32 per spin): f_g = rot_g G^T of a chunk of walkers at a time,
    (chunk, naux, o, o) complex, then the Coulomb traces c_g = tr f_g and the
    exchange sum_g tr(f_g f_g).  Every precision measures here in fp64."""
    gc = torch.complex(g[0], g[1]).view(nw, nspin, o, n)
    naux = meas_ctx.rot_chol.shape[1]
    out = torch.empty((nw,), dtype=torch.complex128, device=g.device)
    chunk = max(1, budget_bytes // (2 * naux * o * o * 16))
    for a in range(0, nw, chunk):
        b = min(nw, a + chunk)
        cs, exs, e1 = [], [], 0.0
        for s in range(nspin):
            rot = meas_ctx.rot_chol[s].reshape(naux * o, n)
            gs = gc[a:b, s]                                    # (m, o, n)
            gt = gs.transpose(1, 2).contiguous()               # (m, n, o)
            if rot.is_complex():
                f = torch.matmul(rot, gt)
            else:
                f = torch.complex(torch.matmul(rot, gt.real.contiguous()), torch.matmul(rot, gt.imag.contiguous()))
            f = f.view(b - a, naux, o, o)
            cs.append(torch.diagonal(f, dim1=-2, dim2=-1).sum(-1))              # (m, naux)
            exs.append((f * f.transpose(-1, -2)).sum(dim=(-1, -2)).sum(-1))    # (m,)
            e1 = e1 + (gs * meas_ctx.rot_h1[s]).sum(dim=(-1, -2))
            del f
        if nspin == 1:
            e2 = 2.0 * (cs[0] * cs[0]).sum(-1) - exs[0]
            e1 = 2.0 * e1
        else:
            e2 = ((cs[0] * cs[0]).sum(-1) + (cs[1] * cs[1]).sum(-1) + 2.0 * (cs[0] * cs[1]).sum(-1)
                  - exs[0] - exs[1]) / 2.0
        out[a:b] = h0 + e1 + e2
    return out


def quadratic_exchange(n: int, o: int, naux: int) -> bool:
    """The exchange as a quadratic form costs 4 (o n)^2 per walker and sector
    against 4 naux o^2 n for the sum over g of tr(f_g f_g): cheaper whenever
    n < naux (molecular Cholesky sets have naux of 4-10 n), and Y must fit
    (4 GiB in fp64)."""
    return n < naux and (o * n) ** 2 * 8 <= (4 << 30)


def _energy_quadratic(greens, nw: int, n: int, o: int, nspin: int, meas_ctx: rhf_meas_ctx,
                      h0: float) -> torch.Tensor:
    """meas/rhf.py:36-50 / 75-105 with the same sums regrouped:
      c_g (summed over spin sectors) = the force-bias contraction with alpha 1;
      Coulomb: restricted 2 sum_g c_g^2, unrestricted (sum_g (c_up + c_dn)_g^2) / 2;
      exchange of sector s: sum_g tr(f_g f_g) = vec(G_s)^T Y_s vec(G_s), one GEMM
      Z = 2 G_s L_s^T = G_s Y_s over all walkers (L_s: Y_s's lower triangle,
      diagonal halved; the ctx's precision, fp64 output);
      one-body: tr(rot_h1_s G_s);
    the dot products and sums on afq_energy_assemble (fp64)."""
    g = greens.greens
    on = o * n
    c = _force_bias_rows(g, meas_ctx, nspin, 1.0, greens.split)
    z = torch.empty((nspin, 2 * nw, on), dtype=torch.float64, device=g.device)
    for sec in range(nspin):
        rows = g.reshape(2 * nw, nspin * on)[:, sec * on:(sec + 1) * on]
        if nspin > 1:
            rows = rows.contiguous()
        lo = meas_ctx.exchange_operand(sec)
        if meas_ctx.precision == "fp64":
            kernels.gemm_nt_lower(rows, lo, z[sec], alpha=2.0, tag="exchange")
        elif meas_ctx.precision == "fp32":
            kernels.gemm_nt_lower(kernels.convert_f32(rows), lo, z[sec], alpha=2.0, tag="exchange")
        else:
            gh, gl = greens.split if nspin == 1 and greens.split is not None else kernels.split_bf16(rows)
            kernels.gemm_nt_lower((gh, gl), lo, z[sec], alpha=2.0, tag="exchange")
    return kernels.energy_assemble(g, z, c, meas_ctx.rot_h1[:nspin].contiguous(), nw, o, n, nspin, h0)


def _energy(walkers, ham_data: ham_chol, meas_ctx: rhf_meas_ctx, trial_data, greens=None):
    nspin = _nspin(walkers)
    w0 = walkers[0] if nspin == 2 else walkers
    nw, n, o = w0.shape
    if greens is None or greens.greens is None:
        greens = trial_rhf.greens(walkers, trial_data)
    if (meas_ctx.exchange == "auto" and not meas_ctx.is_complex
            and quadratic_exchange(n, o, meas_ctx.rot_chol.shape[1])):
        return _energy_quadratic(greens, nw, n, o, nspin, meas_ctx, float(ham_data.h0))
    if o > ENERGY_KERNEL_MAX_O or meas_ctx.is_complex:
        return _energy_library(greens.greens, nw, n, o, nspin, meas_ctx, float(ham_data.h0))
    rot_pad, rot_h1 = meas_ctx.energy_operands(nspin)
    return kernels.energy(greens.greens, nw, n, o, nspin, rot_pad, rot_h1, float(ham_data.h0))


def energy_kernel_r(walkers, ham_data, meas_ctx, trial_data, greens=None):
    """meas/rhf.py:36-50, batched: (nw,) complex."""
    return _energy(walkers, ham_data, meas_ctx, trial_data, greens)


def energy_kernel_u(walkers, ham_data, meas_ctx, trial_data, greens=None):
    """meas/rhf.py:75-105 (per-spin integrals for a UHF trial), batched: (nw,) complex."""
    return _energy(walkers, ham_data, meas_ctx, trial_data, greens)


def make_rhf_meas_ops(sys: system, precision: str = "fp64", exchange: str = "auto") -> meas_ops:
    """meas/rhf.py:153-176; also serves UHF trials (per-spin integrals).
    precision selects the force-bias and energy contractions: "fp64", "fp32"
    or "bf16x3"; exchange the energy's exchange form ("auto" or "per_g", see
    rhf_meas_ctx)."""
    from functools import partial
    if precision not in ("fp64", "fp32", "bf16x3"):
        raise ValueError(f"unknown precision {precision!r}")
    if exchange not in ("auto", "per_g"):
        raise ValueError(f"unknown exchange {exchange!r}")
    build_meas_ctx_p = partial(build_meas_ctx, precision=precision, exchange=exchange)
    wk = sys.walker_kind.lower()
    # the bf16x3 force bias takes the Green's function rows as bf16 planes,
    # which the Green's function kernel writes in the same pass
    greens_p = partial(trial_rhf.greens, split=precision == "bf16x3")
    # the step's new walkers need only what the next force bias reads
    step_greens = trial_rhf.greens_split if precision == "bf16x3" else None
    if wk == "restricted":
        return meas_ops(overlap=trial_rhf.overlap_r, build_meas_ctx=build_meas_ctx_p,
                        kernels={k_force_bias: force_bias_kernel_r, k_energy: energy_kernel_r},
                        log_overlap=trial_rhf.log_overlap, greens=greens_p,
                        force_bias_planes=force_bias_planes, step_greens=step_greens)
    if wk == "unrestricted":
        return meas_ops(overlap=trial_rhf.overlap_u, build_meas_ctx=build_meas_ctx_p,
                        kernels={k_force_bias: force_bias_kernel_u, k_energy: energy_kernel_u},
                        log_overlap=trial_rhf.log_overlap, greens=greens_p,
                        force_bias_planes=force_bias_planes, step_greens=step_greens)
    if wk == "generalized":
        # meas/rhf.py:169-174: overlap and force bias, no energy kernel
        return meas_ops(overlap=trial_rhf.overlap_g, build_meas_ctx=build_meas_ctx_p,
                        kernels={k_force_bias: force_bias_kernel_g},
                        log_overlap=trial_rhf.log_overlap_g, greens=trial_rhf.greens_g,
                        force_bias_planes=force_bias_planes_g)
    raise ValueError(f"unknown walker_kind: {sys.walker_kind}")


make_uhf_meas_ops = make_rhf_meas_ops
