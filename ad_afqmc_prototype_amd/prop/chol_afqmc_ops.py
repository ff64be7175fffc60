"""Cholesky-AFQMC propagator (prop/chol_afqmc_ops.py:15-262 of the reference).

apply_trotter is batched over walkers: one fp64 MFMA GEMM builds every
walker's VHS = sum_g x_g L_g (prop/chol_afqmc_ops.py:70-74), then the fused
Trotter kernel applies exp(-dt/2 h1eff) . Taylor_n(i sqrt_dt VHS) . exp(-dt/2 h1eff)
(prop/chol_afqmc_ops.py:173-196, prop/utils.py:20-31).  n_chunks bounds the
VHS workspace (2 * chunk * n * n doubles) exactly as the reference's chunked
vmap bounds memory; results do not depend on it.

precision: "fp64" (plain step), "fp32" (VHS GEMM and Trotter step in fp32) or
"bf16x3" (VHS GEMM in split bf16 with fp32 accumulation, Trotter step in fp32).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Callable, NamedTuple

import torch

from .. import kernels
from ..ham.chol import ham_chol


@dataclass(frozen=True)
class chol_afqmc_ctx:
    dt: float
    sqrt_dt: float
    exp_h1_half: torch.Tensor       # (n, n)
    mf_shifts: torch.Tensor         # (n_fields,) complex128, purely imaginary
    h0_prop: complex
    exp_h1_half_pad: torch.Tensor   # (n, ldv) device layout of exp_h1_half
    exp_h1_half_pad32: torch.Tensor = None  # float32 copy for the mixed step
    # exp(-dt h1_eff / 2) - I as upper 16 x 16 blocks of split-bf16 pair words
    # (int32, split_pairs of the float32 entries): the bf16x3 Trotter's E operand
    exp_h1_half_sym_pairs: torch.Tensor = None
    # E^2 - I (E^2 = exp(-dt h1_eff), the product of two halves in fp64) in the
    # same layout: the first half of a step after one that left out its last
    exp_h1_sym_pairs: torch.Tensor = None


class chol_afqmc_ops(NamedTuple):
    n_fields: Callable[[], int]
    build_prop_ctx: Callable[[Any, float], chol_afqmc_ctx]
    apply_trotter: Callable[..., Any]       # (walkers, fields (nw, G), ctx, n_terms) -> walkers
    apply_trotter_planes: Callable[..., Any]  # (walkers, xbar planes (2, nw, G), ctx, n_terms)
    precision: str = "fp64"
    # the VHS contraction takes xbar as bf16 planes: the step asks the field
    # shift kernel for them (apply_trotter_planes(..., xbar_split=(hi, lo)))
    xbar_split: bool = False
    # apply_trotter_planes takes merged_first / last_half: a step may leave out
    # its final one-body half and the next apply E^2 as its first
    merges_halves: bool = False


def _as_total_rdm1_restricted(dm: torch.Tensor) -> torch.Tensor:
    if dm.ndim == 3 and dm.shape[0] == 2:
        return dm[0] + dm[1]
    return dm


def spin_orbital_rdm1(trial_rdm1: torch.Tensor) -> torch.Tensor:
    """A generalized basis takes the trial's one-body density matrix in the
    spin-orbital basis (nso, nso); a per-spin stack (2, n, n) (the RHF / UHF
    trials' get_rdm1) becomes its block diagonal diag(dm_up, dm_dn)."""
    if trial_rdm1.ndim == 3 and trial_rdm1.shape[0] == 2:
        n = trial_rdm1.shape[1]
        out = torch.zeros((2 * n, 2 * n), dtype=trial_rdm1.dtype, device=trial_rdm1.device)
        out[:n, :n] = trial_rdm1[0]
        out[n:, n:] = trial_rdm1[1]
        return out
    return trial_rdm1


def _mf_shifts(ham_data: ham_chol, trial_rdm1: torch.Tensor) -> torch.Tensor:
    """prop/chol_afqmc_ops.py:59-63: 1j * tr(L_g dm), dm the total density
    (restricted basis) or the spin-orbital one (generalized basis)."""
    dm = torch.as_tensor(trial_rdm1, device=ham_data.device)
    dm = dm.to(torch.complex128 if dm.is_complex() else torch.float64)
    if ham_data.basis == "restricted":
        dm = _as_total_rdm1_restricted(dm)
    else:
        dm = spin_orbital_rdm1(dm)
    return 1.0j * torch.einsum("gij,ji->g", ham_data.chol.to(dm.dtype), dm).to(torch.complex128)


def _build_prop_ctx(ham_data: ham_chol, trial_rdm1, dt: float) -> chol_afqmc_ctx:
    """prop/chol_afqmc_ops.py:77-103: in the restricted basis
    h1_eff = h1 - 1/2 sum_g L_g L_g - sum_g Re(i mf_g) L_g; a generalized
    basis keeps h1_eff = h1 (prop/chol_afqmc_ops.py:86-94)."""
    mf = _mf_shifts(ham_data, trial_rdm1)
    h0_prop = complex(-ham_data.h0 - 0.5 * torch.sum(mf ** 2).item())
    chol = ham_data.chol
    g, n, _ = chol.shape
    h1_eff = ham_data.h1
    if ham_data.basis == "restricted":
        v0m = 0.5 * torch.einsum("gik,gkj->ij", chol, chol)
        mf_r = (1.0j * mf).real
        v1m = torch.einsum("g,gik->ik", mf_r, chol)
        h1_eff = ham_data.h1 - v0m - v1m
    exp_h1_half = torch.linalg.matrix_exp(-0.5 * dt * h1_eff)
    # the exponential of a symmetric matrix, made exactly symmetric (the
    # rounding of the matrix products leaves ~1e-16 asymmetry), so that every
    # layout that stores one triangle holds the same matrix
    exp_h1_half = 0.5 * (exp_h1_half + exp_h1_half.T)
    ldv = kernels.ldv_for(n)
    pad = torch.zeros((n, ldv), dtype=torch.float64, device=chol.device)
    pad[:, :n] = exp_h1_half
    exp_h1 = exp_h1_half @ exp_h1_half
    exp_h1 = 0.5 * (exp_h1 + exp_h1.T)
    return chol_afqmc_ctx(dt=float(dt), sqrt_dt=float(dt) ** 0.5, exp_h1_half=exp_h1_half,
                          mf_shifts=mf.contiguous(), h0_prop=h0_prop, exp_h1_half_pad=pad,
                          exp_h1_half_pad32=pad.to(torch.float32),
                          exp_h1_half_sym_pairs=minus_identity_pairs(upper_blocks(exp_h1_half).to(torch.float32), n),
                          exp_h1_sym_pairs=minus_identity_pairs(upper_blocks(exp_h1).to(torch.float32), n))


def chol_transposed(chol: torch.Tensor) -> torch.Tensor:
    """L^T (n*ldv, n_fields) with zero rows at q >= n: the B operand of the VHS GEMM."""
    g, n, _ = chol.shape
    ldv = kernels.ldv_for(n)
    pad = torch.zeros((g, n, ldv), dtype=torch.float64, device=chol.device)
    pad[:, :, :n] = chol
    return pad.reshape(g, n * ldv).T.contiguous()


def chol_sym_rows(chol: torch.Tensor) -> torch.Tensor:
    """The rows of L^T for the upper 16 x 16 blocks of the (p, q) plane only:
    (sym_blocks(n) * 256, n_fields), block (bi, bj), bi <= bj, in row-major
    order, element (pi, qi) at pi * 16 + qi, zero rows past n.  The B operand
    of the symmetric VHS contraction (V = sum_g x_g L_g is symmetric)."""
    g, n, _ = chol.shape
    nb = (n + 15) // 16
    bi, bj = torch.triu_indices(nb, nb, device=chol.device)
    ar = torch.arange(16, device=chol.device)
    p = (bi[:, None, None] * 16 + ar[None, :, None]).expand(-1, 16, 16).reshape(-1)
    q = (bj[:, None, None] * 16 + ar[None, None, :]).expand(-1, 16, 16).reshape(-1)
    out = torch.zeros((p.numel(), g), dtype=torch.float64, device=chol.device)
    ok = (p < n) & (q < n)
    out[ok] = chol[:, p[ok], q[ok]].T
    return out


def upper_blocks(mat: torch.Tensor) -> torch.Tensor:
    """The upper 16 x 16 blocks of a symmetric n x n matrix, flat
    (sym_blocks(n) * 256,) in the order of chol_sym_rows, zero past n."""
    return chol_sym_rows(mat[None]).reshape(-1)


def split_pairs(x32: torch.Tensor) -> torch.Tensor:
    """float32 values as split-bf16 pair words (int32): bf16(x) in the low 16
    bits, bf16(x - bf16(x)) in the high 16 bits, both rounded to nearest even
    (the split the bf16x3 kernels make of an fp32 operand)."""
    hi = x32.to(torch.bfloat16)
    lo = (x32 - hi.to(torch.float32)).to(torch.bfloat16)
    return ((hi.view(torch.int16).to(torch.int32) & 0xFFFF) | (lo.view(torch.int16).to(torch.int32) << 16)).contiguous()


def minus_identity_pairs(e_blocks32: torch.Tensor, n: int) -> torch.Tensor:
    """Split pairs of E - I from E's float32 upper blocks: the identity is
    subtracted in float32 on the diagonal entries of the first n rows, as the
    full-row bf16x3 Trotter does on load."""
    eye = upper_blocks(torch.eye(n, dtype=torch.float64, device=e_blocks32.device)).to(torch.float32)
    return split_pairs(e_blocks32 - eye)


def _chunks(nw: int, n_chunks: int):
    if nw % n_chunks != 0:
        raise ValueError(f"n_walkers={nw} is not divisible into n_chunks={n_chunks}")
    cs = nw // n_chunks
    return [(i * cs, (i + 1) * cs) for i in range(n_chunks)]


PRECISIONS = ("fp64", "fp32", "bf16x3")


def make_chol_afqmc_ops(ham_data: ham_chol, walker_kind: str,
                        precision: str = "fp64") -> chol_afqmc_ops:
    """prop/chol_afqmc_ops.py:215-262: a restricted basis with restricted,
    unrestricted or generalized walkers, or a generalized (spin-orbital) basis
    with generalized walkers."""
    if precision not in PRECISIONS:
        raise ValueError(f"unknown precision {precision!r}; expected one of {PRECISIONS}")
    walker_kind = walker_kind.lower()
    nf = int(ham_data.chol.shape[0])
    n = int(ham_data.chol.shape[1])
    if ham_data.basis == "generalized" and walker_kind != "generalized":
        raise ValueError("ham_data.basis='generalized' requires walker_kind='generalized'")
    if walker_kind not in ("restricted", "unrestricted", "generalized"):
        raise ValueError(f"unknown walker_kind: {walker_kind}")
    # generalized walkers in the restricted basis: both spin blocks of the
    # (2n, o') walker with the n x n VHS; in a generalized basis the (nso, o')
    # walker is propagated whole with the nso x nso VHS, as a restricted walker
    # of nso orbitals (prop/chol_afqmc_ops.py:258-262)
    generalized = walker_kind == "generalized" and ham_data.basis == "restricted"

    # the VHS contraction against the upper blocks only, each entry stored twice
    chol_s = chol_sym_rows(ham_data.chol)
    ldv = kernels.ldv_for(n)
    if precision == "fp32":
        chol_s32 = kernels.convert_f32(chol_s)
        del chol_s
    elif precision == "bf16x3":
        chol_hi, chol_lo = kernels.split_bf16(chol_s)
        del chol_s

    def n_fields_() -> int:
        return nf

    # bf16x3 (n <= 224): the Trotter reads V as the upper blocks themselves,
    # so the contraction's output is used as it comes (0.58 of the bytes)
    v_blocks = precision == "bf16x3" and n <= kernels.BF16X3_SYM_MAX_N
    # the upper-block Trotter takes its two one-body halves separately
    merges = v_blocks and not generalized

    def make_vhs(xbar: torch.Tensor, xbar_split=None) -> torch.Tensor:
        """VHS planes in the step's precision, from the symmetric contraction
        (upper 16 x 16 blocks): (2, nw, sym_blocks(n) * 256) upper blocks for the
        bf16x3 Trotter, else (2, nw, n, ldv) rows (each entry stored twice)."""
        nwc = xbar.shape[1]
        flat = xbar.reshape(2 * nwc, nf)
        if v_blocks:
            # V leaves the contraction as split-bf16 pair words
            out = torch.empty((2 * nwc, chol_hi.shape[0]), dtype=torch.int32, device=xbar.device)
            xh, xl = xbar_split if xbar_split is not None else kernels.split_bf16(flat)
            kernels.gemm_bf16x3_nt(xh, xl, chol_hi, chol_lo, out, tag="vhs")
            return out.view(2, nwc, -1)
        if precision == "fp64":
            out = kernels.vhs_sym(flat, chol_s, n)
        elif precision == "fp32":
            out = kernels.vhs_sym(kernels.convert_f32(flat), chol_s32, n)
        else:
            out = kernels.vhs_sym(kernels.split_bf16(flat), (chol_hi, chol_lo), n)
        return out.view(2, nwc, n, ldv)

    def e_half(ctx: chol_afqmc_ctx) -> torch.Tensor:
        if v_blocks:
            return ctx.exp_h1_half_sym_pairs
        return ctx.exp_h1_half_pad if precision == "fp64" else ctx.exp_h1_half_pad32

    def build_prop_ctx(trial_rdm1, dt: float) -> chol_afqmc_ctx:
        return _build_prop_ctx(ham_data, trial_rdm1, dt)

    def apply_trotter_planes(w, xbar: torch.Tensor, ctx: chol_afqmc_ctx, n_terms: int,
                             n_chunks: int = 1, xbar_split=None, merged_first: bool = False,
                             last_half: bool = True):
        """``xbar_split``: optional (hi, lo) bf16 planes of xbar's rows (from
        kernels.field_shifts(split=True)), used by the bf16x3 VHS when n_chunks == 1.
        ``merged_first``: the first one-body factor is E^2 (the previous step
        left out its last half); ``last_half=False`` leaves out this step's last
        half (W_out = Taylor . E w).  Only where ``merges_halves``."""
        if (merged_first or not last_half) and not merges:
            raise ValueError("this propagator applies both one-body halves in every step")
        halves = (ctx.exp_h1_sym_pairs if merged_first else None, last_half)
        if generalized:
            # _apply_trotter_g_from_restricted (prop/chol_afqmc_ops.py:200-212): the
            # spin blocks of each (2n, o') walker with the walker's one VHS
            top, bot = apply_spin_blocks((w[:, :n, :].contiguous(), w[:, n:, :].contiguous()), xbar,
                                         ctx, n_terms, n_chunks, xbar_split)
            return torch.cat([top, bot], dim=1)
        return apply_spin_blocks(w, xbar, ctx, n_terms, n_chunks, xbar_split, halves)

    def apply_spin_blocks(w, xbar: torch.Tensor, ctx: chol_afqmc_ctx, n_terms: int, n_chunks: int,
                          xbar_split=None, halves=(None, True)):
        unr = isinstance(w, tuple)
        w0 = w[0] if unr else w
        nw = w0.shape[0]
        e = e_half(ctx)
        e_first = e if halves[0] is None else halves[0]
        # the last half: E (the same operand: the plain launch), or none
        kw = {} if (halves[0] is None and halves[1]) else {"e_last": e if halves[1] else None}
        if n_chunks == 1:
            return kernels.trotter(w, make_vhs(xbar, xbar_split), e_first, ctx.sqrt_dt, n_terms,
                                   arith=precision, **kw)
        out = (torch.empty_like(w[0]), torch.empty_like(w[1])) if unr else torch.empty_like(w0)
        for a, b in _chunks(nw, n_chunks):
            v = make_vhs(xbar[:, a:b].contiguous())
            wc = (w[0][a:b], w[1][a:b]) if unr else w0[a:b]
            oc = (out[0][a:b], out[1][a:b]) if unr else out[a:b]
            kernels.trotter(wc, v, e_first, ctx.sqrt_dt, n_terms, out=oc, arith=precision, **kw)
        return out

    def apply_trotter(w, fields, ctx: chol_afqmc_ctx, n_terms: int, n_chunks: int = 1):
        """fields (nw, n_fields), real or complex."""
        f = torch.as_tensor(fields, device=ham_data.device)
        if f.is_complex():
            xbar = torch.stack([f.real, f.imag]).to(torch.float64).contiguous()
        else:
            xbar = torch.stack([f, torch.zeros_like(f)]).to(torch.float64).contiguous()
        return apply_trotter_planes(w, xbar, ctx, n_terms, n_chunks)

    return chol_afqmc_ops(n_fields_, build_prop_ctx, apply_trotter, apply_trotter_planes, precision, v_blocks,
                          merges)
