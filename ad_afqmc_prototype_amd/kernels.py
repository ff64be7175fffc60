"""Typed torch-tensor wrappers around the C ABI of libafqmc_hip.so.

Every function checks device, dtype, contiguity and shape on the host, then
enqueues the kernel on the current HIP stream of the tensors' device.  There is
no fallback: a tensor that is not on a HIP device raises.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib

f64 = torch.float64
f32 = torch.float32
c128 = torch.complex128


def _stream(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


# Optional per-entry-point timing with HIP events recorded on the launch stream
# (used by the benchmark; off by default).
_timing = None


class timing:
    """Context manager: records (start, end) HIP events around every library
    call, on the stream the kernel is enqueued on.

    ``pool`` event pairs are created (and recorded once, then synchronized)
    before the context is entered: a freshly created HIP event costs ~10x a
    launch on its first record, which would otherwise land inside the timed
    region.  Calls beyond the pool create events on the fly."""

    def __init__(self, pool: int = 0):
        self.events = {}
        self._pool = []
        if pool > 0:
            st = torch.cuda.current_stream()
            for _ in range(pool):
                pair = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                pair[0].record(st)
                pair[1].record(st)
                self._pool.append(pair)
            torch.cuda.synchronize()
            self._pool.reverse()

    def _pair(self):
        if self._pool:
            return self._pool.pop()
        return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def __enter__(self):
        global _timing
        _timing = self
        return self

    def __exit__(self, *exc):
        global _timing
        _timing = None
        return False

    def summary(self):
        torch.cuda.synchronize()
        out = {}
        for name, evs in self.events.items():
            ms = [a.elapsed_time(b) for a, b in evs]
            out[name] = {"count": len(ms), "avg_ms": sum(ms) / len(ms), "total_ms": sum(ms)}
        return out


def _call(name, stream_tensor, *args, tag=None):
    if _timing is None:
        _lib.call(name, *args)
        return
    st = torch.cuda.current_stream(stream_tensor.device)
    e0, e1 = _timing._pair()
    e0.record(st)
    _lib.call(name, *args)
    e1.record(st)
    _timing.events.setdefault(name if tag is None else f"{name}:{tag}", []).append((e0, e1))


def _ptr(t):
    return None if t is None else t.data_ptr()


def _check(t: torch.Tensor, dtype, shape, name: str, device=None):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch.Tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{name}: tensor must live on a HIP device (got {t.device}); "
                         "the HIP path has no CPU fallback")
    if device is not None and t.device != device:
        raise ValueError(f"{name}: tensor on {t.device}, expected {device}")
    if t.dtype != dtype:
        raise TypeError(f"{name}: dtype {t.dtype}, expected {dtype}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: tensor must be contiguous")


def ldv_for(n: int) -> int:
    """Row pitch of VHS / exp_h1_half / rotated Cholesky rows: a multiple of 16
    entries, so fp32 rows start on 64-byte boundaries and the 64-byte row
    segments the VHS contraction stores are whole memory sectors."""
    return (n + 15) // 16 * 16


def dgemm_nt(A, B, C, alpha=1.0, beta=0.0, tag=None):
    """C = alpha * A @ B.T + beta * C (fp64 MFMA)."""
    M, K = A.shape
    N = B.shape[0]
    _check(A, f64, (M, K), "A")
    _check(B, f64, (N, K), "B", A.device)
    _check(C, f64, (M, N), "C", A.device)
    _call("afq_dgemm_nt", A, M, N, K, float(alpha), A.data_ptr(), K, B.data_ptr(), K, float(beta),
              C.data_ptr(), N, _stream(A), tag=tag)
    return C


def gemm_nt_lower(A, B, C, alpha=1.0, tag=None):
    """C = alpha * A @ B.T for a B whose row j is zero past column j (every
    column tile's k-loop stops at its last column; same result as the full
    GEMM on the same B).  A, B float64 (fp64, C float64), float32 (fp32, C
    float32 or float64) or (hi, lo) bf16 plane pairs from split_bf16 (C
    float32 or float64)."""
    if isinstance(A, tuple):
        (ah, al), (bh, bl) = A, B
        M, K = ah.shape
        N = bh.shape[0]
        for t, nm, sh in ((ah, "a_hi", (M, K)), (al, "a_lo", (M, K)), (bh, "b_hi", (N, K)), (bl, "b_lo", (N, K))):
            _check(t, torch.int16, sh, nm, ah.device)
        if C.dtype not in (f32, f64):
            raise TypeError("C must be float32 or float64")
        _check(C, C.dtype, (M, N), "C", ah.device)
        _call("afq_gemm_bf16x3_nt_lower", ah, M, N, K, float(alpha), ah.data_ptr(), al.data_ptr(), K,
              bh.data_ptr(), bl.data_ptr(), K, C.data_ptr(), N, int(C.dtype == f64), _stream(ah), tag=tag)
        return C
    M, K = A.shape
    N = B.shape[0]
    if A.dtype == f64:
        _check(A, f64, (M, K), "A")
        _check(B, f64, (N, K), "B", A.device)
        _check(C, f64, (M, N), "C", A.device)
        _call("afq_dgemm_nt_lower", A, M, N, K, float(alpha), A.data_ptr(), K, B.data_ptr(), K, C.data_ptr(), N,
              _stream(A), tag=tag)
        return C
    _check(A, f32, (M, K), "A")
    _check(B, f32, (N, K), "B", A.device)
    if C.dtype not in (f32, f64):
        raise TypeError("C must be float32 or float64")
    _check(C, C.dtype, (M, N), "C", A.device)
    _call("afq_sgemm_nt_lower", A, M, N, K, float(alpha), A.data_ptr(), K, B.data_ptr(), K, C.data_ptr(), N,
          int(C.dtype == f64), _stream(A), tag=tag)
    return C


def sym_blocks(n: int) -> int:
    """Number of upper 16 x 16 blocks of an n x n symmetric matrix."""
    nb = (n + 15) // 16
    return nb * (nb + 1) // 2


def vhs_sym(a, b, n, out=None, tag="vhs"):
    """Symmetric VHS contraction: full VHS rows (rows, n, ldv) from shifted-field
    rows ``a`` (rows, k) and the upper-block Cholesky rows ``b``
    (sym_blocks(n) * 256, k).  ``a``, ``b`` float64 (fp64, float64 out),
    float32 (fp32, float32 out) or (hi, lo) bf16 plane pairs from split_bf16
    (bf16x3, float32 out)."""
    if isinstance(a, tuple):
        (ah, al), (bh, bl) = a, b
        arith, out_dtype = 2, f32
        rows, k = ah.shape
        for t, nm, sh in ((ah, "a_hi", (rows, k)), (al, "a_lo", (rows, k)),
                          (bh, "b_hi", (sym_blocks(n) * 256, k)), (bl, "b_lo", (sym_blocks(n) * 256, k))):
            _check(t, torch.int16, sh, nm, ah.device)
    else:
        ah, al, bh, bl = a, None, b, None
        if a.dtype not in (f64, f32):
            raise TypeError("vhs_sym: float64, float32 or bf16 plane pairs")
        arith, out_dtype = (0, f64) if a.dtype == f64 else (1, f32)
        rows, k = a.shape
        _check(a, a.dtype, (rows, k), "a")
        _check(b, a.dtype, (sym_blocks(n) * 256, k), "b", a.device)
    ldv = ldv_for(n)
    if out is None:
        out = torch.empty((rows, n, ldv), dtype=out_dtype, device=ah.device)
    _check(out, out_dtype, (rows, n, ldv), "out", ah.device)
    _call("afq_vhs_sym", ah, rows, n, ldv, k, arith, ah.data_ptr(), _ptr(al), k, bh.data_ptr(), _ptr(bl), k,
          out.data_ptr(), _stream(ah), tag=tag)
    return out


def sgemm_nt(A, B, C, alpha=1.0, beta=0.0, tag=None):
    """C = alpha * A @ B.T + beta * C, fp32 operands, C float32 or float64."""
    M, K = A.shape
    N = B.shape[0]
    _check(A, f32, (M, K), "A")
    _check(B, f32, (N, K), "B", A.device)
    if C.dtype not in (f32, f64):
        raise TypeError("C must be float32 or float64")
    _check(C, C.dtype, (M, N), "C", A.device)
    _call("afq_sgemm_nt", A, M, N, K, float(alpha), A.data_ptr(), K, B.data_ptr(), K, float(beta),
          C.data_ptr(), N, int(C.dtype == f64), _stream(A), tag=tag)
    return C


def pad8(k: int) -> int:
    return (k + 7) // 8 * 8


def split_bf16(x: torch.Tensor, ld: int | None = None):
    """float64 (rows, cols) -> (hi, lo) bf16 bit planes (uint16 as int16) of
    shape (rows, ld), ld = cols rounded up to 8 by default (zero padding)."""
    if x.dim() != 2:
        x = x.reshape(-1, x.shape[-1])
    rows, cols = x.shape
    _check(x, f64, (rows, cols), "x")
    ld = pad8(cols) if ld is None else ld
    hi = torch.empty((rows, ld), dtype=torch.int16, device=x.device)
    lo = torch.empty((rows, ld), dtype=torch.int16, device=x.device)
    _call("afq_split_bf16", x, rows, cols, x.data_ptr(), cols, hi.data_ptr(), lo.data_ptr(), ld,
          _stream(x))
    return hi, lo


def convert_f32(x: torch.Tensor, ld: int | None = None):
    """float64 (rows, cols) -> float32 (rows, ld) with zero padding."""
    if x.dim() != 2:
        x = x.reshape(-1, x.shape[-1])
    rows, cols = x.shape
    _check(x, f64, (rows, cols), "x")
    ld = cols if ld is None else ld
    out = torch.empty((rows, ld), dtype=f32, device=x.device)
    _call("afq_convert_f32", x, rows, cols, x.data_ptr(), cols, out.data_ptr(), ld, _stream(x))
    return out


def splitk_for(M: int, N: int, K: int) -> int:
    """K chunks for the bf16x3 GEMM: problems that the 256 x 256 tiles fill run
    whole; otherwise enough chunks (at most 4, each >= 512 deep) to give the
    128 x 128 tile grid about four workgroups per CU (256 CUs)."""
    if ((M + 255) // 256) * ((N + 255) // 256) >= 512:
        return 1
    tiles = ((M + 127) // 128) * ((N + 127) // 128)
    return max(1, min(4, 1024 // tiles, K // 512))


def gemm_bf16x3_nt(a_hi, a_lo, b_hi, b_lo, C, alpha=1.0, tag=None, splits=None):
    """C (M, N) = alpha * (Ahi Bhi^T + Ahi Blo^T + Alo Bhi^T); planes from split_bf16.
    ``splits`` > 1 cuts K into that many chunks run side by side (fp32
    partials in a workspace, summed in chunk order in fp64); None picks it
    from the shape (splitk_for)."""
    M, K = a_hi.shape
    N = b_hi.shape[0]
    if splits is None:
        splits = splitk_for(M, N, K)
    for t, nm, sh in ((a_hi, "a_hi", (M, K)), (a_lo, "a_lo", (M, K)), (b_hi, "b_hi", (N, K)),
                      (b_lo, "b_lo", (N, K))):
        _check(t, torch.int16, sh, nm, a_hi.device)
    if C.dtype not in (f32, f64, torch.int32):
        raise TypeError("C must be float32, float64 or int32 (split-bf16 pair words)")
    _check(C, C.dtype, (M, N), "C", a_hi.device)
    mode = {f32: 0, f64: 1, torch.int32: 2}[C.dtype]
    if splits > 1:
        ws = torch.empty((splits, M, N), dtype=f32, device=a_hi.device)
        _call("afq_gemm_bf16x3_nt_splitk", a_hi, M, N, K, float(alpha), a_hi.data_ptr(), a_lo.data_ptr(), K,
              b_hi.data_ptr(), b_lo.data_ptr(), K, C.data_ptr(), N, mode, int(splits),
              ws.data_ptr(), _stream(a_hi), tag=tag)
        return C
    _call("afq_gemm_bf16x3_nt", a_hi, M, N, K, float(alpha), a_hi.data_ptr(), a_lo.data_ptr(), K,
          b_hi.data_ptr(), b_lo.data_ptr(), K, C.data_ptr(), N, mode, _stream(a_hi),
          tag=tag)
    return C


def overlap_greens(walkers, trial_up, trial_dn=None, greens=None, logdet=None, split=False):
    """Returns (greens planes (2, nw, nspin*o*n), logdet complex128 (nw, nspin)),
    and with ``split`` also the (hi, lo) bf16 planes of the greens rows
    (2 nw, pad8(nspin*o*n)) that split_bf16 would make, from the same pass."""
    wu, wd = (walkers if isinstance(walkers, tuple) else (walkers, None))
    nspin = 1 if wd is None else 2
    nw, n, o = wu.shape
    _check(wu, c128, (nw, n, o), "walkers")
    dev = wu.device
    if wd is not None:
        _check(wd, c128, (nw, n, o), "walkers_dn", dev)
    ctrial = trial_up.is_complex() or (trial_dn is not None and trial_dn.is_complex())
    tdt = c128 if ctrial else f64
    if ctrial:
        trial_up = trial_up.to(c128)
        trial_dn = None if trial_dn is None else trial_dn.to(c128)
    _check(trial_up, tdt, (n, o), "trial_up", dev)
    if nspin == 2:
        trial_dn = trial_up if trial_dn is None else trial_dn
        _check(trial_dn, tdt, (n, o), "trial_dn", dev)
    if ctrial:
        # complex trial orbitals: m = C^H w in the ctrial kernel, the bf16
        # planes (if asked) split from its fp64 rows
        if greens is None:
            greens = torch.empty((2, nw, nspin * o * n), dtype=f64, device=dev)
        if logdet is None:
            logdet = torch.empty((nw, nspin), dtype=c128, device=dev)
        _call("afq_overlap_greens_ctrial", wu, nw, n, o, nspin, wu.data_ptr(), _ptr(wd), trial_up.data_ptr(),
              _ptr(trial_dn if nspin == 2 else None), greens.data_ptr(), logdet.data_ptr(), _stream(wu))
        if split:
            return greens, logdet, split_bf16(greens.view(2 * nw, nspin * o * n))
        return greens, logdet
    if greens is None:
        greens = torch.empty((2, nw, nspin * o * n), dtype=f64, device=dev)
    if logdet is None:
        logdet = torch.empty((nw, nspin), dtype=c128, device=dev)
    _check(greens, f64, (2, nw, nspin * o * n), "greens", dev)
    _check(logdet, c128, (nw, nspin), "logdet", dev)
    if split:
        ldk = pad8(nspin * o * n)
        hi = torch.empty((2 * nw, ldk), dtype=torch.int16, device=dev)
        lo = torch.empty((2 * nw, ldk), dtype=torch.int16, device=dev)
        _call("afq_overlap_greens_split", wu, nw, n, o, nspin, wu.data_ptr(), _ptr(wd), trial_up.data_ptr(),
              _ptr(trial_dn if nspin == 2 else None), greens.data_ptr(), logdet.data_ptr(), hi.data_ptr(),
              lo.data_ptr(), ldk, _stream(wu))
        return greens, logdet, (hi, lo)
    _call("afq_overlap_greens", wu, nw, n, o, nspin, wu.data_ptr(), _ptr(wd), trial_up.data_ptr(),
              _ptr(trial_dn if nspin == 2 else None), greens.data_ptr(), logdet.data_ptr(),
              _stream(wu))
    return greens, logdet


def overlap_greens_bf16x3(walkers, trial_up, trial_dn=None):
    """log det (nw, nspin) complex128 and the (hi, lo) bf16 planes (2 nw,
    pad8(nspin*o*n)) of the Green's function rows, without the fp64 rows: the
    rows are a split-bf16 product of w and the fp64 inverse of C^T w (o <= 32).
    The operands of the bf16x3 force bias."""
    wu, wd = (walkers if isinstance(walkers, tuple) else (walkers, None))
    nspin = 1 if wd is None else 2
    nw, n, o = wu.shape
    _check(wu, c128, (nw, n, o), "walkers")
    dev = wu.device
    if wd is not None:
        _check(wd, c128, (nw, n, o), "walkers_dn", dev)
    _check(trial_up, f64, (n, o), "trial_up", dev)
    if nspin == 2:
        trial_dn = trial_up if trial_dn is None else trial_dn
        _check(trial_dn, f64, (n, o), "trial_dn", dev)
    logdet = torch.empty((nw, nspin), dtype=c128, device=dev)
    ldk = pad8(nspin * o * n)
    hi = torch.empty((2 * nw, ldk), dtype=torch.int16, device=dev)
    lo = torch.empty((2 * nw, ldk), dtype=torch.int16, device=dev)
    _call("afq_overlap_greens_bf16x3", wu, nw, n, o, nspin, wu.data_ptr(), _ptr(wd), trial_up.data_ptr(),
          _ptr(trial_dn if nspin == 2 else None), logdet.data_ptr(), hi.data_ptr(), lo.data_ptr(), ldk,
          _stream(wu))
    return logdet, (hi, lo)


def splitk_for_f64(M: int, N: int, K: int) -> int:
    """K chunks for the fp64 GEMM (128 x 128 tiles, two workgroups per CU):
    whole when the tile grid covers 256 CUs twice; otherwise enough chunks (at
    most 8, each >= 512 deep) for about four workgroups per CU."""
    tiles = ((M + 127) // 128) * ((N + 127) // 128)
    if tiles >= 512:
        return 1
    return max(1, min(8, 1024 // tiles, K // 512))


def force_bias(greens, rot, alpha, fb=None, splits=None):
    """fb planes (2, nw, naux) = alpha * greens_rows . rot^T.  ``splits`` > 1
    cuts the contraction (K = nspin*o*n) into that many chunks run side by side
    and summed in order; None picks it from the shape (splitk_for_f64)."""
    _, nw, k = greens.shape
    naux = rot.shape[0]
    _check(greens, f64, (2, nw, k), "greens")
    _check(rot, f64, (naux, k), "rot", greens.device)
    if fb is None:
        fb = torch.empty((2, nw, naux), dtype=f64, device=greens.device)
    _check(fb, f64, (2, nw, naux), "fb", greens.device)
    if splits is None:
        splits = splitk_for_f64(2 * nw, naux, k)
    if splits > 1:
        ws = torch.empty((splits, 2 * nw, naux), dtype=f64, device=greens.device)
        _call("afq_dgemm_nt_splitk", greens, 2 * nw, naux, k, float(alpha), greens.data_ptr(), k, rot.data_ptr(), k,
              fb.data_ptr(), naux, int(splits), ws.data_ptr(), _stream(greens), tag="force_bias")
        return fb
    _call("afq_force_bias", greens, nw, naux, k, float(alpha), greens.data_ptr(), rot.data_ptr(),
              fb.data_ptr(), _stream(greens))
    return fb


def field_shifts(fields, fb, mf, sqrt_dt, xbar=None, shift_term=None, fb_term=None, split=False):
    """(xbar planes (2, nw, naux), shift_term, fb_term); with ``split`` also the
    (hi, lo) bf16 planes of the xbar rows (2 nw, pad8(naux)) that split_bf16
    would make, from the same pass."""
    nw, naux = fields.shape
    dev = fields.device
    _check(fields, f64, (nw, naux), "fields")
    _check(fb, f64, (2, nw, naux), "fb", dev)
    _check(mf, c128, (naux,), "mf_shifts", dev)
    if xbar is None:
        xbar = torch.empty((2, nw, naux), dtype=f64, device=dev)
    if shift_term is None:
        shift_term = torch.empty((nw,), dtype=c128, device=dev)
    if fb_term is None:
        fb_term = torch.empty((nw,), dtype=c128, device=dev)
    if split:
        ldk = pad8(naux)
        hi = torch.empty((2 * nw, ldk), dtype=torch.int16, device=dev)
        lo = torch.empty((2 * nw, ldk), dtype=torch.int16, device=dev)
        _call("afq_field_shifts_split", fields, nw, naux, float(sqrt_dt), fields.data_ptr(), fb.data_ptr(),
              mf.data_ptr(), xbar.data_ptr(), shift_term.data_ptr(), fb_term.data_ptr(), hi.data_ptr(),
              lo.data_ptr(), ldk, _stream(fields))
        return xbar, shift_term, fb_term, (hi, lo)
    _call("afq_field_shifts", fields, nw, naux, float(sqrt_dt), fields.data_ptr(), fb.data_ptr(),
              mf.data_ptr(), xbar.data_ptr(), shift_term.data_ptr(), fb_term.data_ptr(),
              _stream(fields))
    return xbar, shift_term, fb_term


def vhs(xbar, chol_t, n, vhs_out=None):
    """VHS planes (2, nw, n, ldv) from shifted-field planes (2, nw, naux)."""
    _, nw, naux = xbar.shape
    ldv = ldv_for(n)
    _check(xbar, f64, (2, nw, naux), "xbar")
    _check(chol_t, f64, (n * ldv, naux), "chol_t", xbar.device)
    if vhs_out is None:
        vhs_out = torch.empty((2, nw, n, ldv), dtype=f64, device=xbar.device)
    _check(vhs_out, f64, (2, nw, n, ldv), "vhs", xbar.device)
    _call("afq_vhs", xbar, nw, n, ldv, naux, xbar.data_ptr(), chol_t.data_ptr(), vhs_out.data_ptr(),
              _stream(xbar))
    return vhs_out


BF16X3_SYM_MAX_N = 224  # on-chip capacity of the bf16x3 Trotter


_SAME = object()


def trotter(walkers, vhs_planes, exp_h1_half_pad, sqrt_dt, n_terms, out=None, arith=None, e_last=_SAME):
    """Fused Trotter step; fp64 when vhs / exp_h1_half are float64.  With
    float32 operands ``arith`` picks the products: "fp32" (default) or
    "bf16x3" (split-bf16 MFMAs, fp32 sums).  Walkers are complex128 either way.
    bf16x3 also takes V and E - I as upper 16 x 16 blocks of split-bf16 pair
    words: vhs_planes (2, nw, sym_blocks(n) * 256) and exp_h1_half_pad
    (sym_blocks(n) * 256,) int32 (n <= 224).  In that layout ``e_last`` may
    name another pair-word matrix F - I for the last one-body half (the first
    uses exp_h1_half_pad), or None to leave the last half out
    (afq_trotter_bf16x3_sym_halves); launches without a last half are timed
    under the tag "deferred"."""
    wu, wd = (walkers if isinstance(walkers, tuple) else (walkers, None))
    nspin = 1 if wd is None else 2
    nw, n, o = wu.shape
    ldv = ldv_for(n)
    dev = wu.device
    _check(wu, c128, (nw, n, o), "walkers")
    if wd is not None:
        _check(wd, c128, (nw, n, o), "walkers_dn", dev)
    rt = vhs_planes.dtype
    if vhs_planes.dim() == 3:
        if arith != "bf16x3" or rt != torch.int32:
            raise ValueError("upper-block VHS planes are a bf16x3 operand (int32 split-bf16 pairs)")
        if n > BF16X3_SYM_MAX_N:
            raise ValueError(f"upper-block VHS planes need n <= {BF16X3_SYM_MAX_N}, got n={n}")
        _check(vhs_planes, torch.int32, (2, nw, sym_blocks(n) * 256), "vhs", dev)
        _check(exp_h1_half_pad, torch.int32, (sym_blocks(n) * 256,), "exp_h1_half - I (upper-block pairs)", dev)
        if out is None:
            out = (torch.empty_like(wu), torch.empty_like(wd)) if nspin == 2 else torch.empty_like(wu)
        ou, od = (out if isinstance(out, tuple) else (out, None))
        if e_last is _SAME or e_last is exp_h1_half_pad:
            _call("afq_trotter_bf16x3_sym", wu, nw, n, o, nspin, int(n_terms), float(sqrt_dt),
                  exp_h1_half_pad.data_ptr(), vhs_planes.data_ptr(), wu.data_ptr(), _ptr(wd), ou.data_ptr(),
                  _ptr(od), _stream(wu))
            return out
        if e_last is not None:
            _check(e_last, torch.int32, (sym_blocks(n) * 256,), "e_last (upper-block pairs)", dev)
        _call("afq_trotter_bf16x3_sym_halves", wu, nw, n, o, nspin, int(n_terms), float(sqrt_dt),
              exp_h1_half_pad.data_ptr(), _ptr(e_last), vhs_planes.data_ptr(), wu.data_ptr(), _ptr(wd),
              ou.data_ptr(), _ptr(od), _stream(wu), tag="deferred" if e_last is None else None)
        return out
    if e_last is not _SAME:
        raise ValueError("separate one-body halves need the upper-block pair-word (bf16x3) operands")
    if rt not in (f64, f32):
        raise TypeError("vhs must be float64 or float32")
    if arith is None:
        arith = "fp64" if rt == f64 else "fp32"
    if (rt == f64) != (arith == "fp64") or arith not in ("fp64", "fp32", "bf16x3"):
        raise ValueError(f"arith={arith!r} does not match {rt} operands")
    _check(vhs_planes, rt, (2, nw, n, ldv), "vhs", dev)
    _check(exp_h1_half_pad, rt, (n, ldv), "exp_h1_half", dev)
    if out is None:
        out = (torch.empty_like(wu), torch.empty_like(wd)) if nspin == 2 else torch.empty_like(wu)
    ou, od = (out if isinstance(out, tuple) else (out, None))
    entry = {"fp64": "afq_trotter", "fp32": "afq_trotter_f32", "bf16x3": "afq_trotter_bf16x3"}[arith]
    _call(entry, wu, nw, n, o, nspin, ldv,
          int(n_terms), float(sqrt_dt),
              exp_h1_half_pad.data_ptr(), vhs_planes.data_ptr(), wu.data_ptr(), _ptr(wd),
              ou.data_ptr(), _ptr(od), _stream(wu))
    return out


def phaseless_update(logdet_new, logovlp_old, shift_term, fb_term, weights, pop_shift, *,
                     e_estimate, h0_prop, dt, sqrt_dt, weight_floor, weight_cap, damping):
    nw, nspin = logdet_new.shape
    dev = logdet_new.device
    _check(logdet_new, c128, (nw, nspin), "logdet_new")
    _check(logovlp_old, c128, (nw,), "log_overlaps", dev)
    _check(shift_term, c128, (nw,), "shift_term", dev)
    _check(fb_term, c128, (nw,), "fb_term", dev)
    _check(weights, f64, (nw,), "weights", dev)
    _check(pop_shift, f64, (), "pop_control_ene_shift", dev)
    lo_new = torch.empty((nw,), dtype=c128, device=dev)
    w_new = torch.empty((nw,), dtype=f64, device=dev)
    pop_new = torch.empty((), dtype=f64, device=dev)
    h0 = complex(h0_prop)
    _call("afq_phaseless_update", logdet_new, nw, nspin, logdet_new.data_ptr(), logovlp_old.data_ptr(),
              shift_term.data_ptr(), fb_term.data_ptr(), weights.data_ptr(), pop_shift.data_ptr(),
              float(e_estimate), h0.real, h0.imag, float(dt), float(sqrt_dt), float(weight_floor),
              float(weight_cap), float(damping), lo_new.data_ptr(), w_new.data_ptr(),
              pop_new.data_ptr(), _stream(logdet_new))
    return lo_new, w_new, pop_new


def orthonormalize(w, out=None):
    nw, n, o = w.shape
    _check(w, c128, (nw, n, o), "walkers")
    if out is None:
        out = torch.empty_like(w)
    _call("afq_orthonormalize", w, nw, n, o, w.data_ptr(), out.data_ptr(), _stream(w))
    return out


def energy_assemble(greens, z, c, rot_h1, nw, o, n, nspin, h0, c_scale=1.0, out=None):
    """Local energies (nw,) complex128 with the exchange as a quadratic form:
    greens (2, nw, nspin*o*n) rows, z (nspin, 2*nw, o*n) = each sector's rows
    times Y_s, c (2, nw, naux) planes (c_scale times the force-bias contraction
    summed over sectors), rot_h1 (nspin, o, n)."""
    dev = greens.device
    on = o * n
    _check(greens, f64, (2, nw, nspin * on), "greens")
    _check(z, f64, (nspin, 2 * nw, on), "z", dev)
    if c.dim() != 3 or c.shape[:2] != (2, nw):
        raise ValueError(f"c must be (2, {nw}, naux), got {tuple(c.shape)}")
    _check(c, f64, tuple(c.shape), "c", dev)
    _check(rot_h1, f64, (nspin, o, n), "rot_h1", dev)
    if out is None:
        out = torch.empty((nw,), dtype=c128, device=dev)
    _check(out, c128, (nw,), "energy", dev)
    _call("afq_energy_assemble", greens, nw, on, nspin, c.shape[2], float(h0), float(c_scale), greens.data_ptr(),
          z.data_ptr(), c.data_ptr(), rot_h1.data_ptr(), out.data_ptr(), _stream(greens))
    return out


def energy(greens, nw, n, o, nspin, rot_chol_pad, rot_h1, h0, out=None):
    """Local energies (nw,) complex128.  rot_chol_pad: (nspin, naux, opad,
    ldr) float64 or float32 (fp64 / fp32 contractions), or a (hi, lo) pair of
    bf16 planes (nspin * naux * opad, ldr32), ldr32 = n rounded up to 32
    (split-bf16 contractions); sums over g are fp64 in every mode."""
    dev = greens.device
    opad = (o + 15) // 16 * 16
    _check(greens, f64, (2, nw, nspin * o * n), "greens")
    _check(rot_h1, f64, (nspin, o, n), "rot_h1", dev)
    if out is None:
        out = torch.empty((nw,), dtype=c128, device=dev)
    _check(out, c128, (nw,), "energy", dev)
    if isinstance(rot_chol_pad, tuple):
        hi, lo = rot_chol_pad
        rows, ldr = hi.shape
        naux = rows // (nspin * opad)
        if naux * nspin * opad != rows or ldr % 32 or ldr < n:
            raise ValueError(f"bf16 energy operand {tuple(hi.shape)} does not match nspin={nspin}, "
                             f"o={o}, n={n}")
        _check(hi, torch.int16, (rows, ldr), "rot_chol_hi", dev)
        _check(lo, torch.int16, (rows, ldr), "rot_chol_lo", dev)
        _call("afq_energy_bf16x3", greens, nw, n, o, nspin, naux, float(h0), greens.data_ptr(),
              hi.data_ptr(), lo.data_ptr(), ldr, rot_h1.data_ptr(), out.data_ptr(), _stream(greens))
        return out
    naux = rot_chol_pad.shape[1]
    ldr = ldv_for(n)
    rt = rot_chol_pad.dtype
    if rt not in (f64, f32):
        raise TypeError("rot_chol_pad must be float64 or float32")
    _check(rot_chol_pad, rt, (nspin, naux, opad, ldr), "rot_chol_pad", dev)
    _call("afq_energy" if rt == f64 else "afq_energy_f32", greens, nw, n, o, nspin, naux, float(h0),
          greens.data_ptr(), rot_chol_pad.data_ptr(), ldr, rot_h1.data_ptr(), out.data_ptr(),
          _stream(greens))
    return out


def block_energy(energies, weights, pop_shift, *, e_ref, thresh, shift_ema):
    nw = energies.shape[0]
    dev = energies.device
    _check(energies, c128, (nw,), "energies")
    _check(weights, f64, (nw,), "weights", dev)
    _check(pop_shift, f64, (), "pop_control_ene_shift", dev)
    out = torch.empty((3,), dtype=f64, device=dev)
    _call("afq_block_energy", energies, nw, energies.data_ptr(), weights.data_ptr(), float(e_ref),
              float(thresh), float(shift_ema), pop_shift.data_ptr(), out.data_ptr(),
              _stream(energies))
    return out


def comb_indices(weights, zeta):
    """Comb indices and new weights; zeta a float, or a 0-d float64 device
    tensor (read on the device, no host round trip)."""
    nw = weights.shape[0]
    dev = weights.device
    _check(weights, f64, (nw,), "weights")
    cw = torch.empty((nw,), dtype=f64, device=dev)
    idx = torch.empty((nw,), dtype=torch.int32, device=dev)
    w_new = torch.empty((nw,), dtype=f64, device=dev)
    if isinstance(zeta, torch.Tensor):
        _check(zeta, f64, (), "zeta", dev)
        _call("afq_comb_indices_dev", weights, nw, weights.data_ptr(), zeta.data_ptr(), cw.data_ptr(),
              idx.data_ptr(), w_new.data_ptr(), _stream(weights))
    else:
        _call("afq_comb_indices", weights, nw, weights.data_ptr(), float(zeta), cw.data_ptr(),
              idx.data_ptr(), w_new.data_ptr(), _stream(weights))
    return idx, w_new


def rng_key(seed: int, device) -> torch.Tensor:
    """jax.random.PRNGKey(seed) as a device uint32[2] (held in an int32 tensor)."""
    s = int(seed) & 0xFFFFFFFFFFFFFFFF
    words = np.array([s >> 32, s & 0xFFFFFFFF], dtype=np.uint32).view(np.int32)
    return torch.tensor(words, dtype=torch.int32, device=device)


def rng_split(key):
    """(new key, subkey) = split(key), on the device."""
    _check(key, torch.int32, (2,), "key")
    nk = torch.empty_like(key)
    sk = torch.empty_like(key)
    _call("afq_rng_split", key, key.data_ptr(), nk.data_ptr(), sk.data_ptr(), _stream(key))
    return nk, sk


def rng_normal(key, shape, out=None):
    """normal(key, shape) float64 on the device."""
    _check(key, torch.int32, (2,), "key")
    if out is None:
        out = torch.empty(shape, dtype=f64, device=key.device)
    _check(out, f64, tuple(shape), "out", key.device)
    _call("afq_rng_normal", key, key.data_ptr(), out.numel(), out.data_ptr(), _stream(key))
    return out


def rng_uniform(key, out=None):
    """uniform(key) in [0, 1): a 0-d float64 device tensor."""
    _check(key, torch.int32, (2,), "key")
    if out is None:
        out = torch.empty((), dtype=f64, device=key.device)
    _check(out, f64, (), "out", key.device)
    _call("afq_rng_uniform", key, key.data_ptr(), out.data_ptr(), _stream(key))
    return out


def threefry2x32(k0: int, k1: int, c0, c1):
    """Threefry-2x32-20 block function on uint32 counter pairs (int32 tensors)."""
    n = c0.numel()
    _check(c0, torch.int32, (n,), "c0")
    _check(c1, torch.int32, (n,), "c1", c0.device)
    o0 = torch.empty_like(c0)
    o1 = torch.empty_like(c1)
    _call("afq_threefry2x32", c0, int(k0) & 0xFFFFFFFF, int(k1) & 0xFFFFFFFF, n, c0.data_ptr(),
          c1.data_ptr(), o0.data_ptr(), o1.data_ptr(), _stream(c0))
    return o0, o1


def gather_rows(src, idx, nplanes=1, out=None):
    """out[p, k] = src[p, idx[k]] over the leading (plane, row) axes; idx int32
    with every entry in [0, rows)."""
    if nplanes == 1:
        rows = src.shape[0]
        row_len = src[0].numel()
    else:
        rows = src.shape[1]
        row_len = src[0, 0].numel()
    if src.dtype == c128:
        row_len *= 2
    elif src.dtype != f64:
        raise TypeError("gather_rows: float64 or complex128 rows only")
    if src.device.type != "cuda" or not src.is_contiguous():
        raise ValueError("gather_rows: contiguous HIP tensor required")
    _check(idx, torch.int32, (idx.shape[0],), "idx", src.device)
    nrows = idx.shape[0]
    if out is None:
        shape = (nrows,) + tuple(src.shape[1:]) if nplanes == 1 else (nplanes, nrows) + tuple(src.shape[2:])
        out = torch.empty(shape, dtype=src.dtype, device=src.device)
    plane_src = rows * row_len
    plane_dst = nrows * row_len
    _call("afq_gather_rows", src, nrows, row_len, nplanes, plane_src, plane_dst, idx.data_ptr(),
              src.data_ptr(), out.data_ptr(), _stream(src))
    return out
