"""Hamiltonian ingestion on the GPU: modified Cholesky decomposition of the
two-electron integrals into the Cholesky vectors of ham_chol
(integrals.py:3-46 of the reference, called from default.py:18-21).

The input is the reference's: the 4-fold symmetric ERIs (pq|rs) over pairs
p >= q, r >= s, a (size x size) matrix with size = n (n + 1) / 2 (what
ao2mo.restore(4, ...) returns).  The decomposition runs on the HIP kernels
of csrc/cholesky.hip: iterations are enqueued in chunks without host
synchronisation, and the host reads the device's stop flag between chunks.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _lib
from .. import kernels

f64 = torch.float64


def pack_eri(eri) -> torch.Tensor:
    """(n, n, n, n) ERIs with the 8-fold symmetry -> (size, size) over pairs
    p >= q, r >= s (the layout of ao2mo.restore(4, ...)), on eri's device."""
    eri = torch.as_tensor(eri, dtype=f64)
    n = eri.shape[0]
    rows, cols = torch.tril_indices(n, n, device=eri.device)
    sub = eri[rows, cols]                # (size, n, n)
    return sub[:, rows, cols].contiguous()


def modified_cholesky(mat, max_error: float = 1e-6, *, max_vectors: int | None = None,
                      chunk: int = 32, device=None) -> torch.Tensor:
    """Cholesky vectors (nchol, n, n) float64 on the device with
    mat ~= sum_g vec(L_g) vec(L_g)^T over packed pairs, by the reference's
    pivoted iteration: largest residual diagonal first, stopping once it is at
    most ``max_error`` (or at ``max_vectors``, default size, minus one)."""
    if device is None:
        device = mat.device if isinstance(mat, torch.Tensor) else torch.device("cuda", torch.cuda.current_device())
    m = torch.as_tensor(np.asarray(mat) if not isinstance(mat, torch.Tensor) else mat, dtype=f64)
    m = m.to(device).contiguous()
    if m.device.type != "cuda":
        raise ValueError("modified_cholesky runs on a HIP device only")
    size = m.shape[0]
    if m.dim() != 2 or m.shape[1] != size:
        raise ValueError(f"mat must be square, got {tuple(m.shape)}")
    n = int(round((-1 + (1 + 8 * size) ** 0.5) / 2))
    if n * (n + 1) // 2 != size:
        raise ValueError(f"mat size {size} is not n (n + 1) / 2")
    limit = size if max_vectors is None else int(max_vectors)
    if not 1 <= limit <= size:
        raise ValueError(f"max_vectors={limit} not in 1..{size}")
    r_len, p_len = ctypes.c_longlong(), ctypes.c_longlong()
    _lib.call("afq_mcd_workspace", size, ctypes.byref(r_len), ctypes.byref(p_len))
    vecs = torch.empty((limit, size), dtype=f64, device=device)
    resid = torch.empty((r_len.value,), dtype=f64, device=device)
    partial = torch.empty((p_len.value,), dtype=f64, device=device)
    state = torch.zeros((4,), dtype=f64, device=device)
    st = kernels._stream(m)
    args = (size, m.data_ptr(), size, float(max_error), limit)
    kernels._call("afq_mcd_start", m, *args, vecs.data_ptr(), resid.data_ptr(), partial.data_ptr(),
                  state.data_ptr(), st)
    done = bool(state[3].item())
    while not done:
        kernels._call("afq_mcd_iterate", m, *args, int(chunk), vecs.data_ptr(), resid.data_ptr(),
                      partial.data_ptr(), state.data_ptr(), st)
        done = bool(state[3].item())
    nchol = int(state[2].item())
    chol = torch.empty((nchol, n, n), dtype=f64, device=device)
    kernels._call("afq_unpack_pairs", m, nchol, n, vecs.data_ptr(), size, chol.data_ptr(), st)
    return chol
