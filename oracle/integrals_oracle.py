"""CPU restatement (numpy) of the reference's Hamiltonian ingestion:
modified (pivoted, incomplete) Cholesky of the two-electron integrals
(integrals.py:3-46) and the pair packing its front end feeds it
(default.py:18-21, the 4-fold symmetric ERIs of ao2mo.restore(4, ...)).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the package.

Pinned (tests/test_integrals.py) by what the mathematics fixes: ERIs built
from known Cholesky vectors are decomposed into exactly their rank and
reconstructed; a diagonal matrix is decomposed pivot by pivot in descending
order; a full-rank SPD matrix is reconstructed exactly when the threshold is
zero, and to the threshold otherwise; the stopping rule's vector count.
"""
from __future__ import annotations

import numpy as np


def pair_index(m: int, p: int) -> int:
    a, b = (m, p) if m >= p else (p, m)
    return a * (a + 1) // 2 + b


def pack_eri(eri: np.ndarray) -> np.ndarray:
    """(n, n, n, n) ERIs (pq|rs) with the 8-fold symmetry -> (size, size) over
    pairs p >= q, r >= s (the layout of ao2mo.restore(4, ...))."""
    n = eri.shape[0]
    idx = [(m, p) for m in range(n) for p in range(m + 1)]
    size = len(idx)
    out = np.empty((size, size))
    for a, (m, p) in enumerate(idx):
        for b, (r, s) in enumerate(idx):
            out[a, b] = eri[m, p, r, s]
    return out


def eri_from_chol(chol: np.ndarray) -> np.ndarray:
    """(pq|rs) = sum_g L_g[p,q] L_g[r,s]."""
    return np.einsum("gpq,grs->pqrs", chol, chol)


def modified_cholesky(mat: np.ndarray, max_error: float = 1e-6, max_vectors: int | None = None):
    """integrals.py:3-46.  Returns (chol (nchol, n, n), pivots (nchol,), packed vectors)."""
    size = mat.shape[0]
    n = int(round((-1 + (1 + 8 * size) ** 0.5) / 2))
    limit = size if max_vectors is None else max_vectors
    diag = np.diagonal(mat).copy()
    vecs = np.zeros((limit, size))
    nu = int(np.argmax(diag))
    d = diag[nu]
    pivots = [nu]
    vecs[0] = mat[nu] / d ** 0.5
    approx = np.zeros(size)
    k = 0
    while abs(d) > max_error and k + 1 < limit:
        approx = approx + vecs[k] * vecs[k]
        resid = np.abs(diag - approx)
        nu = int(np.argmax(resid))
        d = resid[nu]
        r = vecs[: k + 1, nu] @ vecs[: k + 1]
        vecs[k + 1] = (mat[nu] - r) / (d + 1e-10) ** 0.5
        pivots.append(nu)
        k += 1
    v = vecs[:k]
    chol = np.zeros((k, n, n))
    for m in range(n):
        for p in range(m + 1):
            chol[:, m, p] = v[:, pair_index(m, p)]
            chol[:, p, m] = v[:, pair_index(m, p)]
    return chol, np.array(pivots[:k], dtype=np.int64), v
