"""CPU checks of the modified-Cholesky restatement (oracle/integrals_oracle.py,
the reference's integrals.py:3-46) and of the front end's host logic."""
import numpy as np
import pytest

from oracle import integrals_oracle as io


def _sym_chol(rng, g, n, scale=1.0):
    a = rng.standard_normal((g, n, n)) * scale
    return 0.5 * (a + a.transpose(0, 2, 1))


def test_pack_eri_layout():
    rng = np.random.default_rng(0)
    n = 4
    eri = io.eri_from_chol(_sym_chol(rng, 3, n))
    mat = io.pack_eri(eri)
    assert mat.shape == (10, 10)
    assert mat[io.pair_index(3, 1), io.pair_index(2, 2)] == eri[3, 1, 2, 2] == eri[1, 3, 2, 2]
    np.testing.assert_allclose(mat, mat.T)


@pytest.mark.parametrize("n,g", [(5, 4), (7, 12)])
def test_exact_rank_is_recovered(n, g):
    """ERIs built from g linearly independent symmetric vectors: exactly g vectors
    (once the residual reaches rounding level), and the ERIs are reconstructed."""
    rng = np.random.default_rng(n + g)
    chol_true = _sym_chol(rng, g, n)
    mat = io.pack_eri(io.eri_from_chol(chol_true))
    chol, piv, v = io.modified_cholesky(mat, max_error=1e-8)
    assert chol.shape[0] == g
    np.testing.assert_allclose(io.eri_from_chol(chol), io.eri_from_chol(chol_true), atol=1e-9)
    assert len(set(piv.tolist())) == g


def test_diagonal_matrix_pivots_descend():
    d = np.array([0.3, 2.0, 0.7, 1.1, 0.05, 0.9])
    mat = np.diag(d)
    chol, piv, v = io.modified_cholesky(mat, max_error=0.06)
    order = np.argsort(-d)
    # vectors are kept while the next pivot exceeds the threshold: 0.05 is not
    assert piv.tolist() == order[:5].tolist()
    for k, p in enumerate(piv):
        expect = np.zeros(6)
        expect[p] = d[p] / np.sqrt(d[p] + (0.0 if k == 0 else 1e-10))
        np.testing.assert_allclose(v[k], expect, rtol=1e-9, atol=1e-15)


def test_full_rank_spd_reconstructed():
    rng = np.random.default_rng(3)
    size = 15  # n = 5
    a = rng.standard_normal((size, size))
    mat = a @ a.T + size * np.eye(size)
    chol, piv, v = io.modified_cholesky(mat, max_error=0.0)
    # the stopping rule keeps at most size - 1 vectors
    assert v.shape[0] == size - 1
    # the reconstruction misses only the last pivot's rank-one part
    resid = mat - v.T @ v
    last = sorted(set(range(size)) - set(piv.tolist()))
    assert len(last) == 1
    off = np.delete(np.delete(resid, last, 0), last, 1)
    assert np.abs(off).max() < 1e-8 * np.abs(mat).max()


def test_threshold_bounds_the_residual():
    """The residual M - V V^T is the PSD Schur complement of the kept pivots, so
    every entry is at most its largest diagonal, which the stopping rule bounds
    by the threshold (rank 12 < 21 pairs: the vector cap is not what stops it)."""
    rng = np.random.default_rng(4)
    n, g = 6, 12
    chol_true = _sym_chol(rng, g, n) * np.exp(-0.6 * np.arange(g))[:, None, None]
    mat = io.pack_eri(io.eri_from_chol(chol_true))
    counts = []
    for tol in (1e-2, 1e-4, 1e-6):
        chol, piv, v = io.modified_cholesky(mat, max_error=tol)
        assert v.shape[0] < mat.shape[0] - 1
        resid = mat - v.T @ v
        assert np.abs(np.diagonal(resid)).max() <= tol
        assert np.abs(resid).max() <= tol * (1 + 1e-9)
        err = np.abs(io.eri_from_chol(chol) - io.eri_from_chol(chol_true)).max()
        assert err <= tol * (1 + 1e-9)
        counts.append(v.shape[0])
    assert counts == sorted(counts) and counts[-1] <= g


def test_max_vectors_limit():
    rng = np.random.default_rng(5)
    mat = io.pack_eri(io.eri_from_chol(_sym_chol(rng, 8, 5)))
    chol, piv, v = io.modified_cholesky(mat, max_error=1e-12, max_vectors=4)
    assert chol.shape[0] == 3  # the loop stops before a 4th: nchol + 1 < max_vectors


def test_product_refuses_host_tensors():
    import torch
    from ad_afqmc_prototype_amd.ham.integrals import modified_cholesky
    with pytest.raises(ValueError):
        modified_cholesky(torch.eye(3, dtype=torch.float64), device=torch.device("cpu"))


def test_front_end_needs_pyscf_for_mean_field_objects():
    import importlib.util
    from ad_afqmc_prototype_amd.default import Rhf
    if importlib.util.find_spec("pyscf") is not None:
        pytest.skip("pyscf is installed")
    with pytest.raises(ImportError):
        Rhf(object())
