"""The exchange of the local energy as a quadratic form (host logic, CPU).

energy_kernel_r/u (meas/rhf.py:45-47) sums tr(f_g f_g) over the Cholesky
index with f_g = R_g G^T (o x o).  The measurement regroups that sum as
vec(G)^T Y vec(G) with Y built once per context from R^T R
(rhf_meas_ctx.exchange_operand), so the step runs one GEMM Z = G Y over all
walkers.  Y is checked here against the oracle's per-g sum on complex Green's
functions (where a transposed regroup or a conjugation would show), for odd
sizes and both spin sectors; its symmetry, and the cost rule that picks the
form."""
import numpy as np
import pytest
import torch

from ad_afqmc_prototype_amd.meas.rhf import quadratic_exchange, rhf_meas_ctx
from oracle import afqmc_oracle as orc


def _ctx(rot_chol):
    ns, g, o, n = rot_chol.shape
    rot = torch.tensor(rot_chol)
    return rhf_meas_ctx(rot_h1=torch.zeros((ns, o, n), dtype=torch.float64), rot_chol=rot,
                        rot_chol_flat=rot[0].reshape(g, o * n).contiguous())


@pytest.mark.parametrize("o,n,g", [(1, 1, 1), (3, 5, 7), (4, 9, 13), (6, 11, 40)])
def test_quadratic_form_equals_per_g_exchange(o, n, g):
    rng = np.random.default_rng(o * 100 + n)
    rot = rng.standard_normal((2, g, o, n))
    ctx = _ctx(rot)
    nw = 5
    gr = rng.standard_normal((nw, o, n)) + 1j * rng.standard_normal((nw, o, n))
    for s in range(2):
        y = ctx.exchange_operand(s).numpy()
        assert y.shape == (o * n, o * n)
        np.testing.assert_array_equal(y, y.T)
        v = gr.reshape(nw, o * n)
        quad = np.einsum("wi,ij,wj->w", v, y, v)
        _, exc = orc._coulomb_exchange(gr, rot[s])
        np.testing.assert_allclose(quad, exc, rtol=1e-12, atol=1e-12 * np.abs(exc).max())


def test_quadratic_form_transposed_regroup_fails():
    """A planted mistake: Y regrouped without the (b, p) <-> (b, q) swap is not
    the exchange (so the equality above is not vacuous)."""
    rng = np.random.default_rng(3)
    o, n, g = 3, 5, 7
    rot = rng.standard_normal((1, g, o, n))
    r = rot[0].reshape(g, o * n)
    wrong = (r.T @ r)            # Y[(a,p),(b,q)] instead of Y[(a,q),(b,p)]
    gr = rng.standard_normal((4, o, n)) + 1j * rng.standard_normal((4, o, n))
    v = gr.reshape(4, o * n)
    _, exc = orc._coulomb_exchange(gr, rot[0])
    assert np.abs(np.einsum("wi,ij,wj->w", v, wrong, v) - exc).max() > 1e-3 * np.abs(exc).max()


def test_quadratic_exchange_rule():
    assert quadratic_exchange(200, 32, 1000)
    assert quadratic_exchange(120, 20, 600)
    assert not quadratic_exchange(200, 32, 200)      # n >= naux: the per-g form is cheaper
    assert not quadratic_exchange(800, 64, 4000)     # Y of (o n)^2 fp64 beyond 4 GiB
