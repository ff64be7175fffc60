"""The allocation-free comb (afq_comb_indices_noalloc, scratch in a static
device array of the library) against afq_comb_indices_dev (caller scratch),
through the C ABI: two calls in a row, the second with fewer walkers, so the
second reads the array the first left behind."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _comb_dev(lib, w, zeta):
    import torch
    nw = w.shape[0]
    cw = torch.empty(nw, dtype=torch.float64, device=w.device)
    idx = torch.full((nw,), -7, dtype=torch.int32, device=w.device)
    wn = torch.full((nw,), float("nan"), dtype=torch.float64, device=w.device)
    lib.call("afq_comb_indices_dev", nw, w.data_ptr(), zeta.data_ptr(), cw.data_ptr(), idx.data_ptr(),
             wn.data_ptr(), torch.cuda.current_stream().cuda_stream)
    return idx, wn


def _comb_noalloc(lib, w, zeta):
    import torch
    nw = w.shape[0]
    idx = torch.full((nw,), -7, dtype=torch.int32, device=w.device)
    wn = torch.full((nw,), float("nan"), dtype=torch.float64, device=w.device)
    lib.call("afq_comb_indices_noalloc", nw, w.data_ptr(), zeta.data_ptr(), idx.data_ptr(), wn.data_ptr(),
             torch.cuda.current_stream().cuda_stream)
    return idx, wn


def test_noalloc_comb_matches_caller_scratch(gpu_device):
    import torch
    from ad_afqmc_prototype_amd import _lib
    rng = np.random.default_rng(23)
    for nw in (5000, 777):  # the second call is smaller than the first
        w = np.abs(rng.standard_normal(nw)) * rng.choice([1e-3, 1.0, 30.0], size=nw)
        w[rng.random(nw) < 0.2] = 0.0
        wt = torch.tensor(w, device=gpu_device)
        zeta = torch.tensor(float(rng.random()), dtype=torch.float64, device=gpu_device)
        a = _comb_noalloc(_lib, wt, zeta)
        b = _comb_dev(_lib, wt, zeta)
        torch.cuda.synchronize()
        assert torch.equal(a[0], b[0]), f"nw={nw}: indices differ"
        assert torch.equal(a[1], b[1]), f"nw={nw}: new weights differ"
        assert int(a[0].min()) >= 0 and int(a[0].max()) < nw


def test_noalloc_comb_rejects_oversize(gpu_device):
    from ad_afqmc_prototype_amd import _lib
    with pytest.raises(ValueError, match="exceeds"):
        _lib.call("afq_comb_indices_noalloc", (1 << 18) + 1, 1, 1, 1, 1, 0)
