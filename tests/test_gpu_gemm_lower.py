"""GEMMs for a B whose row j is zero past column j (afq_dgemm_nt_lower,
afq_sgemm_nt_lower, afq_gemm_bf16x3_nt_lower): every output column tile stops
its k-loop at its last column, and the result equals the full-K GEMM on the
same B bit for bit (the skipped products are zeros).  Shapes cover the
exchange's (M = 4096 walker rows, N = K = o n = 6400), ragged tiles, N != K,
and the bf16x3 256 x 256 tiling (>= 512 tiles).  The exchange itself, with
L = Y's lower triangle (diagonal halved) and Z = 2 G L^T, is checked against
the oracle by the energy tests."""
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(4096, 6400, 6400), (37, 300, 304), (130, 264, 256), (64, 520, 512), (4096, 8192, 8192)]


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("M,N,K", SHAPES)
def test_lower_gemm_equals_full_gemm(gpu_device, prec, M, N, K):
    import torch
    from ad_afqmc_prototype_amd import kernels
    if prec == "fp64" and M * N * K > 4096 * 6400 * 6400:
        pytest.skip("fp64 covered at the exchange shape")
    print(f"lower-B GEMM {prec} M={M} N={N} K={K}", flush=True)
    g = torch.Generator(device="cpu").manual_seed(M + N + K)
    A = torch.randn((M, K), dtype=torch.float64, generator=g).to(gpu_device)
    B = torch.tril(torch.randn((N, K), dtype=torch.float64, generator=g)).to(gpu_device)
    full = torch.full((M, N), float("nan"), dtype=torch.float64, device=gpu_device)
    low = torch.full((M, N), float("nan"), dtype=torch.float64, device=gpu_device)
    if prec == "fp64":
        kernels.dgemm_nt(A, B, full, alpha=2.0)
        kernels.gemm_nt_lower(A, B, low, alpha=2.0)
    elif prec == "fp32":
        a32, b32 = kernels.convert_f32(A), kernels.convert_f32(B)
        kernels.sgemm_nt(a32, b32, full, alpha=2.0)
        kernels.gemm_nt_lower(a32, b32, low, alpha=2.0)
    else:
        ah, al = kernels.split_bf16(A)
        bh, bl = kernels.split_bf16(B)
        kernels.gemm_bf16x3_nt(ah, al, bh, bl, full, alpha=2.0, splits=1)
        kernels.gemm_nt_lower((ah, al), (bh, bl), low, alpha=2.0)
    torch.cuda.synchronize()
    assert torch.isfinite(low).all()
    assert torch.equal(low, full)
    # and it is the product (fp32-rounding tolerance for the reduced precisions)
    ref = 2.0 * (A @ B.T)
    tol = 1e-12 if prec == "fp64" else 1e-4
    assert float((low - ref).abs().max()) <= tol * float(ref.abs().max())
