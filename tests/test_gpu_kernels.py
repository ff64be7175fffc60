"""HIP kernels vs the CPU oracle, one entry point at a time (fp64: 1e-10
relative, norm-wise; integer outputs bit-exact)."""
import numpy as np
import pytest

from oracle import afqmc_oracle as orc

from ._pairs import assert_rel, make_pair, rel, random_walkers, to_dev, to_np

pytestmark = pytest.mark.gpu
TOL = 1e-10

SHAPES = [
    # seed, n, o, naux, uhf, kind
    (0, 24, 6, 96, False, "restricted"),
    (1, 40, 7, 150, True, "unrestricted"),
    (5, 45, 11, 130, False, "unrestricted"),
    (2, 200, 32, 1000, False, "restricted"),
    # more than 32 occupied orbitals per spin: column groups in the Trotter,
    # the block Green's function, the library-GEMM energy
    (8, 60, 48, 90, True, "unrestricted"),
    (9, 70, 64, 80, False, "restricted"),
    # the 120-orbital UHF production shape (fp64 Trotter: both spins' 40
    # columns in one workgroup, two k-blocks of V in LDS) and 128 rows with
    # one LDS k-block
    (1, 120, 20, 600, True, "unrestricted"),
    (10, 128, 16, 200, False, "restricted"),
]


@pytest.fixture(scope="module")
def torch_mod(gpu_device):
    import torch
    return torch


def test_dgemm_matches_torch_fp64(gpu_device):
    import torch
    from ad_afqmc_prototype_amd import kernels
    g = torch.Generator(device="cpu").manual_seed(0)
    for (M, N, K) in [(1, 1, 1), (17, 33, 5), (128, 128, 16), (300, 257, 1001), (64, 4096, 96)]:
        A = torch.randn((M, K), dtype=torch.float64, generator=g)
        B = torch.randn((N, K), dtype=torch.float64, generator=g)
        C0 = torch.randn((M, N), dtype=torch.float64, generator=g)
        C = C0.to(gpu_device)
        kernels.dgemm_nt(A.to(gpu_device), B.to(gpu_device), C, alpha=0.5, beta=-1.5)
        ref = 0.5 * (A @ B.T) - 1.5 * C0
        err = rel(C.cpu().numpy(), ref.numpy())
        assert err < 1e-13, f"M={M} N={N} K={K} rel err {err:.3e}"


@pytest.mark.parametrize("nw,naux,k,splits", [
    (3, 5, 7, 2),          # K shorter than one chunk of 16: a single chunk runs
    (40, 150, 630, 3),     # last chunk short (224-deep chunks: 224, 224, 182)
    (64, 96, 1000, 8),
    (100, 600, 4800, None),  # cfg2-like: the shape picks 8 chunks
])
def test_force_bias_split_k(gpu_device, nw, naux, k, splits):
    """The split-K fp64 force bias against a CPU product (each chunk sums its
    k range, the chunks are added in order: fp64 rounding only), and identical
    across two calls (fixed order)."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    g = torch.Generator(device="cpu").manual_seed(nw + k)
    G = torch.randn((2, nw, k), dtype=torch.float64, generator=g)
    R = torch.randn((naux, k), dtype=torch.float64, generator=g)
    ref = 2.0 * torch.einsum("pwk,gk->pwg", G, R)
    Gd, Rd = G.to(gpu_device), R.to(gpu_device)
    fb1 = kernels.force_bias(Gd, Rd, 2.0, splits=splits)
    fb2 = kernels.force_bias(Gd, Rd, 2.0, splits=splits)
    whole = kernels.force_bias(Gd, Rd, 2.0, splits=1)
    assert torch.equal(fb1, fb2)
    assert rel(fb1.cpu().numpy(), ref.numpy()) < 1e-13
    assert rel(whole.cpu().numpy(), ref.numpy()) < 1e-13


@pytest.mark.parametrize("seed,n,o,naux,uhf,kind", SHAPES)
def test_greens_and_log_overlap(gpu_device, seed, n, o, naux, uhf, kind):
    from ad_afqmc_prototype_amd.trial import rhf as trial_rhf
    p = make_pair(seed, n, o, naux, uhf=uhf, kind=kind, device=gpu_device)
    rng = np.random.default_rng(seed + 10)
    nw = 16
    w = random_walkers(rng, kind, nw, n, o, p.s.mo_up, p.s.mo_dn)
    gc = trial_rhf.greens(to_dev(w, gpu_device), p.g["trial"])
    lo = gc.log_overlap.cpu().numpy()
    lo_ref = orc.overlap(w, p.sd.trial, kind, mode="log")
    np.testing.assert_allclose(lo.real, lo_ref.real, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(np.exp(1j * lo.imag), np.exp(1j * lo_ref.imag), atol=1e-10)
    G = gc.greens.cpu().numpy()
    nspin = 1 if kind == "restricted" else 2
    Gc = (G[0] + 1j * G[1]).reshape(nw, nspin, o, n)
    ws = (w,) if kind == "restricted" else w
    cs = (p.s.mo_up,) if kind == "restricted" else (p.s.mo_up, p.s.mo_dn)
    for s_ in range(nspin):
        assert_rel(Gc[:, s_], orc.half_green(ws[s_], cs[s_]), TOL)


@pytest.mark.parametrize("seed,n,o,naux,uhf,kind", SHAPES)
def test_force_bias(gpu_device, seed, n, o, naux, uhf, kind):
    from ad_afqmc_prototype_amd.core.ops import k_force_bias
    p = make_pair(seed, n, o, naux, uhf=uhf, kind=kind, device=gpu_device)
    rng = np.random.default_rng(seed + 11)
    w = random_walkers(rng, kind, 24, n, o, p.s.mo_up, p.s.mo_dn)
    fb = p.g["m_ops"].require_kernel(k_force_bias)(to_dev(w, gpu_device), p.g["ham"], p.g["m_ctx"], p.g["trial"])
    assert_rel(fb.cpu().numpy(), orc.force_bias(w, p.sd.trial, p.sd.mctx, kind), TOL)


@pytest.mark.parametrize("seed,n,o,naux,uhf,kind", SHAPES)
def test_field_shifts(gpu_device, seed, n, o, naux, uhf, kind):
    import torch
    from ad_afqmc_prototype_amd import kernels
    p = make_pair(seed, n, o, naux, uhf=uhf, kind=kind, device=gpu_device)
    rng = np.random.default_rng(seed + 12)
    nw = 33
    x = rng.standard_normal((nw, naux))
    fb = rng.standard_normal((nw, naux)) + 1j * rng.standard_normal((nw, naux))
    fbp = torch.tensor(np.stack([fb.real, fb.imag]), device=gpu_device)
    xbar, st, ft = kernels.field_shifts(torch.tensor(x, device=gpu_device), fbp, p.g["p_ctx"].mf_shifts,
                                        p.g["p_ctx"].sqrt_dt)
    sh, st_ref, ft_ref = orc.field_shift_terms(x, fb, p.sd.pctx)
    xb = xbar.cpu().numpy()
    assert_rel(xb[0] + 1j * xb[1], sh, 1e-13)
    assert_rel(st.cpu().numpy(), st_ref, 1e-12)
    assert_rel(ft.cpu().numpy(), ft_ref, 1e-12)


@pytest.mark.parametrize("seed,n,o,naux,uhf,kind", SHAPES)
def test_vhs_and_trotter(gpu_device, seed, n, o, naux, uhf, kind):
    import torch
    from ad_afqmc_prototype_amd import kernels
    from ad_afqmc_prototype_amd.prop.chol_afqmc_ops import chol_transposed
    p = make_pair(seed, n, o, naux, uhf=uhf, kind=kind, device=gpu_device)
    rng = np.random.default_rng(seed + 13)
    nw = 8
    w = random_walkers(rng, kind, nw, n, o, p.s.mo_up, p.s.mo_dn)
    x = 0.3 * (rng.standard_normal((nw, naux)) + 0.5j * rng.standard_normal((nw, naux)))
    xp = torch.tensor(np.stack([x.real, x.imag]), device=gpu_device)
    v = kernels.vhs(xp, chol_transposed(p.g["ham"].chol), n).cpu().numpy()[..., :n]
    assert_rel(v[0] + 1j * v[1], orc.make_vhs(p.sd.chol, x), TOL)
    out = p.g["p_ops"].apply_trotter_planes(to_dev(w, gpu_device), xp, p.g["p_ctx"], 6)
    ref = orc.apply_trotter(w, x, p.sd.pctx, p.sd.chol, 6, kind)
    if kind == "restricted":
        assert_rel(to_np(out), ref, TOL)
    else:
        assert_rel(to_np(out)[0], ref[0], TOL)
        assert_rel(to_np(out)[1], ref[1], TOL)
    # n_chunks does not change the result
    out2 = p.g["p_ops"].apply_trotter_planes(to_dev(w, gpu_device), xp, p.g["p_ctx"], 6, n_chunks=2)
    a, b = (to_np(out), to_np(out2))
    if kind == "restricted":
        np.testing.assert_array_equal(a, b)
    else:
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])


def test_trotter_zero_terms_and_identity(gpu_device):
    """n_terms = 0 gives E E w; zero VHS with E = I returns w exactly."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    rng = np.random.default_rng(3)
    nw, n, o = 5, 37, 9
    w = rng.standard_normal((nw, n, o)) + 1j * rng.standard_normal((nw, n, o))
    ldv = kernels.ldv_for(n)
    e = rng.standard_normal((n, n)) * 0.2
    ep = np.zeros((n, ldv)); ep[:, :n] = e
    v = torch.zeros((2, nw, n, ldv), dtype=torch.float64, device=gpu_device)
    wd = torch.tensor(w, device=gpu_device)
    out = kernels.trotter(wd, v, torch.tensor(ep, device=gpu_device), 0.1, 0).cpu().numpy()
    assert_rel(out, e @ (e @ w), 1e-13)
    ep_i = np.zeros((n, ldv)); ep_i[:, :n] = np.eye(n)
    out = kernels.trotter(wd, v, torch.tensor(ep_i, device=gpu_device), 0.1, 6).cpu().numpy()
    np.testing.assert_array_equal(out, w)


@pytest.mark.parametrize("exchange", ["auto", "per_g"])
@pytest.mark.parametrize("seed,n,o,naux,uhf,kind", SHAPES)
def test_energy(gpu_device, seed, n, o, naux, uhf, kind, exchange):
    """Local energies against the oracle at 1e-10, the exchange both as the
    quadratic form vec(G)^T Y vec(G) ("auto" where n < naux) and on the fused
    per-g kernels."""
    from ad_afqmc_prototype_amd.core.ops import k_energy
    from ad_afqmc_prototype_amd.meas.rhf import make_rhf_meas_ops
    p = make_pair(seed, n, o, naux, uhf=uhf, kind=kind, device=gpu_device)
    rng = np.random.default_rng(seed + 14)
    nw = 4 if n >= 200 else 12
    w = random_walkers(rng, kind, nw, n, o, p.s.mo_up, p.s.mo_dn)
    m_ops = make_rhf_meas_ops(p.g["sys"], exchange=exchange)
    ctx = m_ops.build_meas_ctx(p.g["ham"], p.g["trial"])
    e = m_ops.require_kernel(k_energy)(to_dev(w, gpu_device), p.g["ham"], ctx, p.g["trial"])
    e_ref = orc.energy(w, p.sd.h0, p.sd.trial, p.sd.mctx, kind)
    assert_rel(e.cpu().numpy(), e_ref, TOL)


@pytest.mark.parametrize("seed,n,o,naux,uhf,kind", [
    # stacked-row energy (o packs poorly into 16-row tiles): q = 64 // o g's per
    # chunk, naux not a multiple of q, so the last chunk is ragged
    (3, 60, 20, 97, True, "unrestricted"),    # q = 3, last chunk 1 g
    (4, 50, 17, 50, False, "restricted"),     # q = 3, last chunk 2 g's
    (6, 64, 21, 40, False, "unrestricted"),   # q = 3, last chunk 1 g
])
def test_energy_stacked_rows(gpu_device, seed, n, o, naux, uhf, kind):
    """The fp64 energy on the shapes that take stacked g chunks, against the
    oracle at 1e-10."""
    test_energy(gpu_device, seed, n, o, naux, uhf, kind, "per_g")


@pytest.mark.parametrize("n,o", [(24, 6), (120, 20), (200, 32), (64, 64)])
def test_orthonormalize(gpu_device, n, o):
    from ad_afqmc_prototype_amd import kernels
    import torch
    rng = np.random.default_rng(n + o)
    nw = 6
    w = rng.standard_normal((nw, n, o)) + 1j * rng.standard_normal((nw, n, o))
    q = kernels.orthonormalize(torch.tensor(w, device=gpu_device)).cpu().numpy()
    q_ref = orc.orthonormalize(w, "restricted")
    assert_rel(q, q_ref, 1e-12)


def test_comb_indices_bit_exact(gpu_device):
    import torch
    from ad_afqmc_prototype_amd import kernels
    rng = np.random.default_rng(7)
    for nw in (1, 5, 128, 1000, 2048, 5000):
        for trial in range(3):
            w = np.abs(rng.standard_normal(nw)) * rng.choice([1e-3, 1.0, 30.0], size=nw)
            w[rng.random(nw) < 0.2] = 0.0
            if trial == 2:
                w[:] = 0.0
            z = float(rng.random())
            idx, w_new = kernels.comb_indices(torch.tensor(w, device=gpu_device), z)
            np.testing.assert_array_equal(idx.cpu().numpy(), orc.sr_indices(w, z, nw))
            np.testing.assert_array_equal(w_new.cpu().numpy(), np.full(nw, np.cumsum(np.abs(w))[-1] / nw))


def test_comb_indices_nan_weights(gpu_device):
    """A NaN weight: every tooth lands on the first NaN of the cumulative sum
    (searchsorted's NaN order), all new weights NaN -- as the oracle."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    rng = np.random.default_rng(17)
    for nw, pos in ((8, 0), (8, 3), (8, 7), (2048, 1500)):
        w = rng.random(nw)
        w[pos] = np.nan
        if nw > 8:
            w[pos + 10] = np.nan
        idx, w_new = kernels.comb_indices(torch.tensor(w, device=gpu_device), 0.41)
        np.testing.assert_array_equal(idx.cpu().numpy(), orc.sr_indices(w, 0.41, nw))
        np.testing.assert_array_equal(idx.cpu().numpy(), np.full(nw, pos))
        assert np.all(np.isnan(w_new.cpu().numpy()))


def test_block_energy_hand_cases(gpu_device):
    """Clip window sqrt(2/dt) (0.99 t kept, 1.01 t replaced), all-zero weights
    (block energy = e_ref) and the shift EMA, against known answers."""
    import math
    import torch
    from ad_afqmc_prototype_amd import kernels
    from .test_oracle_pinning import _block_cases
    for e, w, e_ref, pop, dt, ema, (eb, ws, pn) in _block_cases():
        out = kernels.block_energy(torch.tensor(np.array(e, dtype=complex), device=gpu_device),
                                   torch.tensor(np.array(w), device=gpu_device),
                                   torch.tensor(pop, dtype=torch.float64, device=gpu_device),
                                   e_ref=e_ref, thresh=math.sqrt(2.0 / dt), shift_ema=ema).cpu().numpy()
        assert out[0] == pytest.approx(eb, rel=1e-14, abs=1e-14)
        assert out[1] == ws
        assert out[2] == pytest.approx(pn, rel=1e-14, abs=1e-14)
        ref = orc.block_energy(np.array(e), np.array(w), e_ref, pop, dt, ema)
        assert out[0] == pytest.approx(ref[0], rel=1e-15, abs=1e-15)


def test_gather_rows(gpu_device):
    import torch
    from ad_afqmc_prototype_amd import kernels
    rng = np.random.default_rng(8)
    src = rng.standard_normal((2, 9, 5, 3))
    idx = np.array([8, 0, 0, 3, 8, 1, 2, 2, 7], dtype=np.int32)
    out = kernels.gather_rows(torch.tensor(src, device=gpu_device), torch.tensor(idx, device=gpu_device), nplanes=2)
    np.testing.assert_array_equal(out.cpu().numpy(), src[:, idx])
    c = rng.standard_normal((9, 4)) + 1j * rng.standard_normal((9, 4))
    out = kernels.gather_rows(torch.tensor(c, device=gpu_device), torch.tensor(idx, device=gpu_device))
    np.testing.assert_array_equal(out.cpu().numpy(), c[idx])


def test_phaseless_update_cases(gpu_device):
    import torch
    from ad_afqmc_prototype_amd import kernels
    rng = np.random.default_rng(9)
    nw = 700
    lo_old = rng.standard_normal(nw) + 1j * rng.uniform(-3, 3, nw)
    dl = 0.05 * rng.standard_normal(nw) + 1j * rng.uniform(-2.5, 2.5, nw)
    dl[:20] += 20.0     # above the cap
    dl[20:40] -= 20.0   # below the floor
    lo_new = lo_old + dl
    st = rng.standard_normal(nw) + 1j * rng.standard_normal(nw)
    ft = 0.01 * (rng.standard_normal(nw) + 1j * rng.standard_normal(nw))
    w = rng.uniform(0.5, 2.0, nw)
    w[40] = np.nan
    pctx = orc.prop_ctx(dt=0.01, sqrt_dt=0.1, exp_h1_half=None, mf_shifts=None, h0_prop=0.3 + 0j)
    prm = orc.params(dt=0.01)
    w_ref, pop_ref = orc.phaseless_update(lo_new, lo_old, st, ft, w, -1.5, -2.0, pctx, prm, "log")
    d = gpu_device
    logdet = torch.tensor((lo_new / 2.0).reshape(nw, 1), device=d)
    out_lo, out_w, out_pop = kernels.phaseless_update(
        logdet, torch.tensor(lo_old, device=d), torch.tensor(st, device=d), torch.tensor(ft, device=d),
        torch.tensor(w, device=d), torch.tensor(-1.5, dtype=torch.float64, device=d), e_estimate=-2.0,
        h0_prop=0.3, dt=0.01, sqrt_dt=0.1, weight_floor=1e-3, weight_cap=100.0, damping=0.1)
    ow = out_w.cpu().numpy()
    np.testing.assert_array_equal(np.isnan(ow), np.isnan(w_ref))
    np.testing.assert_array_equal(ow == 0, w_ref == 0)
    m = ~np.isnan(w_ref)
    np.testing.assert_allclose(ow[m], w_ref[m], rtol=1e-12)
    assert np.isnan(out_pop.item()) and np.isnan(pop_ref)
    # without the NaN walker the population shift is finite and matches
    w[40] = 1.0
    w_ref, pop_ref = orc.phaseless_update(lo_new, lo_old, st, ft, w, -1.5, -2.0, pctx, prm, "log")
    _, out_w, out_pop = kernels.phaseless_update(
        logdet, torch.tensor(lo_old, device=d), torch.tensor(st, device=d), torch.tensor(ft, device=d),
        torch.tensor(w, device=d), torch.tensor(-1.5, dtype=torch.float64, device=d), e_estimate=-2.0,
        h0_prop=0.3, dt=0.01, sqrt_dt=0.1, weight_floor=1e-3, weight_cap=100.0, damping=0.1)
    np.testing.assert_allclose(out_w.cpu().numpy(), w_ref, rtol=1e-12)
    np.testing.assert_allclose(out_pop.item(), pop_ref, rtol=1e-12)


def test_invalid_arguments_raise(gpu_device):
    import torch
    from ad_afqmc_prototype_amd import _lib, kernels
    w = torch.zeros((2, 10, 3), dtype=torch.complex128, device=gpu_device)
    with pytest.raises(ValueError):
        kernels.overlap_greens(w, torch.zeros((10, 4), dtype=torch.float64, device=gpu_device))
    with pytest.raises(ValueError):
        kernels.orthonormalize(w.cpu())
    with pytest.raises(ValueError, match="exceeds"):
        _lib.call("afq_trotter", 1, 400, 40, 1, 400, 6, 0.1, 1, 1, 2, 0, 3, 0, 0)
    with pytest.raises(ValueError, match="aliases"):
        _lib.call("afq_trotter", 1, 8, 2, 1, 8, 6, 0.1, 1, 1, 2, 0, 2, 0, 0)


@pytest.mark.parametrize("n,naux,nw,prec", [
    (1, 3, 2, "fp64"), (17, 40, 5, "fp64"), (45, 130, 3, "fp64"), (200, 1000, 16, "fp64"),
    (1, 3, 2, "fp32"), (17, 40, 5, "fp32"), (200, 1000, 16, "fp32"),
    (1, 8, 2, "bf16x3"), (24, 96, 8, "bf16x3"), (45, 130, 3, "bf16x3"), (200, 1000, 16, "bf16x3"),
    (200, 1000, 2048, "bf16x3"),  # the 256 x 256 tile kernel
])
def test_symmetric_vhs_matches_full_gemm(gpu_device, n, naux, nw, prec):
    """The symmetric contraction (upper 16 x 16 blocks, each entry stored at
    (p, q) and (q, p)) equals the full-row GEMM bit for bit (same k order per
    entry; L_g exactly symmetric), padding columns zero, every entry written."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    from ad_afqmc_prototype_amd.prop.chol_afqmc_ops import chol_sym_rows, chol_transposed
    print(f"symmetric VHS n={n} naux={naux} rows={2 * nw} {prec}", flush=True)
    rng = np.random.default_rng(n)
    a = rng.standard_normal((naux, n, n))
    chol = torch.tensor(0.5 * (a + a.transpose(0, 2, 1)), device=gpu_device)
    x = torch.tensor(rng.standard_normal((2 * nw, naux)), device=gpu_device)
    ldv = kernels.ldv_for(n)
    full_b, sym_b = chol_transposed(chol), chol_sym_rows(chol)
    odt = torch.float64 if prec == "fp64" else torch.float32
    ref = torch.empty((2 * nw, n * ldv), dtype=odt, device=gpu_device)
    if prec == "fp64":
        kernels.dgemm_nt(x, full_b, ref)
        a_op, b_op = x, sym_b
    elif prec == "fp32":
        x32 = kernels.convert_f32(x)
        kernels.sgemm_nt(x32, kernels.convert_f32(full_b), ref)
        a_op, b_op = x32, kernels.convert_f32(sym_b)
    else:
        a_op = kernels.split_bf16(x)
        kernels.gemm_bf16x3_nt(*a_op, *kernels.split_bf16(full_b), ref)
        b_op = kernels.split_bf16(sym_b)
    out = torch.full((2 * nw, n, ldv), float("nan"), dtype=odt, device=gpu_device)
    kernels.vhs_sym(a_op, b_op, n, out=out)
    assert torch.equal(out.view(2 * nw, n * ldv), ref)
    assert torch.equal(out[..., n:], torch.zeros_like(out[..., n:]))
