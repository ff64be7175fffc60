"""Mixed-precision kernels (fp32 and split-bf16 contractions, fp32 Trotter
step) against fp64.  Tolerances are stated per test: fp32 ~ 1e-6 relative,
bf16x3 (hi/lo split, fp32 accumulation) ~ 1e-4 relative."""
import numpy as np
import pytest

from ad_afqmc_prototype_amd import synthetic as syn
from oracle import afqmc_oracle as orc

from ._pairs import assert_rel, make_pair, rel, to_dev, to_np

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,N,K", [(1, 1, 8), (17, 33, 40), (256, 130, 1000), (300, 257, 512)])
def test_sgemm_and_bf16x3_gemm(gpu_device, M, N, K):
    import torch
    from ad_afqmc_prototype_amd import kernels
    g = torch.Generator().manual_seed(M + N + K)
    A = torch.randn((M, K), dtype=torch.float64, generator=g)
    B = torch.randn((N, K), dtype=torch.float64, generator=g)
    ref = (A @ B.T).numpy()
    Ad, Bd = A.to(gpu_device), B.to(gpu_device)
    C = torch.empty((M, N), dtype=torch.float64, device=gpu_device)
    kernels.sgemm_nt(kernels.convert_f32(Ad), kernels.convert_f32(Bd), C)
    assert_rel(C.cpu().numpy(), ref, 2e-6)
    C32 = torch.empty((M, N), dtype=torch.float32, device=gpu_device)
    kernels.sgemm_nt(kernels.convert_f32(Ad), kernels.convert_f32(Bd), C32, alpha=0.5)
    assert_rel(C32.cpu().numpy(), 0.5 * ref, 2e-6)
    ah, al = kernels.split_bf16(Ad)
    bh, bl = kernels.split_bf16(Bd)
    kernels.gemm_bf16x3_nt(ah, al, bh, bl, C, alpha=2.0)
    err = rel(C.cpu().numpy(), 2.0 * ref)
    assert err < 1e-4
    # the split is exact to ~2^-16: hi + lo reproduces x to that precision
    hi = (ah.cpu().to(torch.int32) & 0xFFFF) << 16
    lo = (al.cpu().to(torch.int32) & 0xFFFF) << 16
    x = hi.view(torch.float32).double() + lo.view(torch.float32).double()
    assert_rel(x[:, :K].numpy(), A.numpy(), 2e-5)


@pytest.mark.parametrize("M,N,K,c64", [(4100, 8200, 1000, False), (4096, 8192, 968, True)])
def test_bf16x3_gemm_large_tiles(gpu_device, M, N, K, c64):
    """The 256 x 256 tile path (taken at >= 512 tiles) with partial M / N tiles
    and a partial last k-step, against fp64: same 1e-4 bound as above."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    g = torch.Generator(device=gpu_device).manual_seed(M + K)
    A = torch.randn((M, K), dtype=torch.float64, device=gpu_device, generator=g)
    B = torch.randn((N, K), dtype=torch.float64, device=gpu_device, generator=g)
    ref = A @ B.T
    ah, al = kernels.split_bf16(A)
    bh, bl = kernels.split_bf16(B)
    C = torch.full((M, N), float("nan"), dtype=torch.float64 if c64 else torch.float32, device=gpu_device)
    kernels.gemm_bf16x3_nt(ah, al, bh, bl, C, alpha=0.5)
    err = (torch.linalg.norm(C.double() - 0.5 * ref) / torch.linalg.norm(0.5 * ref)).item()
    assert err < 1e-4, err
    # every element written, and the largest single-element error is small too
    assert torch.isfinite(C).all()
    assert ((C.double() - 0.5 * ref).abs().max() / (0.5 * ref).abs().max()).item() < 1e-4


def test_split_bf16_zero_padding_and_plain_bf16(gpu_device):
    import torch
    from ad_afqmc_prototype_amd import kernels
    x = torch.tensor([[1.0, -2.5, 3.0e-3]], dtype=torch.float64, device=gpu_device)
    hi, lo = kernels.split_bf16(x)
    assert hi.shape == (1, 8)
    assert torch.all(hi[0, 3:] == 0) and torch.all(lo[0, 3:] == 0)
    # values exactly representable in bf16 have a zero low part
    assert int(lo[0, 0]) == 0 and int(lo[0, 1]) == 0
    assert int(hi[0, 0]) == 0x3F80 and (int(hi[0, 1]) & 0xFFFF) == 0xC020


@pytest.mark.parametrize("precision,tol", [("fp32", 2e-5), ("bf16x3", 5e-4)])
@pytest.mark.parametrize("case", [
    dict(seed=0, n=24, o=6, g=96, uhf=False, kind="restricted"),
    dict(seed=2, n=200, o=32, g=1000, uhf=False, kind="restricted"),
    dict(seed=1, n=120, o=20, g=600, uhf=True, kind="unrestricted"),
])
def test_mixed_trotter_matches_fp64(gpu_device, precision, tol, case):
    import torch
    from ad_afqmc_prototype_amd.prop.chol_afqmc_ops import make_chol_afqmc_ops
    p = make_pair(case["seed"], case["n"], case["o"], case["g"], uhf=case["uhf"], kind=case["kind"],
                  dt=0.005, chol_scale=0.2, device=gpu_device)
    rng = np.random.default_rng(5)
    nw = 8
    st = orc.init_state(p.sd, nw, overlap_mode="log")
    w = st.walkers
    x = rng.standard_normal((nw, case["g"])) + 0.3j * rng.standard_normal((nw, case["g"]))
    xp = torch.tensor(np.stack([x.real, x.imag]), device=gpu_device)
    ops = make_chol_afqmc_ops(p.g["ham"], case["kind"], precision=precision)
    ctx = ops.build_prop_ctx(p.g["t_ops"].get_rdm1(p.g["trial"]), 0.005)
    out = to_np(ops.apply_trotter_planes(to_dev(w, gpu_device), xp, ctx, 6))
    ref = orc.apply_trotter(w, x, p.sd.pctx, p.sd.chol, 6, case["kind"])
    if case["kind"] == "restricted":
        assert_rel(out, ref, tol)
    else:
        assert max(rel(out[0], ref[0]), rel(out[1], ref[1])) < tol


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("bf16x3", 3e-4)])
def test_mixed_force_bias(gpu_device, precision, tol):
    from ad_afqmc_prototype_amd.core.ops import k_force_bias
    from ad_afqmc_prototype_amd.meas.rhf import make_rhf_meas_ops
    p = make_pair(2, 200, 32, 1000, device=gpu_device)
    m_ops = make_rhf_meas_ops(p.g["sys"], precision=precision)
    ctx = m_ops.build_meas_ctx(p.g["ham"], p.g["trial"])
    st = orc.init_state(p.sd, 16, overlap_mode="log")
    rng = np.random.default_rng(1)
    w = st.walkers + 0.1 * (rng.standard_normal(st.walkers.shape) + 1j * rng.standard_normal(st.walkers.shape))
    fb = m_ops.require_kernel(k_force_bias)(to_dev(w, gpu_device), p.g["ham"], ctx, p.g["trial"])
    assert_rel(fb.cpu().numpy(), orc.force_bias(w, p.sd.trial, p.sd.mctx, "restricted"), tol)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_mixed_block_energy_close_to_fp64(gpu_device, precision):
    """Common random numbers: the mixed run and the fp64 oracle see the same
    fields and comb uniforms; block energies agree to well below 0.1 mHa on the
    24-orbital workload at Cholesky scale 0.25 (the phaseless walk's working
    regime: no local energy is clipped)."""
    import dataclasses
    import torch
    from ad_afqmc_prototype_amd.meas.rhf import make_rhf_meas_ops
    from ad_afqmc_prototype_amd.prop.afqmc import make_key
    from ad_afqmc_prototype_amd.prop.blocks import afqmc_block
    from ad_afqmc_prototype_amd.prop.chol_afqmc_ops import make_chol_afqmc_ops
    from ad_afqmc_prototype_amd.prop.types import prop_state
    from ad_afqmc_prototype_amd.trial import rhf as trial_rhf
    seed, n, o, g, nw, dt = 0, 24, 6, 96, 128, 0.01
    p = make_pair(seed, n, o, g, dt=dt, chol_scale=0.25, device=gpu_device)
    st0 = orc.init_state(p.sd, nw, overlap_mode="log")
    fields = [syn.fields_for_step(seed, t, nw, g) for t in range(40)]
    zetas = [syn.zeta_for_event(seed, k) for k in range(4)]
    _, blocks = orc.run_cadenced(st0, p.sd, p.prm, n_steps=40, qr_every=5, sr_every=10,
                                 fields_fn=lambda t: fields[t], zeta_fn=lambda k: zetas[k],
                                 overlap_mode="log")
    m_ops = make_rhf_meas_ops(p.g["sys"], precision=precision)
    p_ops = make_chol_afqmc_ops(p.g["ham"], "restricted", precision=precision)
    p_ctx = p_ops.build_prop_ctx(p.g["t_ops"].get_rdm1(p.g["trial"]), dt)
    m_ctx = m_ops.build_meas_ctx(p.g["ham"], p.g["trial"])
    w = to_dev(st0.walkers, gpu_device)
    gs = prop_state(walkers=w, weights=torch.ones(nw, dtype=torch.float64, device=gpu_device),
                    log_overlaps=torch.tensor(st0.overlaps, device=gpu_device), rng_key=make_key(0),
                    pop_control_ene_shift=torch.tensor(st0.e_estimate, dtype=torch.float64, device=gpu_device),
                    e_estimate=st0.e_estimate, greens=trial_rhf.greens(w, p.g["trial"]))
    params = dataclasses.replace(p.g["params"], n_prop_steps=10, n_ortho=5)
    worst = 0.0
    for k in range(4):
        f = [torch.tensor(fields[10 * k + t], device=gpu_device) for t in range(10)]
        gs, obs = afqmc_block(gs, sys=p.g["sys"], params=params, ham_data=p.g["ham"],
                              trial_data=p.g["trial"], meas_ops=m_ops, prop_ops=p_ops, prop_ctx=p_ctx,
                              meas_ctx=m_ctx, fields=f, zeta=zetas[k])
        worst = max(worst, abs(obs.scalars["energy"].item() - blocks[k]["energy"]))
    assert worst < 1e-4, worst  # Ha
    # the walk is live: block energies move and none is the clipped estimate
    e = np.array([b["energy"] for b in blocks])
    assert np.ptp(e) > 1e-3 and np.all(np.abs(e - st0.e_estimate) > 1e-6), e


@pytest.mark.parametrize("case", [
    dict(seed=2, n=200, o=32, g=1000, uhf=False, kind="restricted", nw=5),
    dict(seed=1, n=120, o=20, g=600, uhf=True, kind="unrestricted", nw=7),
    dict(seed=0, n=24, o=6, g=96, uhf=False, kind="restricted", nw=9),
    # stacked g chunks (o = 7: 9 g's per chunk, o = 12: 5, o = 19: 3), the last
    # chunk ragged
    dict(seed=5, n=40, o=7, g=101, uhf=False, kind="restricted", nw=5),
    dict(seed=6, n=48, o=12, g=61, uhf=True, kind="unrestricted", nw=6),
    dict(seed=7, n=70, o=19, g=130, uhf=False, kind="restricted", nw=3),
])
@pytest.mark.parametrize("exchange", ["auto", "per_g"])
def test_energy_fp32_contractions(gpu_device, case, exchange):
    """fp32 f_g contractions with fp64 sums (or the fp32 G Y product of the
    exchange as a quadratic form): per-walker energies within 2e-5 of
    |E| of the fp64 oracle (two walkers per workgroup, odd walker count).  The
    random walkers here are far from the trial (|E| up to 2e4 Ha, large
    imaginary parts), the hardest case for fp32."""
    from ad_afqmc_prototype_amd.core.ops import k_energy
    from ad_afqmc_prototype_amd.meas.rhf import make_rhf_meas_ops
    p = make_pair(case["seed"], case["n"], case["o"], case["g"], uhf=case["uhf"], kind=case["kind"],
                  device=gpu_device)
    m_ops = make_rhf_meas_ops(p.g["sys"], precision="fp32", exchange=exchange)
    ctx = m_ops.build_meas_ctx(p.g["ham"], p.g["trial"])
    rng = np.random.default_rng(3)
    from ._pairs import random_walkers
    w = random_walkers(rng, case["kind"], case["nw"], case["n"], case["o"], p.s.mo_up, p.s.mo_dn)
    e = m_ops.require_kernel(k_energy)(to_dev(w, gpu_device), p.g["ham"], ctx, p.g["trial"]).cpu().numpy()
    e_ref = orc.energy(w, p.sd.h0, p.sd.trial, p.sd.mctx, case["kind"])
    err = np.abs(e - e_ref) / np.abs(e_ref)
    assert err.max() < 2e-5, err.max()


def _trotter_inputs(rng, nw, n, o, nspin, dt, vnorm):
    """Complex-symmetric VHS (as sum_g xbar_g L_g is) with sqrt(dt) ||V_w||_2 =
    vnorm, E = expm(-dt h1 / 2) for a symmetric h1 of norm ~2, walkers."""
    import scipy.linalg
    h1 = rng.standard_normal((n, n)) / np.sqrt(n)
    h1 = h1 + h1.T
    e = scipy.linalg.expm(-0.5 * dt * h1)
    a = rng.standard_normal((nw, n, n)) + 1j * rng.standard_normal((nw, n, n))
    v = 0.5 * (a + np.transpose(a, (0, 2, 1)))
    v *= vnorm / (np.sqrt(dt) * np.linalg.norm(v, 2, axis=(1, 2)))[:, None, None]
    ws = [rng.standard_normal((nw, n, o)) + 1j * rng.standard_normal((nw, n, o)) for _ in range(nspin)]
    return e, v, ws


def _planes(e, v, ldv):
    n = e.shape[0]
    ep = np.zeros((n, ldv))
    ep[:, :n] = e
    vp = np.zeros((2, v.shape[0], n, ldv))
    vp[0, :, :, :n] = v.real
    vp[1, :, :, :n] = v.imag
    return ep, vp


@pytest.mark.parametrize("nw,n,o,nspin", [
    (8, 24, 6, 1),     # all of V in registers
    (5, 100, 20, 2),   # odd row-tile count, o not a multiple of 16, nw not a multiple of 8
    (8, 180, 32, 1),   # one k-block of V in LDS
    (16, 200, 32, 1),  # two k-blocks in LDS (the 200-orbital shape)
    (3, 224, 17, 2),   # largest on-chip size
    (2, 250, 8, 1),    # beyond on-chip capacity: the fp32 step runs
    (6, 80, 48, 2),    # more than 32 occupied orbitals (three column tiles)
])
def test_trotter_bf16x3_vs_fp64(gpu_device, nw, n, o, nspin):
    """Split-bf16 Trotter step against the fp64 step (pinned to the oracle)
    on the same E, VHS and walkers, sqrt(dt) ||V||_2 = 0.3, dt = 0.005.  The
    split (~2^-17 relative) enters through sqrt(dt) V / k and E - I, so each
    walker is within 2e-5 (norm-wise relative); the fp32 step within 2e-6."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    rng = np.random.default_rng(n + 7 * o + nspin)
    dt = 0.005
    e, v, ws = _trotter_inputs(rng, nw, n, o, nspin, dt, 0.3)
    ldv = kernels.ldv_for(n)
    ep, vp = _planes(e, v, ldv)
    ep64, vp64 = torch.tensor(ep, device=gpu_device), torch.tensor(vp, device=gpu_device)
    ep32, vp32 = ep64.float(), vp64.float()
    wd = tuple(torch.tensor(x, device=gpu_device) for x in ws)
    wd = wd if nspin == 2 else wd[0]
    ref = to_np(kernels.trotter(wd, vp64, ep64, dt ** 0.5, 6))
    got = to_np(kernels.trotter(wd, vp32, ep32, dt ** 0.5, 6, arith="bf16x3"))
    f32 = to_np(kernels.trotter(wd, vp32, ep32, dt ** 0.5, 6, arith="fp32"))
    refs, gots, f32s = [x if nspin == 2 else (x,) for x in (ref, got, f32)]
    for r, g, f in zip(refs, gots, f32s):
        err = max(rel(g[i], r[i]) for i in range(nw))
        assert err < 2e-5, err
        err32 = max(rel(f[i], r[i]) for i in range(nw))
        assert err32 < 2e-6, err32


def test_trotter_bf16x3_exact_identity(gpu_device):
    """E = I and V = 0: E - I is exactly zero, so the step returns the fp32
    rounding of the walkers bit for bit (X is added exactly, not split)."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    rng = np.random.default_rng(11)
    nw, n, o = 8, 200, 32
    ldv = kernels.ldv_for(n)
    w = rng.standard_normal((nw, n, o)) + 1j * rng.standard_normal((nw, n, o))
    ep = np.zeros((n, ldv), dtype=np.float32)
    ep[:, :n] = np.eye(n)
    v = torch.zeros((2, nw, n, ldv), dtype=torch.float32, device=gpu_device)
    out = kernels.trotter(torch.tensor(w, device=gpu_device), v, torch.tensor(ep, device=gpu_device),
                          0.07, 6, arith="bf16x3").cpu().numpy()
    np.testing.assert_array_equal(out, w.astype(np.complex64).astype(np.complex128))


@pytest.mark.parametrize("case,tol", [
    (dict(seed=2, n=200, o=32, g=1000, uhf=False, kind="restricted", nw=5, spread=0.05), 2e-6),
    (dict(seed=1, n=120, o=20, g=600, uhf=True, kind="unrestricted", nw=3, spread=0.05), 2e-6),
    (dict(seed=0, n=24, o=6, g=96, uhf=False, kind="restricted", nw=9, spread=0.05), 1e-5),
    (dict(seed=2, n=200, o=32, g=1000, uhf=False, kind="restricted", nw=4, spread=0.3), 1e-4),
])
@pytest.mark.parametrize("exchange", ["auto", "per_g"])
def test_energy_bf16x3_contractions(gpu_device, case, tol, exchange):
    """Split-bf16 f_g contractions with fp64 sums (or the split-bf16 G Y
    product of the exchange as a quadratic form) against the fp64 oracle.
    The split keeps ~2^-17 of each operand, and that error enters the
    Coulomb square and the exchange coherently, so the bound is relative to
    |E|: 2e-6 for walkers near the trial at the working Cholesky scale, 1e-5
    for the 24-orbital system at scale 1 (two-body parts a larger share of
    |E|; still below one unit of 2^-17), 1e-4 for walkers far from the trial
    (|E| ~ 1e3, large Coulomb and exchange parts that cancel).  Odd walker
    counts leave a half-filled two-walker workgroup."""
    from ad_afqmc_prototype_amd.core.ops import k_energy
    from ad_afqmc_prototype_amd.meas.rhf import make_rhf_meas_ops
    from ._pairs import random_walkers
    p = make_pair(case["seed"], case["n"], case["o"], case["g"], uhf=case["uhf"], kind=case["kind"],
                  chol_scale=0.03 if case["n"] == 200 else 1.0, device=gpu_device)
    m_ops = make_rhf_meas_ops(p.g["sys"], precision="bf16x3", exchange=exchange)
    ctx = m_ops.build_meas_ctx(p.g["ham"], p.g["trial"])
    rng = np.random.default_rng(4)
    w = random_walkers(rng, case["kind"], case["nw"], case["n"], case["o"], p.s.mo_up, p.s.mo_dn,
                       spread=case["spread"])
    e = m_ops.require_kernel(k_energy)(to_dev(w, gpu_device), p.g["ham"], ctx, p.g["trial"]).cpu().numpy()
    e_ref = orc.energy(w, p.sd.h0, p.sd.trial, p.sd.mctx, case["kind"])
    err = np.abs(e - e_ref) / np.abs(e_ref)
    assert err.max() < tol, err.max()


@pytest.mark.parametrize("nw,n,o,nspin", [(8, 24, 6, 1), (5, 40, 7, 2), (6, 45, 11, 2), (16, 200, 32, 1),
                                          (3, 60, 40, 1)])
def test_greens_split_planes_equal_split_of_greens(gpu_device, nw, n, o, nspin):
    """afq_overlap_greens_split: the same greens and logdet as afq_overlap_greens,
    and bf16 planes equal to split_bf16 of the greens rows bit for bit (zero
    padding columns included); o > 32 takes the block kernel and a split."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    print(f"greens split nw={nw} n={n} o={o} nspin={nspin}", flush=True)
    rng = np.random.default_rng(n + o)
    trial = torch.tensor(np.linalg.qr(rng.standard_normal((n, o)))[0], device=gpu_device)
    ws = [torch.tensor(rng.standard_normal((nw, n, o)) + 1j * rng.standard_normal((nw, n, o)),
                       device=gpu_device) for _ in range(nspin)]
    walk = ws[0] if nspin == 1 else tuple(ws)
    td = None if nspin == 1 else trial
    g0, ld0 = kernels.overlap_greens(walk, trial, td)
    g1, ld1, (hi, lo) = kernels.overlap_greens(walk, trial, td, split=True)
    assert torch.equal(g0, g1) and torch.equal(ld0, ld1)
    rh, rl = kernels.split_bf16(g0.reshape(2 * nw, -1))
    assert hi.shape == rh.shape
    assert torch.equal(hi, rh) and torch.equal(lo, rl)


def test_reconfigure_gathers_greens_planes(gpu_device):
    """After the comb the state's bf16 planes are those of its gathered greens."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    from ad_afqmc_prototype_amd.meas.rhf import make_rhf_meas_ops
    from ad_afqmc_prototype_amd.prop.afqmc import init_prop_state
    from ad_afqmc_prototype_amd.prop.blocks import reconfigure
    from ad_afqmc_prototype_amd.trial.rhf import make_rhf_trial_ops
    p = make_pair(0, 24, 6, 96, device=gpu_device)
    m_ops = make_rhf_meas_ops(p.g["sys"], precision="bf16x3")
    st = init_prop_state(sys=p.g["sys"], n_walkers=32, seed=3, ham_data=p.g["ham"],
                         trial_ops_=make_rhf_trial_ops(p.g["sys"]), trial_data=p.g["trial"], meas_ops=m_ops)
    gen = torch.Generator(device=gpu_device).manual_seed(4)
    st = st._replace(weights=torch.rand(32, dtype=torch.float64, device=gpu_device, generator=gen) + 0.1)
    st2, idx = reconfigure(st, sys=p.g["sys"], trial_data=p.g["trial"], meas_ops_=m_ops, zeta=0.3)
    rh, rl = kernels.split_bf16(st2.greens.greens.reshape(64, -1))
    assert torch.equal(st2.greens.split[0], rh) and torch.equal(st2.greens.split[1], rl)


@pytest.mark.parametrize("M,N,K", [(4096, 1000, 6400), (37, 300, 1000), (130, 257, 8), (256, 96, 2400)])
def test_bf16x3_gemm_split_k(gpu_device, M, N, K):
    """Split K: every chunk count gives the unsplit product to fp32 rounding of
    the partial sums, repeated calls are bit-identical (fixed-order
    reduction), and K shorter than one chunk per split runs as fewer chunks."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    print(f"split-K bf16x3 M={M} N={N} K={K}", flush=True)
    g = torch.Generator(device="cpu").manual_seed(M + K)
    A = torch.randn((M, K), dtype=torch.float64, generator=g).to(gpu_device)
    B = torch.randn((N, K), dtype=torch.float64, generator=g).to(gpu_device)
    ah, al = kernels.split_bf16(A)
    bh, bl = kernels.split_bf16(B)
    base = torch.empty((M, N), dtype=torch.float64, device=gpu_device)
    kernels.gemm_bf16x3_nt(ah, al, bh, bl, base, alpha=0.5, splits=1)
    ref = 0.5 * (A @ B.T)
    scale = float(ref.abs().max())
    for s in (2, 3, 4, 5):
        c1 = torch.empty_like(base)
        c2 = torch.empty_like(base)
        kernels.gemm_bf16x3_nt(ah, al, bh, bl, c1, alpha=0.5, splits=s)
        kernels.gemm_bf16x3_nt(ah, al, bh, bl, c2, alpha=0.5, splits=s)
        assert torch.equal(c1, c2)
        assert float((c1 - base).abs().max()) <= 2e-6 * scale * max(1.0, K / 1000) ** 0.5
        assert float((c1 - ref).abs().max()) <= 2e-5 * scale


@pytest.mark.parametrize("c_dtype", ["float64", "float32"])
def test_bf16x3_split_k_chunks_equal_unsplit_chunk_products(gpu_device, c_dtype):
    """The force-bias shape (M 4096, N 1000, K 6400, 4 chunks of 1600) runs its
    chunks on the 256 x 256 tiles; each chunk's fp32 partial equals the
    128 x 128 kernel's unsplit product over the chunk's K range, so the result
    equals those partials summed in fp64 in chunk order, times alpha, bit for
    bit."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    M, N, K, s, alpha = 4096, 1000, 6400, 4, 0.5
    print(f"split-K chunk partials M={M} N={N} K={K} splits={s} ({c_dtype})", flush=True)
    g = torch.Generator(device="cpu").manual_seed(11)
    A = torch.randn((M, K), dtype=torch.float64, generator=g).to(gpu_device)
    B = torch.randn((N, K), dtype=torch.float64, generator=g).to(gpu_device)
    ah, al = kernels.split_bf16(A)
    bh, bl = kernels.split_bf16(B)
    dt = getattr(torch, c_dtype)
    out = torch.empty((M, N), dtype=dt, device=gpu_device)
    kernels.gemm_bf16x3_nt(ah, al, bh, bl, out, alpha=alpha, splits=s)
    kc = 1600
    acc = None
    for z in range(s):
        part = torch.empty((M, N), dtype=torch.float32, device=gpu_device)
        sl = slice(z * kc, (z + 1) * kc)
        kernels.gemm_bf16x3_nt(ah[:, sl].contiguous(), al[:, sl].contiguous(), bh[:, sl].contiguous(),
                               bl[:, sl].contiguous(), part, alpha=1.0, splits=1)
        acc = part.double() if acc is None else acc + part.double()
    ref = (alpha * acc).to(dt)
    assert torch.equal(out, ref)


@pytest.mark.parametrize("nw,n,o,nspin", [(8, 24, 6, 1), (5, 45, 11, 2), (3, 17, 5, 1), (16, 200, 32, 1),
                                          (4, 224, 16, 2)])
def test_trotter_bf16x3_upper_block_vhs_equals_full_rows(gpu_device, nw, n, o, nspin):
    """The bf16x3 Trotter on V and E stored as upper 16 x 16 blocks equals the
    one on full (n, ldv) rows bit for bit: the same contraction fills both V
    layouts, E is exactly symmetric, and the lower-block fragments read from
    the stored transposes are the same floats the full rows hold."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    from ad_afqmc_prototype_amd.prop.chol_afqmc_ops import chol_sym_rows, minus_identity_pairs, upper_blocks
    print(f"upper-block Trotter nw={nw} n={n} o={o} nspin={nspin}", flush=True)
    rng = np.random.default_rng(n + nw)
    g = 40
    a = rng.standard_normal((g, n, n)) * 0.05
    chol = torch.tensor(0.5 * (a + a.transpose(0, 2, 1)), device=gpu_device)
    x = torch.tensor(rng.standard_normal((2 * nw, g)), device=gpu_device)
    xs = kernels.split_bf16(x)
    cs = kernels.split_bf16(chol_sym_rows(chol))
    v_full = kernels.vhs_sym(xs, cs, n).view(2, nw, n, kernels.ldv_for(n))
    v_blk = torch.empty((2 * nw, cs[0].shape[0]), dtype=torch.int32, device=gpu_device)
    kernels.gemm_bf16x3_nt(*xs, *cs, v_blk)
    v_blk = v_blk.view(2, nw, -1)
    # the pair words hold the float32 products split: hi + lo is within the
    # split of the float32 output of the same contraction
    v32 = torch.empty((2 * nw, cs[0].shape[0]), dtype=torch.float32, device=gpu_device)
    kernels.gemm_bf16x3_nt(*xs, *cs, v32)
    hv = ((v_blk.reshape(-1) & 0xFFFF) << 16).view(torch.float32).double()
    lv = (v_blk.reshape(-1) & ~0xFFFF).view(torch.float32).double()
    assert float(((hv + lv) - v32.reshape(-1).double()).abs().max()) <= 2.0 ** -16 * float(v32.abs().max())
    ldv = kernels.ldv_for(n)
    h = rng.standard_normal((n, n)) * 0.01
    e = torch.zeros((n, ldv), dtype=torch.float32, device=gpu_device)
    e[:, :n] = torch.tensor(np.eye(n) + 0.5 * (h + h.T), dtype=torch.float32, device=gpu_device)
    ws = [torch.tensor(rng.standard_normal((nw, n, o)) + 1j * rng.standard_normal((nw, n, o)),
                       device=gpu_device) for _ in range(nspin)]
    walk = ws[0] if nspin == 1 else tuple(ws)
    r_full = kernels.trotter(walk, v_full, e, 0.07, 6, arith="bf16x3")
    e_blk = minus_identity_pairs(upper_blocks(e[:, :n].double()).float(), n)
    r_blk = kernels.trotter(walk, v_blk, e_blk, 0.07, 6, arith="bf16x3")
    for a_, b_ in zip(r_full if nspin == 2 else (r_full,), r_blk if nspin == 2 else (r_blk,)):
        assert torch.equal(a_, b_)


@pytest.mark.parametrize("n,o,nspin", [(200, 32, 1), (180, 16, 2), (120, 20, 1)])
def test_trotter_bf16x3_walker_pairs_equal_per_item_launch(gpu_device, n, o, nspin):
    """With upper-block V, two items per walker (nspin * ceil(o / 16) == 2) and
    at least as many walkers as CUs, persistent workgroups run both items of a
    walker (E W of the second item and the first Taylor sum parked in the output
    rows as fp64, E - I kept from one walker to the next).  Every walker's result
    equals the per-item launch's (the same walkers, fewer of them) bit for bit,
    in both V layouts, and the fp64 step within 2e-5."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    from ad_afqmc_prototype_amd.prop.chol_afqmc_ops import chol_sym_rows, minus_identity_pairs, split_pairs, upper_blocks
    cus = torch.cuda.get_device_properties(gpu_device).multi_processor_count
    nw = cus + 37
    print(f"walker-pair Trotter nw={nw} n={n} o={o} nspin={nspin}", flush=True)
    rng = np.random.default_rng(n + o)
    dt = 0.005
    e, v, ws = _trotter_inputs(rng, nw, n, o, nspin, dt, 0.3)
    ldv = kernels.ldv_for(n)
    ep, vp = _planes(e, v, ldv)
    ep32 = torch.tensor(ep, device=gpu_device).float()
    vp32 = torch.tensor(vp, device=gpu_device).float()
    e_blk = minus_identity_pairs(upper_blocks(torch.tensor(e, device=gpu_device)).float(), n)
    v_blk = split_pairs(chol_sym_rows(vp32.reshape(2 * nw, n, ldv)[:, :, :n].double()).T.float()).view(2, nw, -1)
    wd = [torch.tensor(x, device=gpu_device) for x in ws]
    few = 16
    for lay in ("rows", "blocks"):
        vv, ee = (vp32, ep32) if lay == "rows" else (v_blk, e_blk)
        vf = vv[:, :few].contiguous()
        walk = wd[0] if nspin == 1 else tuple(wd)
        walk_few = wd[0][:few].contiguous() if nspin == 1 else tuple(x[:few].contiguous() for x in wd)
        got = kernels.trotter(walk, vv, ee, dt ** 0.5, 6, arith="bf16x3")
        ref = kernels.trotter(walk_few, vf, ee, dt ** 0.5, 6, arith="bf16x3")
        gots = got if nspin == 2 else (got,)
        refs = ref if nspin == 2 else (ref,)
        for g_, r_ in zip(gots, refs):
            assert torch.equal(g_[:few], r_), lay
    v64 = torch.tensor(vp, device=gpu_device)[:, :few].contiguous()
    e64 = torch.tensor(ep, device=gpu_device)
    walk_few = wd[0][:few].contiguous() if nspin == 1 else tuple(x[:few].contiguous() for x in wd)
    r64 = kernels.trotter(walk_few, v64, e64, dt ** 0.5, 6)
    for g_, r_ in zip(gots, r64 if nspin == 2 else (r64,)):
        g_, r_ = to_np(g_[:few]), to_np(r_)
        assert max(rel(g_[i], r_[i]) for i in range(few)) < 2e-5


@pytest.mark.parametrize("n,o,nspin", [(200, 32, 1), (180, 16, 2), (61, 17, 1)])
def test_trotter_bf16x3_separate_halves(gpu_device, n, o, nspin):
    """afq_trotter_bf16x3_sym_halves: (i) a last-half operand equal to (a copy
    of) the first gives the plain launch's result bit for bit; (ii) with no
    last half, with E^2 first, and with E^2 first and E last, the walker-pair
    launch equals the per-item launch bit for bit; (iii) two steps with the
    halves between them merged, E T2 E^2 (T1 E W), are within 4e-5 of the fp64
    pair of steps E T2 E E T1 E W per walker, and E (T1 E W) within 2e-5 of the
    fp64 first step."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    from ad_afqmc_prototype_amd.prop.chol_afqmc_ops import chol_sym_rows, minus_identity_pairs, split_pairs, upper_blocks
    cus = torch.cuda.get_device_properties(gpu_device).multi_processor_count
    nw = cus + 37
    few = 16
    print(f"separate one-body halves nw={nw} n={n} o={o} nspin={nspin}", flush=True)
    rng = np.random.default_rng(7 * n + o)
    dt = 0.005
    sq = dt ** 0.5
    e, v1, ws = _trotter_inputs(rng, nw, n, o, nspin, dt, 0.3)
    _, v2, _ = _trotter_inputs(rng, nw, n, o, nspin, dt, 0.3)
    e2 = e @ e
    e2 = 0.5 * (e2 + e2.T)
    ldv = kernels.ldv_for(n)

    def blocks(v):
        vp = torch.tensor(np.stack([v.real, v.imag]).reshape(2 * nw, n, n), device=gpu_device)
        return split_pairs(chol_sym_rows(vp.float().double()).T.float()).view(2, nw, -1).contiguous()

    v1b, v2b = blocks(v1), blocks(v2)
    e_blk = minus_identity_pairs(upper_blocks(torch.tensor(e, device=gpu_device)).float(), n)
    e2_blk = minus_identity_pairs(upper_blocks(torch.tensor(e2, device=gpu_device)).float(), n)
    wd = [torch.tensor(x, device=gpu_device) for x in ws]
    walk = wd[0] if nspin == 1 else tuple(wd)
    walk_few = wd[0][:few].contiguous() if nspin == 1 else tuple(x[:few].contiguous() for x in wd)

    def parts(x):
        return x if isinstance(x, tuple) else (x,)

    plain = kernels.trotter(walk, v1b, e_blk, sq, 6, arith="bf16x3")
    sep = kernels.trotter(walk, v1b, e_blk, sq, 6, arith="bf16x3", e_last=e_blk.clone())
    for a, b in zip(parts(plain), parts(sep)):
        assert torch.equal(a, b)
    for ef, el in ((e_blk, None), (e2_blk, None), (e2_blk, e_blk)):
        got = kernels.trotter(walk, v1b, ef, sq, 6, arith="bf16x3", e_last=el)
        ref = kernels.trotter(walk_few, v1b[:, :few].contiguous(), ef, sq, 6, arith="bf16x3", e_last=el)
        for a, b in zip(parts(got), parts(ref)):
            assert torch.equal(a[:few], b), (ef is e2_blk, el is None)

    half = kernels.trotter(walk, v1b, e_blk, sq, 6, arith="bf16x3", e_last=None)
    two = kernels.trotter(half, v2b, e2_blk, sq, 6, arith="bf16x3", e_last=e_blk)
    _, vp1 = _planes(e, v1, ldv)
    ep, vp2 = _planes(e, v2, ldv)
    e64 = torch.tensor(ep, device=gpu_device)
    one64 = kernels.trotter(walk, torch.tensor(vp1, device=gpu_device), e64, sq, 6)
    two64 = kernels.trotter(one64, torch.tensor(vp2, device=gpu_device), e64, sq, 6)
    et = torch.tensor(e, device=gpu_device).to(torch.complex128)
    for h_, o_, t_, r_ in zip(parts(half), parts(one64), parts(two), parts(two64)):
        h_ = to_np(et @ h_)
        o_, t_, r_ = to_np(o_), to_np(t_), to_np(r_)
        assert max(rel(h_[i], o_[i]) for i in range(nw)) < 2e-5
        assert max(rel(t_[i], r_[i]) for i in range(nw)) < 4e-5


@pytest.mark.parametrize("kind,n,o,g,scale", [("restricted", 200, 32, 1000, 0.03),
                                              ("unrestricted", 120, 16, 600, 0.05)])
def test_block_with_merged_one_body_halves(gpu_device, kind, n, o, g, scale):
    """bf16x3 blocks (10 steps, QR after 5) whose steps leave their final
    exp(-dt h1_eff / 2) to the next step, which applies exp(-dt h1_eff) once
    (the walker-pair launch: more walkers than CUs), against the same blocks
    with both halves in every step, on the device stream's fields and comb
    uniforms: identical comb picks; walkers, weights and log-overlaps within
    1e-5; block energies within 2e-5 Ha.  A deferred state is refused by QR,
    measurement and comb, and the fp64 propagator never defers."""
    import dataclasses
    import torch
    from ad_afqmc_prototype_amd.meas.rhf import make_rhf_meas_ops
    from ad_afqmc_prototype_amd.prop.afqmc import afqmc_step, can_merge_halves, init_prop_state
    from ad_afqmc_prototype_amd.prop.blocks import afqmc_block, measure, orthonormalize_state, reconfigure
    from ad_afqmc_prototype_amd.prop.chol_afqmc_ops import make_chol_afqmc_ops
    cus = torch.cuda.get_device_properties(gpu_device).multi_processor_count
    nw, dt = cus + 24, 0.005
    print(f"merged one-body halves {kind} n={n} o={o} nw={nw}", flush=True)
    p = make_pair(3, n, o, g, uhf=kind == "unrestricted", kind=kind, dt=dt, chol_scale=scale, device=gpu_device)
    m_ops = make_rhf_meas_ops(p.g["sys"], precision="bf16x3")
    p_ops = make_chol_afqmc_ops(p.g["ham"], kind, precision="bf16x3")
    p_ctx = p_ops.build_prop_ctx(p.g["t_ops"].get_rdm1(p.g["trial"]), dt)
    m_ctx = m_ops.build_meas_ctx(p.g["ham"], p.g["trial"])
    assert can_merge_halves(p_ops, m_ops, m_ctx, p.g["trial"])
    assert not can_merge_halves(p.g["p_ops"], p.g["m_ops"], p.g["m_ctx"], p.g["trial"])
    st0 = init_prop_state(sys=p.g["sys"], n_walkers=nw, seed=5, ham_data=p.g["ham"], trial_ops_=p.g["t_ops"],
                          trial_data=p.g["trial"], meas_ops=m_ops)
    base = dataclasses.replace(p.g["params"], n_prop_steps=10, n_ortho=5)
    kw = dict(ham_data=p.g["ham"], trial_data=p.g["trial"], meas_ops=m_ops, prop_ops=p_ops, prop_ctx=p_ctx,
              meas_ctx=m_ctx)
    res = {}
    for merge in (True, False):
        params = dataclasses.replace(base, merge_one_body_halves=merge)
        st, es, picks = st0, [], []
        for _ in range(2):
            st, obs, idx = afqmc_block(st, sys=p.g["sys"], params=params, return_indices=True, **kw)
            es.append(obs.scalars["energy"].item())
            picks.append(idx)
        assert not st.deferred
        res[merge] = (st, es, picks)
    (a, ea, pa), (b, eb, pb) = res[True], res[False]
    for x, y in zip(pa, pb):
        assert torch.equal(x, y)
    wa = a.walkers if isinstance(a.walkers, tuple) else (a.walkers,)
    wb = b.walkers if isinstance(b.walkers, tuple) else (b.walkers,)
    for x, y in zip(wa, wb):
        x, y = to_np(x), to_np(y)
        assert max(rel(x[i], y[i]) for i in range(nw)) < 1e-5
    assert rel(to_np(a.weights), to_np(b.weights)) < 1e-5
    assert float((a.log_overlaps - b.log_overlaps).abs().max()) < 1e-5
    assert max(abs(x - y) for x, y in zip(ea, eb)) < 2e-5, (ea, eb)
    assert all(np.isfinite(ea)) and abs(ea[1] - ea[0]) > 1e-6

    d = afqmc_step(st0, params=base, defer_final=True, **kw)
    assert d.deferred
    with pytest.raises(ValueError):
        orthonormalize_state(d, sys=p.g["sys"], trial_data=p.g["trial"], meas_ops_=m_ops)
    with pytest.raises(ValueError):
        measure(d, params=base, ham_data=p.g["ham"], trial_data=p.g["trial"], meas_ops=m_ops, meas_ctx=m_ctx)
    with pytest.raises(ValueError):
        reconfigure(d, sys=p.g["sys"], trial_data=p.g["trial"], meas_ops_=m_ops, zeta=0.5)
    # the next step takes it and returns a whole state
    assert not afqmc_step(d, params=base, **kw).deferred


@pytest.mark.parametrize("nw,naux", [(8, 96), (5, 130), (33, 1000)])
def test_field_shifts_split_planes_equal_split_of_xbar(gpu_device, nw, naux):
    """afq_field_shifts_split: the same xbar, shift and force-bias terms as
    afq_field_shifts, and bf16 planes equal to split_bf16(xbar rows) bit for
    bit (zero padding columns included)."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    print(f"field shifts split nw={nw} naux={naux}", flush=True)
    g = torch.Generator(device="cpu").manual_seed(naux)
    x = torch.randn((nw, naux), dtype=torch.float64, generator=g).to(gpu_device)
    fb = torch.randn((2, nw, naux), dtype=torch.float64, generator=g).to(gpu_device)
    mf = torch.randn((naux,), dtype=torch.complex128, generator=g).to(gpu_device)
    a = kernels.field_shifts(x, fb, mf, 0.07)
    b = kernels.field_shifts(x, fb, mf, 0.07, split=True)
    for u, v in zip(a, b[:3]):
        assert torch.equal(u, v)
    rh, rl = kernels.split_bf16(a[0].reshape(2 * nw, naux))
    assert torch.equal(b[3][0], rh) and torch.equal(b[3][1], rl)


@pytest.mark.parametrize("case", [
    dict(seed=2, n=200, o=32, uhf=False, kind="restricted", nw=67),
    dict(seed=1, n=120, o=20, uhf=True, kind="unrestricted", nw=33),
    dict(seed=0, n=24, o=6, uhf=False, kind="restricted", nw=40),
    dict(seed=5, n=45, o=16, uhf=False, kind="restricted", nw=9),
    dict(seed=6, n=61, o=17, uhf=True, kind="unrestricted", nw=5),
])
def test_greens_bf16x3_split_only(gpu_device, case):
    """afq_overlap_greens_bf16x3 (Gauss-Jordan inverse, rows as a split-bf16
    product) against the fp64 LU kernel: log det to 1e-12, the bf16 planes'
    value hi + lo within 1e-4 of the fp64 rows (relative, norm-wise per walker),
    padding columns zero; walkers near the trial and far from it."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    from ._pairs import random_walkers
    s = syn.make_system(case["seed"], case["n"], case["o"], 10, uhf=case["uhf"])
    rng = np.random.default_rng(case["seed"])
    for spread in (0.3, 3.0):
        w = random_walkers(rng, case["kind"], case["nw"], case["n"], case["o"], s.mo_up, s.mo_dn, spread=spread)
        wd = to_dev(w, gpu_device)
        cu = torch.tensor(s.mo_up, device=gpu_device)
        cd = torch.tensor(s.mo_dn, device=gpu_device)
        g64, ld64, (h64, l64) = kernels.overlap_greens(wd, cu, cd, split=True)
        ld, (hi, lo) = kernels.overlap_greens_bf16x3(wd, cu, cd)
        ref = ld64.cpu().numpy()
        got = ld.cpu().numpy()
        np.testing.assert_allclose(got.real, ref.real, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(np.exp(1j * got.imag), np.exp(1j * ref.imag), atol=1e-12)

        def value(h, l):
            hv = (h.cpu().to(torch.int32) & 0xFFFF) << 16
            lv = (l.cpu().to(torch.int32) & 0xFFFF) << 16
            return hv.view(torch.float32).double() + lv.view(torch.float32).double()

        nw = case["nw"]
        k = g64.shape[2]
        v = value(hi, lo).numpy()
        assert np.all(v[:, k:] == 0.0)
        g = g64.cpu().numpy().reshape(2 * nw, k)
        err = np.linalg.norm(v[:, :k] - g, axis=1) / np.linalg.norm(g, axis=1)
        assert err.max() < 1e-4, (spread, err.max())
