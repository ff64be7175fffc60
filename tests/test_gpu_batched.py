"""Grouped launches over many independent systems (prop/batched.py):

* the grouped fp32 GEMM equals the single-problem GEMM bit for bit, in every
  tile shape;
* a batch of uneven systems (different n, o, naux, walker counts) equals each
  system run alone with afqmc_step / orthonormalize_state / measure /
  reconfigure in fp32 precision: walkers, weights, log-overlaps, population
  shifts, keys and block energies, bit for bit (the same kernel bodies);
* the block end in two grouped launches equals the per-system launches bit
  for bit, in both precisions;
* a block captured as a HIP graph and replayed equals the eager batch;
* the grouped bf16x3 GEMM and operand split equal their single-problem
  entry points bit for bit, and a batch in bf16x3 precision (contractions and
  Trotter) stays within the split-bf16 error of the fp32 batch.
"""
import numpy as np
import pytest

from ad_afqmc_prototype_amd import synthetic as syn

pytestmark = pytest.mark.gpu

SYSTEMS = [  # seed, n, o, naux, walkers: three Trotter buckets (2, 3, 4 row tiles),
    (101, 24, 4, 100, 16),   # two energy buckets (o <= 16, o > 16)
    (102, 40, 6, 190, 40),
    (103, 33, 9, 150, 24),
    (104, 56, 12, 260, 64),
    (105, 48, 20, 220, 24),
]
DT, QR, SR = 0.005, 2, 4


def _members(device, systems=SYSTEMS, scale=0.25):
    from ad_afqmc_prototype_amd.ham.chol import ham_chol
    from ad_afqmc_prototype_amd.prop.batched import batch_system
    from ad_afqmc_prototype_amd.trial.rhf import rhf_trial
    out = []
    for seed, n, o, g, nw in systems:
        s = syn.make_system(seed, n, o, g, chol_scale=scale)
        ham = ham_chol.from_arrays(s.h0, s.h1, s.chol, device=device)
        out.append(batch_system(ham=ham, trial=rhf_trial(s.mo_up), n_walkers=nw, seed=seed))
    return out


def _alone(member, n_steps, qr=QR, sr=SR):
    """The system run by itself through the single-system step and block functions."""
    from ad_afqmc_prototype_amd.core.system import system
    from ad_afqmc_prototype_amd.meas.rhf import make_rhf_meas_ops
    from ad_afqmc_prototype_amd.prop.afqmc import afqmc_step, init_prop_state
    from ad_afqmc_prototype_amd.prop.blocks import measure, orthonormalize_state, reconfigure
    from ad_afqmc_prototype_amd.prop.chol_afqmc_ops import make_chol_afqmc_ops
    from ad_afqmc_prototype_amd.prop.types import afqmc_params
    from ad_afqmc_prototype_amd.trial.rhf import make_rhf_trial_ops
    ham, trial = member.ham, member.trial
    n, o = trial.norb, trial.nocc
    sys_ = system(n, (o, o), "restricted")
    t_ops = make_rhf_trial_ops(sys_)
    m_ops = make_rhf_meas_ops(sys_, precision="fp32")
    p_ops = make_chol_afqmc_ops(ham, "restricted", precision="fp32")
    params = afqmc_params(dt=DT, n_walkers=member.n_walkers)
    p_ctx = p_ops.build_prop_ctx(t_ops.get_rdm1(trial), DT)
    m_ctx = m_ops.build_meas_ctx(ham, trial)
    st = init_prop_state(sys=sys_, n_walkers=member.n_walkers, seed=member.seed, ham_data=ham,
                         trial_ops_=t_ops, trial_data=trial, meas_ops=m_ops)
    blocks = []
    for t in range(1, n_steps + 1):
        st = afqmc_step(st, params=params, ham_data=ham, trial_data=trial, meas_ops=m_ops,
                        prop_ops=p_ops, prop_ctx=p_ctx, meas_ctx=m_ctx)
        if t % sr == 0:
            st = orthonormalize_state(st, sys=sys_, trial_data=trial, meas_ops_=m_ops)
            st, sc, _ = measure(st, params=params, ham_data=ham, trial_data=trial, meas_ops=m_ops,
                                meas_ctx=m_ctx)
            blocks.append((float(sc["energy"]), float(sc["weight"])))
            st, _ = reconfigure(st, sys=sys_, trial_data=trial, meas_ops_=m_ops)
        elif t % qr == 0:
            st = orthonormalize_state(st, sys=sys_, trial_data=trial, meas_ops_=m_ops)
    return st, blocks


def test_grouped_gemm_matches_single(gpu_device):
    import torch
    from ad_afqmc_prototype_amd import _lib, kernels
    from ad_afqmc_prototype_amd.prop import batched as B
    g = torch.Generator(device="cpu").manual_seed(3)
    dims = [(1, 1, 1), (37, 70, 19), (130, 129, 260), (64, 4096, 96), (300, 17, 1001)]
    ops = []
    for i, (M, N, K) in enumerate(dims):
        a = torch.randn((M, K), generator=g).to(gpu_device, torch.float32)
        b = torch.randn((N, K), generator=g).to(gpu_device, torch.float32)
        c64 = i % 2 == 0
        ref = torch.empty((M, N), dtype=torch.float64 if c64 else torch.float32, device=gpu_device)
        kernels.sgemm_nt(a, b, ref, alpha=0.75)
        ops.append((a, b, ref, c64))
    for shape in B.GEMM_SHAPES:
        bm, bn = B.gemm_tile_shape(shape)
        arr = (B._gemm_problem * len(ops))()
        outs, start = [], 0
        for p, (a, b, ref, c64) in zip(arr, ops):
            out = torch.full_like(ref, float("nan"))
            outs.append(out)
            M, K = a.shape
            N = b.shape[0]
            p.M, p.N, p.K, p.tile_start = M, N, K, start
            p.lda, p.ldb, p.ldc = K, K, N
            p.A, p.B, p.C, p.alpha, p.c_fp64 = a.data_ptr(), b.data_ptr(), out.data_ptr(), 0.75, int(c64)
            start += -(-M // bm) * -(-N // bn)
        tab = B._upload(arr).to(gpu_device)
        _lib.call("afq_sgemm_nt_grouped", len(ops), tab.data_ptr(), start, shape,
                  torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        for out, (_, _, ref, _) in zip(outs, ops):
            assert torch.equal(out, ref), f"shape {shape}: grouped GEMM differs"


def _compare_system(batch, i, st, blocks_alone, blocks_batch):
    import torch
    got = batch.system_state(i)
    assert torch.equal(got["walkers"], st.walkers), f"system {i}: walkers"
    assert torch.equal(got["weights"], st.weights), f"system {i}: weights"
    assert torch.equal(got["log_overlaps"], st.log_overlaps), f"system {i}: log-overlaps"
    assert torch.equal(got["pop"], st.pop_control_ene_shift), f"system {i}: pop shift"
    assert torch.equal(got["key"], st.rng_key), f"system {i}: key"
    assert blocks_batch == blocks_alone, f"system {i}: block energies / weights"


@pytest.mark.parametrize("shapes", [(0, 0), (3, 1), (2, 3)])
def test_batch_matches_systems_run_alone(gpu_device, shapes):
    from ad_afqmc_prototype_amd.prop.batched import system_batch
    print(f"uneven batch of {len(SYSTEMS)} systems, GEMM shapes {shapes}", flush=True)
    members = _members(gpu_device)
    batch = system_batch(members, dt=DT, qr_every=QR, sr_every=SR, gemm_shapes=shapes, precision="fp32")
    n_steps = 2 * SR
    per_block = []
    for _ in range(n_steps // SR):
        batch.advance(SR)
        per_block.append(batch.block_stats.cpu().numpy().copy())
    for i, m in enumerate(members):
        st, blocks = _alone(m, n_steps)
        bb = [(float(b[i, 0]), float(b[i, 1])) for b in per_block]
        _compare_system(batch, i, st, blocks, bb)
    # the run is live: weights moved and the comb copied walkers
    w = batch.system_state(3)["weights"]
    assert float(w.min()) > 0.0 and np.isfinite(per_block[-1]).all()


@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
def test_grouped_block_end_matches_per_system(gpu_device, precision):
    """The block end in two grouped launches (afq_block_end_grouped,
    afq_reconfigure_gather_grouped) against the per-system launches: after
    each of two blocks the block energies and weight sums, walkers, weights,
    log-overlaps, population shifts, keys, the Green's-function rows (and in
    bf16x3 their split planes) are equal bit for bit."""
    import torch
    from ad_afqmc_prototype_amd.prop.batched import system_batch
    print(f"grouped vs per-system block end ({precision})", flush=True)
    a = system_batch(_members(gpu_device), dt=DT, qr_every=QR, sr_every=SR, precision=precision)
    b = system_batch(_members(gpu_device), dt=DT, qr_every=QR, sr_every=SR, precision=precision,
                     grouped_block_end=False)
    assert a.grouped_block_end and not b.grouped_block_end
    for blk in range(2):
        a.advance(SR)
        b.advance(SR)
        torch.cuda.synchronize()
        assert torch.equal(a.block_stats, b.block_stats), f"block {blk}: block stats"
        for i in range(len(SYSTEMS)):
            x, y = a.system_state(i), b.system_state(i)
            for k in x:
                assert torch.equal(x[k], y[k]), f"block {blk}, system {i}: {k}"
        assert torch.equal(a.greens.buf, b.greens.buf), f"block {blk}: Green's-function rows"
        if precision == "bf16x3":
            for p in range(2):
                assert torch.equal(a.greens_hl[p].buf, b.greens_hl[p].buf), f"block {blk}: split plane {p}"
    # the run is live and the comb ran (it leaves every weight of a system equal)
    assert float(a.block_stats[:, 1].min()) > 0.0
    w = [a.system_state(i)["weights"] for i in range(len(SYSTEMS))]
    assert all(torch.all(x == x[0]) for x in w)     # the comb leaves equal weights


@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
def test_batch_graph_replay_matches_eager(gpu_device, precision):
    import torch
    from ad_afqmc_prototype_amd.prop.batched import captured_batch_block, system_batch
    print(f"captured batch block vs eager ({precision})", flush=True)
    eager = system_batch(_members(gpu_device), dt=DT, qr_every=QR, sr_every=SR, precision=precision)
    graph_batch = system_batch(_members(gpu_device), dt=DT, qr_every=QR, sr_every=SR, precision=precision)
    cap = captured_batch_block(graph_batch, warmup_blocks=1)   # advances one block eagerly
    eager.advance(SR)
    for rep in range(3):
        stats = cap.replay().clone()
        eager.advance(SR)
        torch.cuda.synchronize()
        assert torch.equal(stats, eager.block_stats), f"replay {rep}: block stats"
        for i in range(len(SYSTEMS)):
            a, b = graph_batch.system_state(i), eager.system_state(i)
            for k in a:
                assert torch.equal(a[k], b[k]), f"replay {rep}, system {i}: {k}"
        # eager work on the program's own stream between replays
        torch.zeros(1 << 20, device=gpu_device).sum().item()


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_tune_chooses_shapes_and_caches(gpu_device, tmp_path, precision):
    from ad_afqmc_prototype_amd.prop.batched import BF16X3_SHAPES, GEMM_SHAPES, system_batch
    shapes = BF16X3_SHAPES if precision == "bf16x3" else GEMM_SHAPES
    batch = system_batch(_members(gpu_device), dt=DT, qr_every=QR, sr_every=SR, precision=precision)
    path = str(tmp_path / "tiles.json")
    entry = batch.tune(reps=2, cache_path=path)
    assert tuple(entry["shapes"]) == batch.gemm_shapes
    assert all(s in shapes for s in batch.gemm_shapes)
    assert all(set(entry["ms"][w]) == {str(s) for s in shapes} for w in ("fb", "vhs"))
    again = system_batch(_members(gpu_device), dt=DT, qr_every=QR, sr_every=SR, precision=precision)
    assert again.tune(cache_path=path)["shapes"] == entry["shapes"]
    # every shape advances the batch identically (same sums, same order)
    import torch
    runs = []
    for fb_shape, vhs_shape in [(shapes[0], shapes[0]), (shapes[-1], shapes[-1])]:
        b = system_batch(_members(gpu_device), dt=DT, qr_every=QR, sr_every=SR, precision=precision,
                         gemm_shapes=(fb_shape, vhs_shape))
        b.advance(SR)
        torch.cuda.synchronize()
        runs.append([b.system_state(i) for i in range(len(SYSTEMS))] + [b.block_stats.clone()])
    for x, y in zip(runs[0][:-1], runs[1][:-1]):
        assert torch.equal(x["walkers"], y["walkers"]) and torch.equal(x["weights"], y["weights"])
    assert torch.equal(runs[0][-1], runs[1][-1])


@pytest.mark.parametrize("shape", [0, 1, 2])
def test_grouped_bf16x3_gemm_and_split_match_single(gpu_device, shape):
    """Every grouped bf16x3 tile shape (128 x 128 through registers, 128 x 256
    and 256 x 256 through buffer_load ... lds) equals afq_gemm_bf16x3_nt on
    each problem, bit for bit, in all three output forms (fp32, fp64, pairs)."""
    import torch
    from ad_afqmc_prototype_amd import _lib, kernels
    from ad_afqmc_prototype_amd.prop import batched as B
    g = torch.Generator(device="cpu").manual_seed(5)
    dims = [(1, 1, 1), (37, 70, 19), (130, 129, 260), (64, 700, 96), (300, 17, 1001), (513, 300, 64)]
    st = torch.cuda.current_stream().cuda_stream
    ops, splits = [], (B._split_problem * len(dims))()
    sources = []  # the split reads them after the loop: keep them allocated
    biggest = 0
    out_dtype = {0: torch.float32, 1: torch.float64, 2: torch.int32}
    for i, (M, N, K) in enumerate(dims):
        a = torch.randn((M, K), generator=g, dtype=torch.float64).to(gpu_device)
        b = torch.randn((N, K), generator=g, dtype=torch.float64).to(gpu_device)
        sources.append(a)
        ah, al = kernels.split_bf16(a)
        bh, bl = kernels.split_bf16(b)
        kp = ah.shape[1]
        # the grouped split of a into fresh planes
        gh, gl = torch.full_like(ah, -1), torch.full_like(al, -1)
        q = splits[i]
        q.src, q.hi, q.lo = a.data_ptr(), gh.data_ptr(), gl.data_ptr()
        q.rows, q.cols, q.ld_src, q.ld_dst = M, K, K, kp
        biggest = max(biggest, M * kp)
        mode = i % 3
        ref = torch.empty((M, N), dtype=out_dtype[mode], device=gpu_device)
        _lib.call("afq_gemm_bf16x3_nt", M, N, kp, 0.75, ah.data_ptr(), al.data_ptr(), kp, bh.data_ptr(),
                  bl.data_ptr(), kp, ref.data_ptr(), N, mode, st)
        ops.append((ah, al, bh, bl, gh, gl, ref, mode))
    stab = B._upload(splits).to(gpu_device)
    _lib.call("afq_split_bf16_grouped", len(dims), stab.data_ptr(), biggest, st)
    bm, bn = B.gemm_tile_shape(shape, "bf16x3")
    assert (bm, bn) == [(128, 128), (128, 256), (256, 256)][shape]
    arr = (B._gemm_problem_bf16x3 * len(ops))()
    outs, start = [], 0
    for p, (ah, al, bh, bl, gh, gl, ref, mode) in zip(arr, ops):
        out = torch.full_like(ref, -7)
        outs.append(out)
        M, kp = ah.shape
        N = bh.shape[0]
        p.M, p.N, p.K, p.tile_start = M, N, kp, start
        p.lda, p.ldb, p.ldc = kp, kp, N
        p.A_hi, p.A_lo, p.B_hi, p.B_lo = gh.data_ptr(), gl.data_ptr(), bh.data_ptr(), bl.data_ptr()
        p.C, p.alpha, p.c_fp64 = out.data_ptr(), 0.75, mode
        start += -(-M // bm) * -(-N // bn)
    tab = B._upload(arr).to(gpu_device)
    _lib.call("afq_gemm_bf16x3_nt_grouped", len(ops), tab.data_ptr(), start, shape, st)
    torch.cuda.synchronize()
    for i, (out, (ah, al, _, _, gh, gl, ref, _)) in enumerate(zip(outs, ops)):
        assert torch.equal(gh, ah) and torch.equal(gl, al), f"problem {i}: grouped split differs"
        assert torch.equal(out, ref), f"problem {i}: grouped bf16x3 GEMM (shape {shape}) differs"


@pytest.mark.parametrize("precision,shape", [("fp32", 0), ("fp32", 3), ("bf16x3", 0), ("bf16x3", 2)])
def test_grouped_lower_b_matches_single(gpu_device, precision, shape):
    """Grouped problems flagged b_lower (B zero past its diagonal, column tiles
    stopping there in zigzag order) equal afq_sgemm_nt_lower /
    afq_gemm_bf16x3_nt_lower on each problem bit for bit, and problems without
    the flag in the same launch equal the full GEMM."""
    import torch
    from ad_afqmc_prototype_amd import _lib, kernels
    from ad_afqmc_prototype_amd.prop import batched as B
    print(f"grouped lower-B GEMM {precision} shape {shape}", flush=True)
    g = torch.Generator(device="cpu").manual_seed(17)
    dims = [(96, 120, 120), (300, 264, 264), (37, 40, 40), (130, 520, 520), (64, 200, 96)]
    st = torch.cuda.current_stream().cuda_stream
    bm, bn = B.gemm_tile_shape(shape, precision)
    bf = precision == "bf16x3"
    arr = ((B._gemm_problem_bf16x3 if bf else B._gemm_problem) * len(dims))()
    keep, outs, refs, start = [], [], [], 0
    for i, (p, (M, N, K)) in enumerate(zip(arr, dims)):
        lower = i != len(dims) - 1
        a = torch.randn((M, K), generator=g, dtype=torch.float64).to(gpu_device)
        b = torch.randn((N, K), generator=g, dtype=torch.float64)
        b = (torch.tril(b) if lower else b).to(gpu_device)
        ref = torch.empty((M, N), dtype=torch.float64, device=gpu_device)
        out = torch.full_like(ref, float("nan"))
        if bf:
            (ah, al), (bh, bl) = kernels.split_bf16(a), kernels.split_bf16(b)
            kp = ah.shape[1]
            if lower:
                kernels.gemm_nt_lower((ah, al), (bh, bl), ref, alpha=2.0)
            else:
                kernels.gemm_bf16x3_nt(ah, al, bh, bl, ref, alpha=2.0, splits=1)
            p.A_hi, p.A_lo, p.B_hi, p.B_lo = ah.data_ptr(), al.data_ptr(), bh.data_ptr(), bl.data_ptr()
            keep += [ah, al, bh, bl]
        else:
            a32, b32 = kernels.convert_f32(a), kernels.convert_f32(b)
            kp = K
            if lower:
                kernels.gemm_nt_lower(a32, b32, ref, alpha=2.0)
            else:
                kernels.sgemm_nt(a32, b32, ref, alpha=2.0)
            p.A, p.B = a32.data_ptr(), b32.data_ptr()
            keep += [a32, b32]
        p.M, p.N, p.K, p.tile_start = M, N, kp, start
        p.lda, p.ldb, p.ldc = kp, kp, N
        p.C, p.alpha, p.c_fp64, p.b_lower = out.data_ptr(), 2.0, 1, int(lower)
        start += -(-M // bm) * -(-N // bn)
        outs.append(out)
        refs.append(ref)
    tab = B._upload(arr).to(gpu_device)
    entry = "afq_gemm_bf16x3_nt_grouped" if bf else "afq_sgemm_nt_grouped"
    _lib.call(entry, len(dims), tab.data_ptr(), start, shape, st)
    torch.cuda.synchronize()
    for i, (out, ref) in enumerate(zip(outs, refs)):
        assert torch.equal(out, ref), f"problem {i}: grouped GEMM differs"


def test_grouped_bf16x3_energy_matches_single(gpu_device):
    """The grouped bf16x3 energy (one launch per occupied-tile bucket, two
    walkers per workgroup where they fit) equals afq_energy_bf16x3 on each
    system alone, bit for bit, on the batch's Green's functions after two steps."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    from ad_afqmc_prototype_amd.prop.batched import system_batch
    batch = system_batch(_members(gpu_device), dt=DT, qr_every=QR, sr_every=SR, precision="bf16x3")
    batch.advance(2)
    batch._orthonormalize()
    st = kernels._stream(batch.desc[0])
    assert {bk["nt"] for bk in batch.energy_buckets} == {1, 2}
    for bk in batch.energy_buckets:
        t = bk["table"]
        kernels._call("afq_energy_bf16x3_grouped", t, bk["count"], t.data_ptr(), bk["blocks"], bk["max_n"],
                      bk["nt"], bk["wpb"], st)
    torch.cuda.synchronize()
    for i, m in enumerate(batch.members):
        nw, n, o = m["nw"], m["n"], m["o"]
        g = batch.greens.part(i, (2, nw, o * n))
        ref = kernels.energy(g, nw, n, o, 1, (m["rot_pad"][0], m["rot_pad"][1]), m["rot_h1"], m["h0"])
        got = batch.eloc.part(i)
        assert torch.equal(got, ref), f"system {i}: grouped bf16x3 energy differs"
        assert bool(torch.isfinite(ref.real).all())


def test_grouped_bf16x3_greens_matches_single(gpu_device):
    """The grouped Gauss-Jordan Green's kernel of the bf16x3 step (force-bias
    planes and log det, no fp64 rows) equals afq_overlap_greens_bf16x3 on each
    system alone, bit for bit, after one step."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    from ad_afqmc_prototype_amd.prop.batched import system_batch
    batch = system_batch(_members(gpu_device), dt=DT, qr_every=QR, sr_every=SR, precision="bf16x3")
    batch.step()  # walkers -> set B, planes and log det of the new walkers
    torch.cuda.synchronize()
    assert {bk["max_o"] <= 16 for bk in batch.greens_split_buckets} == {True, False}
    for i, m in enumerate(batch.members):
        nw, n, o = m["nw"], m["n"], m["o"]
        ldk = kernels.pad8(o * n)
        ref_ld, (ref_hi, ref_lo) = kernels.overlap_greens_bf16x3(batch.walkers[1].part(i, (nw, n, o)),
                                                                  batch.trials.part(i, (n, o)))
        assert torch.equal(batch.logdet.part(i), ref_ld[:, 0]), f"system {i}: log det differs"
        assert torch.equal(batch.greens_hl[0].part(i, (2 * nw, ldk)), ref_hi), f"system {i}: hi plane"
        assert torch.equal(batch.greens_hl[1].part(i, (2 * nw, ldk)), ref_lo), f"system {i}: lo plane"


def test_batch_bf16x3_close_to_fp32(gpu_device):
    """Same fields, same start: after two steps (one QR) the walkers and
    weights of the bf16x3 batch are within the split-bf16 error (~2^-16
    relative per contraction) of the fp32 batch's, and so is the first block
    energy (measured before any comb)."""
    import torch
    from ad_afqmc_prototype_amd.prop.batched import system_batch
    print("bf16x3 batch vs fp32 batch", flush=True)
    a = system_batch(_members(gpu_device), dt=DT, qr_every=QR, sr_every=SR, precision="bf16x3")
    b = system_batch(_members(gpu_device), dt=DT, qr_every=QR, sr_every=SR, precision="fp32")
    a.advance(2)
    b.advance(2)
    torch.cuda.synchronize()
    for i in range(len(SYSTEMS)):
        x, y = a.system_state(i), b.system_state(i)
        scale = float(y["walkers"].abs().max())
        dw = float((x["walkers"] - y["walkers"]).abs().max()) / scale
        assert dw < 1e-4, f"system {i}: walkers differ by {dw:.2e} (relative to max |w|)"
        assert torch.allclose(x["weights"], y["weights"], rtol=1e-4, atol=0), f"system {i}: weights"
        assert torch.equal(x["key"], y["key"]), f"system {i}: key"
    a.advance(SR - 2)
    b.advance(SR - 2)
    ea, eb = a.block_stats.cpu().numpy(), b.block_stats.cpu().numpy()
    assert np.isfinite(ea).all()
    np.testing.assert_allclose(ea[:, 0], eb[:, 0], rtol=1e-5, atol=0)


def test_batch_bf16x3_merged_halves_match_unmerged(gpu_device):
    """The bf16x3 batch whose steps leave their last one-body half to the next
    step (the default: E^2 applied once, E C trial and C^T L E force bias for
    the walkers in between) against the same batch applying both halves in
    every step: after one block (QR and measurement + comb) the block energies
    agree to 2e-5 Ha, walkers to 1e-5 of max |w| and weights to 1e-5; a
    deferred batch refuses QR and state reads until a step applies its last
    half."""
    import torch
    from ad_afqmc_prototype_amd.prop.batched import system_batch
    print("merged vs unmerged one-body halves, bf16x3 batch", flush=True)
    a = system_batch(_members(gpu_device), dt=DT, qr_every=QR, sr_every=SR, precision="bf16x3")
    b = system_batch(_members(gpu_device), dt=DT, qr_every=QR, sr_every=SR, precision="bf16x3",
                     merge_one_body_halves=False)
    assert a.merge and not b.merge
    a.advance(SR - 2)
    b.advance(SR - 2)
    for i in range(len(SYSTEMS)):
        x, y = a.system_state(i), b.system_state(i)
        scale = float(y["walkers"].abs().max())
        assert float((x["walkers"] - y["walkers"]).abs().max()) / scale < 1e-5, f"system {i}: walkers"
        assert torch.allclose(x["weights"], y["weights"], rtol=1e-5, atol=0), f"system {i}: weights"
    a.advance(2)
    b.advance(2)
    ea, eb = a.block_stats.cpu().numpy(), b.block_stats.cpu().numpy()
    assert np.isfinite(ea).all()
    assert np.abs(ea[:, 0] - eb[:, 0]).max() < 2e-5, (ea[:, 0], eb[:, 0])
    a.step(last_half=False)
    assert a.deferred
    with pytest.raises(RuntimeError):
        a.system_state(0)
    a.step()
    assert not a.deferred
    a.system_state(0)


# --------------------------------------------------------------------------
# the 64-system workload: seeds 100..163, N ~ U[40,160], nocc ~ U[6,24],
# naux = round(N U[4,6]), walkers ~ U[64,512]; dt 0.005, QR every 10,
# measurement + comb every 20, Cholesky scale 0.05
# --------------------------------------------------------------------------

def _cfg4_members(device):
    specs = syn.uneven_batch_specs(64)
    return specs, _members(device, [(sp["seed"], sp["norb"], sp["nocc"], sp["naux"], sp["n_walkers"])
                                    for sp in specs], scale=0.05)


def _oracle_block(sp, scale=0.05, dt=0.005, qr=10, sr=20):
    """One block of system ``sp`` on the CPU oracle, with the fields and the
    comb uniform of the system's own Threefry stream (split per step, normals;
    split, uniform at the block end)."""
    from oracle import afqmc_oracle as orc
    from oracle import threefry as tf
    s = syn.make_system(sp["seed"], sp["norb"], sp["nocc"], sp["naux"], chol_scale=scale)
    sd = orc.make_system_data(s.h0, s.h1, s.chol, (s.mo_up, s.mo_dn), "restricted", dt)
    key = tf.prng_key(sp["seed"])
    fields = []
    for _ in range(sr):
        key, sub = tf.split(key)
        fields.append(tf.normal(sub, (sp["n_walkers"], sp["naux"])))
    key, sub = tf.split(key)
    zeta = float(tf.uniform(sub))
    st0 = orc.init_state(sd, sp["n_walkers"], overlap_mode="log")
    st, blocks = orc.run_cadenced(st0, sd, orc.params(dt=dt), n_steps=sr, qr_every=qr, sr_every=sr,
                                  fields_fn=lambda t: fields[t], zeta_fn=lambda k: zeta,
                                  overlap_mode="log")
    return st, blocks[0]


def test_cfg4_batch_block(gpu_device):
    """One block of the 64-system workload in fp32: every system bit for bit
    equal to its single-system fp32 run; the smallest, the largest-n and the
    largest-nocc system against the fp64 oracle (block energy within 0.1 mHa,
    weight sum to 1e-5, walker projectors to 1e-4); the bf16x3 batch's block
    energies within 1e-5 relative of the fp32 batch's."""
    import torch
    from ad_afqmc_prototype_amd.prop.batched import system_batch
    specs, members = _cfg4_members(gpu_device)
    print(f"64-system fp32 batch, {sum(sp['n_walkers'] for sp in specs)} walkers", flush=True)
    batch = system_batch(members, dt=0.005, qr_every=10, sr_every=20, precision="fp32")
    batch.advance(20)
    stats = batch.block_stats.cpu().numpy().copy()
    assert np.isfinite(stats).all()
    for i, m in enumerate(members):
        st, blocks = _alone(m, 20, qr=10, sr=20)
        _compare_system(batch, i, st, blocks, [(float(stats[i, 0]), float(stats[i, 1]))])
    print("64 systems equal their single-system runs", flush=True)
    picks = {min(range(64), key=lambda i: specs[i]["norb"] * specs[i]["n_walkers"]),
             max(range(64), key=lambda i: specs[i]["norb"]),
             max(range(64), key=lambda i: specs[i]["nocc"])}
    for i in sorted(picks):
        st_ref, blk = _oracle_block(specs[i])
        assert abs(stats[i, 0] - blk["energy"]) < 1e-4, (i, stats[i, 0], blk["energy"])
        assert abs(stats[i, 1] - blk["weight"]) / blk["weight"] < 1e-5, i
        w = batch.system_state(i)["walkers"].cpu().numpy()
        proj = w @ np.conj(np.swapaxes(w, 1, 2))
        proj_ref = st_ref.walkers @ np.conj(np.swapaxes(st_ref.walkers, 1, 2))
        err = np.linalg.norm(proj - proj_ref) / np.linalg.norm(proj_ref)
        assert err < 1e-4, (i, err)
        print(f"system {i} (seed {specs[i]['seed']}, n={specs[i]['norb']}, o={specs[i]['nocc']}): "
              f"block energy {stats[i, 0]:.10f} vs oracle {blk['energy']:.10f}", flush=True)
    del batch
    torch.cuda.empty_cache()
    b16 = system_batch(members, dt=0.005, qr_every=10, sr_every=20, precision="bf16x3")
    b16.advance(20)
    s16 = b16.block_stats.cpu().numpy()
    np.testing.assert_allclose(s16[:, 0], stats[:, 0], rtol=1e-5, atol=0)


@pytest.mark.parametrize("first_full,last_half", [(0, 1), (1, 0), (0, 0), (1, 1)])
def test_grouped_bf16x3_trotter_buckets_match_single(gpu_device, first_full, last_half):
    """One grouped bf16x3 Trotter launch per bucket (n = 40, 160, 176, 208:
    no, one and two V k-blocks held in LDS) equals afq_trotter_bf16x3_sym on
    each system alone, bit for bit, with random complex-symmetric V given as
    pair words of its upper blocks; with the halves of merging steps (E^2 - I
    first, no last half) against afq_trotter_bf16x3_sym_halves."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    from ad_afqmc_prototype_amd.prop.batched import system_batch
    from ad_afqmc_prototype_amd.prop.chol_afqmc_ops import chol_sym_rows, split_pairs
    systems = [(111, 40, 6, 170, 24), (112, 160, 24, 700, 20), (113, 176, 17, 800, 12),
               (114, 208, 32, 900, 10)]
    members = _members(gpu_device, systems, scale=0.05)
    batch = system_batch(members, dt=DT, qr_every=QR, sr_every=SR, precision="bf16x3")
    g = torch.Generator(device="cpu").manual_seed(9)
    for i, m in enumerate(batch.members):
        nw, n = m["nw"], m["n"]
        a = torch.randn((2 * nw, n, n), generator=g, dtype=torch.float64)
        a = 0.5 * (a + a.transpose(1, 2))
        a *= 0.2 / (DT ** 0.5 * n ** 0.5)
        # fp32 values, in the layout of the VHS GEMM's pair output: (2 nw, nsym)
        blocks = chol_sym_rows(a.float().double().to(gpu_device)).T.float().contiguous()
        batch.vhs.part(i, (2, nw, m["nsym"])).copy_(split_pairs(blocks).view(2, nw, m["nsym"]))
    st = kernels._stream(batch.desc[0])
    assert len(batch.trotter_buckets) == 4
    plain = (first_full, last_half) == (0, 1)
    for bk in batch.trotter_buckets:
        t = bk["tables"][0]
        if plain:
            kernels._call("afq_trotter_bf16x3_sym_grouped", t, bk["count"], t.data_ptr(), bk["blocks"],
                          bk["max_n"], 6, st)
        else:
            kernels._call("afq_trotter_bf16x3_sym_grouped_halves", t, bk["count"], t.data_ptr(), bk["blocks"],
                          bk["max_n"], 6, first_full, last_half, st)
    torch.cuda.synchronize()
    for i, m in enumerate(batch.members):
        nw, n, o = m["nw"], m["n"], m["o"]
        win = batch.walkers[0].part(i, (nw, n, o))
        e_first = batch.e_full.part(i) if first_full else batch.e_half.part(i)
        kw = {} if plain else {"e_last": batch.e_half.part(i) if last_half else None}
        ref = kernels.trotter(win, batch.vhs.part(i, (2, nw, m["nsym"])), e_first,
                              batch.sqrt_dt, 6, arith="bf16x3", **kw)
        got = batch.walkers[1].part(i, (nw, n, o))
        assert torch.equal(got, ref), f"system {i} (n={n}): grouped bf16x3 Trotter differs"
        assert float((ref - win).abs().max()) > 1e-3  # the step moved the walkers
