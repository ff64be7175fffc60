"""Walker-step throughput of the phaseless-AFQMC hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                    [--precision fp64] [--no-cpu-baseline]

A step is one pass of afqmc_step (force bias, field shifts, VHS, Trotter,
overlaps, phaseless update) over every walker, with QR every 10 steps and a
measurement + comb reconfiguration every 20 steps amortised into the timed
region (the cadence of the 200-orbital workload).  Inputs (Hamiltonian,
trial, walkers, auxiliary fields, comb uniforms) are synthetic and resident in
HBM before the timed region starts.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}
PEAK_HBM_GBS = 8000.0


def flops_per_walker_step(n, o, g, nspin, rhf_trial):
    """Algorithmic real flops of one walker through the step (complex x real
    MAC = 4, complex x complex MAC = 8).  The VHS is symmetric, so only its
    n (n + 1) / 2 distinct entries count."""
    s_fb = 1 if rhf_trial else nspin
    vhs = 2.0 * g * n * (n + 1)
    fb = s_fb * 4.0 * g * o * n
    taylor = nspin * 6 * 8.0 * n * n * o
    one_body = nspin * 2 * 4.0 * n * n * o
    greens = nspin * (4.0 * o * o * n + (8.0 / 3.0) * o ** 3 + 8.0 * o * o * n)
    return dict(vhs=vhs, force_bias=fb, taylor=taylor, one_body=one_body, greens=greens,
                total=vhs + fb + taylor + one_body + greens)


def build(workload: str, device, precision: str = "fp64", chol_scale: float | None = None):
    import torch

    from ad_afqmc_prototype_amd import synthetic as syn
    from ad_afqmc_prototype_amd.core.system import system
    from ad_afqmc_prototype_amd.ham.chol import ham_chol
    from ad_afqmc_prototype_amd.meas.rhf import make_rhf_meas_ops
    from ad_afqmc_prototype_amd.prop.afqmc import init_prop_state
    from ad_afqmc_prototype_amd.prop.chol_afqmc_ops import make_chol_afqmc_ops
    from ad_afqmc_prototype_amd.prop.types import afqmc_params
    from ad_afqmc_prototype_amd.trial.rhf import (make_rhf_trial_ops, make_uhf_trial_ops,
                                                  rhf_trial, uhf_trial)

    cfg = syn.WORKLOADS[workload]
    s = syn.make_system(cfg["seed"], cfg["norb"], cfg["nocc"], cfg["naux"], uhf=cfg["uhf"],
                        chol_scale=cfg["chol_scale"] if chol_scale is None else chol_scale)
    kind = cfg["walker_kind"]
    sys_ = system(norb=s.norb, nelec=(s.nocc, s.nocc), walker_kind=kind)
    ham = ham_chol.from_arrays(s.h0, s.h1, s.chol, device=device)
    if cfg["uhf"]:
        trial, t_ops = uhf_trial(s.mo_up, s.mo_dn), make_uhf_trial_ops(sys_)
    else:
        trial, t_ops = rhf_trial(s.mo_up), make_rhf_trial_ops(sys_)
    m_ops = make_rhf_meas_ops(sys_, precision=precision)
    p_ops = make_chol_afqmc_ops(ham, kind, precision=precision)
    params = afqmc_params(dt=cfg["dt"], n_walkers=cfg["n_walkers"], n_prop_steps=cfg["sr_every"],
                          n_ortho=cfg["qr_every"])
    p_ctx = p_ops.build_prop_ctx(t_ops.get_rdm1(trial), params.dt)
    m_ctx = m_ops.build_meas_ctx(ham, trial)
    state = init_prop_state(sys=sys_, n_walkers=cfg["n_walkers"], seed=cfg["seed"], ham_data=ham,
                            trial_ops_=t_ops, trial_data=trial, meas_ops=m_ops)
    torch.cuda.synchronize()
    return dict(cfg=cfg, s=s, sys=sys_, ham=ham, trial=trial, t_ops=t_ops, m_ops=m_ops,
                p_ops=p_ops, params=params, p_ctx=p_ctx, m_ctx=m_ctx, state=state)


def run_steps(ctx, state, fields, t0, k, cadence=True, energies=None, comb=None):
    """k steps starting at global step index t0 with the QR / measurement +
    reconfiguration cadence of the workload (block energies appended to
    ``energies`` as device scalars, comb indices to ``comb``).  As in
    afqmc_block, a step followed by another step leaves its final one-body
    half to it where the propagator merges halves (params.merge_one_body_halves);
    the state returned has all its halves."""
    from ad_afqmc_prototype_amd import synthetic as syn
    from ad_afqmc_prototype_amd.prop.afqmc import afqmc_step
    from ad_afqmc_prototype_amd.prop.blocks import measure, orthonormalize_state, reconfigure

    cfg = ctx["cfg"]
    kw = dict(ham_data=ctx["ham"], trial_data=ctx["trial"], meas_ops=ctx["m_ops"],
              prop_ops=ctx["p_ops"], prop_ctx=ctx["p_ctx"], meas_ctx=ctx["m_ctx"])
    merge = bool(getattr(ctx["params"], "merge_one_body_halves", False))
    for i in range(k):
        t = t0 + i
        block_end = cadence and ((t + 1) % cfg["sr_every"] == 0 or (t + 1) % cfg["qr_every"] == 0)
        state = afqmc_step(state, params=ctx["params"], fields=fields[t % len(fields)],
                           defer_final=merge and i + 1 < k and not block_end, **kw)
        if not cadence:
            continue
        if (t + 1) % cfg["sr_every"] == 0:
            state = orthonormalize_state(state, sys=ctx["sys"], trial_data=ctx["trial"],
                                         meas_ops_=ctx["m_ops"])
            state, sc, _ = measure(state, params=ctx["params"], ham_data=ctx["ham"],
                                   trial_data=ctx["trial"], meas_ops=ctx["m_ops"], meas_ctx=ctx["m_ctx"])
            if energies is not None:
                energies.append(sc["energy"])
            state, idx = reconfigure(state, sys=ctx["sys"], trial_data=ctx["trial"],
                                     meas_ops_=ctx["m_ops"],
                                     zeta=syn.zeta_for_event(cfg["seed"], (t + 1) // cfg["sr_every"]))
            if comb is not None:
                comb.append(idx)
        elif (t + 1) % cfg["qr_every"] == 0:
            state = orthonormalize_state(state, sys=ctx["sys"], trial_data=ctx["trial"],
                                         meas_ops_=ctx["m_ops"])
    return state


def warm_block_end(ctx, state):
    """QR, measurement and comb once on the state's values, result dropped (the
    functions are pure): their first launches and first buffer allocations then
    fall before the timed pass, which runs them at the block cadence."""
    from ad_afqmc_prototype_amd.prop.blocks import measure, orthonormalize_state, reconfigure

    warm = orthonormalize_state(state, sys=ctx["sys"], trial_data=ctx["trial"], meas_ops_=ctx["m_ops"])
    warm, _, _ = measure(warm, params=ctx["params"], ham_data=ctx["ham"], trial_data=ctx["trial"],
                         meas_ops=ctx["m_ops"], meas_ctx=ctx["m_ctx"])
    warm, _ = reconfigure(warm, sys=ctx["sys"], trial_data=ctx["trial"], meas_ops_=ctx["m_ops"], zeta=0.5)
    del warm


def cpu_baseline(workload: str, single_thread: bool = False, nw: int = 128):
    """The CPU oracle (numpy fp64, op for op the reference algorithm) on a
    subsample of the same workload's walkers, timed on this host with all of
    its BLAS threads, or with one (the reference pins its own CPU runs to one
    thread, config.py:49-54)."""
    import contextlib

    from threadpoolctl import threadpool_info, threadpool_limits

    ctx = threadpool_limits(limits=1, user_api="blas") if single_thread else contextlib.nullcontext()
    with ctx:
        threads = max([i.get("num_threads", 1) for i in threadpool_info()
                       if i.get("user_api") == "blas"] or [1])
        return _cpu_baseline_run(workload, threads, nw)


def _cpu_baseline_run(workload: str, threads: int, nw: int):
    """One block of the workload's cadence (sr_every steps, a QR every
    qr_every, then measurement and comb) on ``nw`` walkers: the work the GPU's
    timed region does per walker-step, at the same amortisation."""
    import os

    import numpy as np

    from ad_afqmc_prototype_amd import synthetic as syn
    from oracle import afqmc_oracle as orc

    cfg = syn.WORKLOADS[workload]
    s = syn.make_system(cfg["seed"], cfg["norb"], cfg["nocc"], cfg["naux"], uhf=cfg["uhf"],
                        chol_scale=cfg["chol_scale"])
    kind = cfg["walker_kind"]
    sd = orc.make_system_data(s.h0, s.h1, s.chol, (s.mo_up, s.mo_dn), kind, cfg["dt"])
    prm = orc.params(dt=cfg["dt"])
    st = orc.init_state(sd, nw, overlap_mode="log")
    # one untimed step (first-touch of the BLAS buffers)
    orc.afqmc_step(st, syn.fields_for_step(cfg["seed"], 0, nw, cfg["naux"]), sd, prm, "log")
    n_steps = cfg["sr_every"]
    t0 = time.perf_counter()
    orc.run_cadenced(st, sd, prm, n_steps=n_steps, qr_every=cfg["qr_every"], sr_every=cfg["sr_every"],
                     fields_fn=lambda t: syn.fields_for_step(cfg["seed"], t + 1, nw, cfg["naux"]),
                     zeta_fn=lambda k: syn.zeta_for_event(cfg["seed"], k), overlap_mode="log")
    el = time.perf_counter() - t0
    return {"value": nw * n_steps / el, "unit": "walker-steps/s", "cores": int(threads),
            "host_cpus": os.cpu_count(), "kind": "port",
            "sample": f"{workload} shapes, {nw} walkers x {n_steps} afqmc_step with a QR every "
                      f"{cfg['qr_every']} and measurement + comb after {cfg['sr_every']} (numpy fp64 "
                      f"oracle, log-overlaps, {el:.1f} s, {threads} BLAS threads)"}


def accuracy_check(workload: str, precision: str, device, chol_scale: float, n_blocks: int):
    """Block energies of the mixed and the fp64 step under common random
    numbers (same fields, same comb uniforms, same initial state) over
    n_blocks blocks of the workload, its Cholesky vectors scaled by chol_scale.

    ``resync``: every mixed block starts from the fp64 state at the start of
    that block, so the difference is the precision error of one block (the
    20 steps, QR, and the measurement).  ``free``: the mixed run continues
    from its own state.  The comb reconfiguration is discontinuous in the
    weights: once a weight differing in the last bits lands on the other
    side of a comb tooth, a different walker is copied, the two runs become
    different samples and their difference grows to the statistical noise of
    a block."""
    import torch

    ctx64 = build(workload, device, "fp64", chol_scale)
    ctxmx = build(workload, device, precision, chol_scale)
    cfg = ctx64["cfg"]
    gen = torch.Generator(device=device)
    gen.manual_seed(cfg["seed"] + 1000)
    sr = cfg["sr_every"]
    fields = [torch.randn((cfg["n_walkers"], cfg["naux"]), dtype=torch.float64, device=device,
                          generator=gen) for _ in range(n_blocks * sr)]
    st64, stfree = ctx64["state"], ctxmx["state"]
    e64, eres, efree, c64, cfree = [], [], [], [], []
    for b in range(n_blocks):
        start = st64
        st64 = run_steps(ctx64, st64, fields, b * sr, sr, cadence=True, energies=e64, comb=c64)
        run_steps(ctxmx, start, fields, b * sr, sr, cadence=True, energies=eres)
        stfree = run_steps(ctxmx, stfree, fields, b * sr, sr, cadence=True, energies=efree, comb=cfree)
    e64, eres, efree = ([float(e.item()) for e in x] for x in (e64, eres, efree))
    # walkers whose comb pick differs between the fp64 and the free-running mixed run
    comb_diff = [int((a != b).sum().item()) for a, b in zip(c64, cfree)]
    del ctx64, ctxmx, fields, st64, stfree, start, c64, cfree
    torch.cuda.empty_cache()
    d_res = [1e3 * abs(a - b) for a, b in zip(eres, e64)]
    d_free = [1e3 * abs(a - b) for a, b in zip(efree, e64)]
    return {"chol_scale": chol_scale, "n_blocks": n_blocks, "precision": precision,
            "block_energies_fp64_Ha": e64, "block_energies_mixed_resync_Ha": eres,
            "resync_abs_diff_mHa": d_res, "max_resync_abs_diff_mHa": max(d_res) if d_res else None,
            "block_energies_mixed_free_Ha": efree, "free_abs_diff_mHa": d_free,
            "free_comb_picks_differing": comb_diff}


def end_to_end(workload: str, precision: str, device, n_eql: int, n_blocks: int):
    """The driver (equilibration + sampling blocks, blocking analysis) on the
    workload with the device random stream (fields and comb uniforms from the
    seed's Threefry keys: the fp64 and the mixed run see the same stream), one
    captured HIP graph per block."""
    import dataclasses

    import torch

    from ad_afqmc_prototype_amd.driver import run_afqmc_energy

    ctx = build(workload, device, precision)
    cfg = ctx["cfg"]
    params = dataclasses.replace(ctx["params"], n_eql_blocks=n_eql, n_blocks=n_blocks, seed=cfg["seed"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mu, se, e_b, w_b = run_afqmc_energy(sys=ctx["sys"], params=params, ham_data=ctx["ham"],
                                        trial_data=ctx["trial"], meas_ops=ctx["m_ops"],
                                        trial_ops=ctx["t_ops"], prop_ops=ctx["p_ops"], verbose=False,
                                        use_graph=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    del ctx
    torch.cuda.empty_cache()
    return {"mean_Ha": mu, "se_Ha": se, "n_eql_blocks": n_eql, "n_blocks": n_blocks,
            "steps_per_block": params.n_prop_steps, "wall_s": round(wall, 2),
            "block_energies_Ha": [float(x) for x in e_b]}


def batched_run(device, n_steps: int, n_systems: int = 64, chol_scale: float = 0.05,
                baseline_steps: int = 20, baselines: bool = True):
    """Uneven-system batch (N ~ U[40,160], nocc ~ U[6,24], naux = round(N U[4,6]),
    walkers ~ U[64,512], seeds 100..), fp32 contractions, dt 0.005, QR every 10
    and measurement + comb every 20 steps: walker-steps/s of the grouped step
    (eager, and one HIP graph per block) and of the same systems stepped one
    after another through afqmc_step (``baseline_steps`` steps)."""
    import numpy as np
    import torch

    from ad_afqmc_prototype_amd import synthetic as syn
    from ad_afqmc_prototype_amd.ham.chol import ham_chol
    from ad_afqmc_prototype_amd.prop.afqmc import afqmc_step
    from ad_afqmc_prototype_amd.prop.batched import batch_system, captured_batch_block, system_batch
    from ad_afqmc_prototype_amd.prop.blocks import measure, orthonormalize_state, reconfigure
    from ad_afqmc_prototype_amd.prop.types import afqmc_params
    from ad_afqmc_prototype_amd.trial.rhf import rhf_trial

    specs = syn.uneven_batch_specs(n_systems)
    members = []
    for sp in specs:
        s = syn.make_system(sp["seed"], sp["norb"], sp["nocc"], sp["naux"], chol_scale=chol_scale)
        members.append(batch_system(ham=ham_chol.from_arrays(s.h0, s.h1, s.chol, device=device),
                                    trial=rhf_trial(s.mo_up), n_walkers=sp["n_walkers"], seed=sp["seed"]))
    dt, qr, sr = 0.005, 10, 20
    batch = system_batch(members, dt=dt, qr_every=qr, sr_every=sr)
    tuned = batch.tune(reps=5)
    nw_tot = batch.total_walkers
    n_steps = max(sr, n_steps // sr * sr)
    out = {"workload": f"{n_systems} uneven systems (seeds {specs[0]['seed']}..{specs[-1]['seed']}), "
                       f"{nw_tot} walkers, chol_scale {chol_scale}",
           "n_systems": n_systems, "total_walkers": nw_tot, "steps": n_steps, "dt": dt,
           "qr_every": qr, "measure_sr_every": sr, "precision": batch.precision, "gemm_shapes": list(batch.gemm_shapes),
           "tune_ms": tuned["ms"], "gflop_per_step": round(batch.flops_per_step() / 1e9, 2)}
    # eager grouped launches (one block of warm-up first)
    batch.advance(sr)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.advance(n_steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    out["grouped_eager"] = {"walker_steps_per_s": nw_tot * n_steps / el, "ms_per_step": 1e3 * el / n_steps}
    # one graph launch per block
    cap = captured_batch_block(batch, warmup_blocks=1)
    cap.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n_steps // sr):
        cap.replay()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    out["grouped_graph"] = {"walker_steps_per_s": nw_tot * n_steps / el, "ms_per_step": 1e3 * el / n_steps,
                            "tflops_algorithmic": round(batch.flops_per_step() * n_steps / el / 1e12, 2)}
    stats = batch.block_stats.cpu().numpy()
    out["block_energy_finite"] = bool(np.isfinite(stats).all())
    del cap, batch
    torch.cuda.empty_cache()
    if not baselines:
        return out
    # the same batch in fp32 (fp32 contractions with tuned tile shapes, fp32 Trotter)
    b32 = system_batch(members, dt=dt, qr_every=qr, sr_every=sr, precision="fp32")
    t32 = b32.tune(reps=5)
    b32.advance(sr)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b32.advance(n_steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    out["grouped_eager_fp32"] = {"walker_steps_per_s": nw_tot * n_steps / el, "ms_per_step": 1e3 * el / n_steps,
                                      "gemm_shapes": list(b32.gemm_shapes), "tune_ms": t32["ms"]}
    del b32
    torch.cuda.empty_cache()
    # the same systems one after another through the single-system step
    from ad_afqmc_prototype_amd.core.system import system
    from ad_afqmc_prototype_amd.meas.rhf import make_rhf_meas_ops
    from ad_afqmc_prototype_amd.prop.afqmc import init_prop_state
    from ad_afqmc_prototype_amd.prop.chol_afqmc_ops import make_chol_afqmc_ops
    from ad_afqmc_prototype_amd.trial.rhf import make_rhf_trial_ops
    runs = []
    for m in members:
        n, o = m.trial.norb, m.trial.nocc
        sys_ = system(n, (o, o), "restricted")
        t_ops, m_ops = make_rhf_trial_ops(sys_), make_rhf_meas_ops(sys_, precision="fp32")
        p_ops = make_chol_afqmc_ops(m.ham, "restricted", precision="fp32")
        kw = dict(params=afqmc_params(dt=dt, n_walkers=m.n_walkers), ham_data=m.ham, trial_data=m.trial,
                  meas_ops=m_ops, prop_ops=p_ops, prop_ctx=p_ops.build_prop_ctx(t_ops.get_rdm1(m.trial), dt),
                  meas_ctx=m_ops.build_meas_ctx(m.ham, m.trial))
        st = init_prop_state(sys=sys_, n_walkers=m.n_walkers, seed=m.seed, ham_data=m.ham, trial_ops_=t_ops,
                             trial_data=m.trial, meas_ops=m_ops)
        runs.append((sys_, kw, st))

    def loop(k):
        for j, (sys_, kw, st) in enumerate(runs):
            for t in range(1, k + 1):
                st = afqmc_step(st, **kw)
                if t % sr == 0:
                    st = orthonormalize_state(st, sys=sys_, trial_data=kw["trial_data"], meas_ops_=kw["meas_ops"])
                    st, _, _ = measure(st, params=kw["params"], ham_data=kw["ham_data"],
                                       trial_data=kw["trial_data"], meas_ops=kw["meas_ops"],
                                       meas_ctx=kw["meas_ctx"])
                    st, _ = reconfigure(st, sys=sys_, trial_data=kw["trial_data"], meas_ops_=kw["meas_ops"])
                elif t % qr == 0:
                    st = orthonormalize_state(st, sys=sys_, trial_data=kw["trial_data"], meas_ops_=kw["meas_ops"])
            runs[j] = (sys_, kw, st)

    loop(2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loop(baseline_steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    out["one_system_at_a_time"] = {"walker_steps_per_s": nw_tot * baseline_steps / el,
                                   "ms_per_step": 1e3 * el / baseline_steps, "steps": baseline_steps}
    out["graph_speedup_vs_one_at_a_time"] = (out["grouped_graph"]["walker_steps_per_s"]
                                             / out["one_system_at_a_time"]["walker_steps_per_s"])
    del runs
    torch.cuda.empty_cache()
    return out


def adjoint_run_bench(device):
    """d<E>/d lambda and d<E>/dL of the 96-orbital workload (16 occupied, 480
    vectors, 512 walkers, dt 0.01, 50 steps in blocks of 10 = one checkpoint per
    10 steps, seed 3, V = 0.01 sym(N(0,1))): wall time of the energy + gradient
    (forward on the HIP kernels, hand-written reverse pass) and of a forward run
    alone, and the agreement with central differences at lambda = +-1e-4."""
    import numpy as np
    import torch

    from ad_afqmc_prototype_amd import synthetic as syn
    from ad_afqmc_prototype_amd.prop import adjoint as adj

    cfg = syn.WORKLOADS["n96_o16_g480_w512_rhf"]
    s = syn.make_system(cfg["seed"], cfg["norb"], cfg["nocc"], cfg["naux"], chol_scale=cfg["chol_scale"])
    n_blocks, n_prop = 5, 10
    fields, zetas = adj.device_random_numbers(cfg["seed"], n_blocks, n_prop, cfg["n_walkers"], cfg["naux"],
                                              device)
    run = adj.adjoint_run(trial=torch.tensor(s.mo_up, device=device), h0=float(s.h0), dt=cfg["dt"],
                          n_walkers=cfg["n_walkers"], n_blocks=n_blocks, n_prop_steps=n_prop,
                          fields=fields, zetas=zetas)
    rng = np.random.default_rng(cfg["seed"] + 1000)
    m = rng.standard_normal((cfg["norb"], cfg["norb"]))
    v = torch.tensor(0.01 * 0.5 * (m + m.T), device=device)
    h1 = torch.tensor(s.h1, device=device)
    chol = torch.tensor(s.chol, device=device)
    adj.energy_gradients(h1, chol, run, v)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = adj.energy_gradients(h1, chol, run, v)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    t0 = time.perf_counter()
    jp = adj.run_blocks(h1 + 1e-4 * v, chol, run)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    jm = adj.run_blocks(h1 - 1e-4 * v, chol, run)
    same = all(bool((a == b).all()) for a, b in zip(jp[3], jm[3]))
    fd = (jp[0] - jm[0]) / 2e-4
    return {"workload": "96 orbitals, 16 occupied, 480 Cholesky vectors, 512 walkers, dt 0.01, "
                        "50 steps (5 blocks of 10, one checkpoint each), fp64",
            "energy_Ha": out["energy"], "d_energy_d_lambda": out["d_lambda"],
            "central_difference": fd, "rel_diff": abs(out["d_lambda"] - fd) / abs(fd),
            "comb_picks_identical_at_pm_lambda": same,
            "s_energy_and_gradient": round(t_grad, 3), "s_forward_only": round(t_fwd, 3),
            "gradient_cost_in_forward_runs": round(t_grad / t_fwd, 2)}


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summaries (profiles/), each scaled by its copy-kernel
    calibration; None when not measured."""
    short = workload.rsplit("_", 1)[0]
    path = os.path.join(ROOT, "profiles", f"traffic_{short}.json")
    try:
        data = json.load(open(path))[workload]
    except (OSError, KeyError, ValueError):
        return None
    row = data.get(kernel)
    cal = data.get("calibration", {})
    if not row or row.get("fetch_kib_raw") is None or row.get("write_kib_raw") is None:
        return None
    return 1024.0 * (row["fetch_kib_raw"] * cal.get("fetch_factor", 1.0)
                     + row["write_kib_raw"] * cal.get("write_factor", 1.0))


def kernel_table(ktimes, n, o, g, nw, nspin, rhf, precision):
    """Algorithmic flops per launch and peak of every timed entry point."""
    fl = flops_per_walker_step(n, o, g, nspin, rhf)
    taylor = (fl["taylor"] + fl["one_body"]) * nw
    spec = {
        "afq_vhs_sym:vhs": (fl["vhs"] * nw, precision),
        "afq_force_bias": (fl["force_bias"] * nw, "fp64"),
        "afq_dgemm_nt_splitk:force_bias": (fl["force_bias"] * nw, "fp64"),
        "afq_sgemm_nt:force_bias": (fl["force_bias"] * nw, "fp32"),
        "afq_gemm_bf16x3_nt:force_bias": (fl["force_bias"] * nw, "bf16x3"),
        "afq_gemm_bf16x3_nt_splitk:force_bias": (fl["force_bias"] * nw, "bf16x3"),
        "afq_trotter": (taylor, "fp64"),
        "afq_trotter_f32": (taylor, "fp32"),
        "afq_trotter_bf16x3": (taylor, "bf16x3"),
        "afq_trotter_bf16x3_sym": (taylor, "bf16x3"),
        "afq_trotter_bf16x3_sym_halves": (taylor, "bf16x3"),
        # without the last one-body half: the Taylor series and one half
        "afq_trotter_bf16x3_sym_halves:deferred": ((fl["taylor"] + fl["one_body"] / 2) * nw, "bf16x3"),
        "afq_gemm_bf16x3_nt:vhs": (fl["vhs"] * nw, "bf16x3"),
        "afq_overlap_greens": (fl["greens"] * nw, "fp64"),
        "afq_energy": (4.0 * g * o * o * n * nspin * nw, "fp64"),
        "afq_energy_f32": (4.0 * g * o * o * n * nspin * nw, "fp32"),
        "afq_energy_bf16x3": (4.0 * g * o * o * n * nspin * nw, "bf16x3"),
        # the exchange as a quadratic form: Z = G Y per spin sector, M = 2 nw
        # (real and imaginary rows), N = K = o n
        "afq_dgemm_nt:exchange": (4.0 * nw * (o * n) ** 2, "fp64"),
        "afq_sgemm_nt:exchange": (4.0 * nw * (o * n) ** 2, "fp32"),
        "afq_gemm_bf16x3_nt:exchange": (4.0 * nw * (o * n) ** 2, "bf16x3"),
        # the same with Y's lower triangle (Z = 2 G L^T, column tiles stopping at
        # the diagonal): 2 nw rows x the on (on + 1) / 2 entries of L, 2 flops each
        "afq_dgemm_nt_lower:exchange": (2.0 * nw * o * n * (o * n + 1), "fp64"),
        "afq_sgemm_nt_lower:exchange": (2.0 * nw * o * n * (o * n + 1), "fp32"),
        "afq_gemm_bf16x3_nt_lower:exchange": (2.0 * nw * o * n * (o * n + 1), "bf16x3"),
    }
    out = {}
    for name, t in ktimes.items():
        row = {"count": t["count"], "avg_ms": round(t["avg_ms"], 4), "total_ms": round(t["total_ms"], 3)}
        if name in spec:
            flops, prec = spec[name]
            # bf16x3 issues three bf16 MFMA products per useful product
            peak = PEAK_TFLOPS["bf16"] / 3.0 if prec == "bf16x3" else PEAK_TFLOPS[prec]
            ach = flops / (t["avg_ms"] * 1e-3) / 1e12
            row.update({"tflops": round(ach, 2), "peak_tflops": round(peak, 1), "frac": round(ach / peak, 4),
                        "precision": prec, "gflop_per_launch": round(flops / 1e9, 3)})
        out[name] = row
    return out


def timed_run(args, precision, device):
    import torch

    from ad_afqmc_prototype_amd import kernels

    ctx = build(args.workload, device, precision)
    cfg = ctx["cfg"]
    nw, g = cfg["n_walkers"], cfg["naux"]
    gen = torch.Generator(device=device)
    gen.manual_seed(cfg["seed"])
    n_fields = min(args.warmup + args.steps, 40)
    fields = [torch.randn((nw, g), dtype=torch.float64, device=device, generator=gen)
              for _ in range(n_fields)]
    state = run_steps(ctx, ctx["state"], fields, 0, args.warmup, cadence=not args.no_cadence)
    if not args.no_cadence:
        warm_block_end(ctx, state)
    torch.cuda.synchronize()
    energies = []
    # the timed pass: nothing but the steps in the stream
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    state = run_steps(ctx, state, fields, args.warmup, args.steps, cadence=not args.no_cadence,
                      energies=energies)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # a second pass over the next K steps with a pair of HIP events around every
    # library call (on its launch stream) for the per-kernel durations; the
    # event records add a little time to each step, reported separately
    with kernels.timing(pool=16 * args.steps + 64) as tm:
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        state = run_steps(ctx, state, fields, args.warmup + args.steps, args.steps, cadence=not args.no_cadence)
        torch.cuda.synchronize()
        elapsed_inst = time.perf_counter() - t1
    ktimes = tm.summary()
    e_block = [float(e.item()) for e in energies]
    # the driver's path: whole blocks (n_prop_steps steps with QR and the
    # measurement + comb) replayed as one HIP graph each, fields and comb
    # uniforms drawn on the device from the state's key inside the graph
    elapsed_graph = None
    if not args.no_cadence and args.steps % ctx["params"].n_prop_steps == 0:
        from ad_afqmc_prototype_amd.prop.blocks import captured_block
        cb = captured_block(state, sys=ctx["sys"], params=ctx["params"], ham_data=ctx["ham"],
                            trial_data=ctx["trial"], meas_ops=ctx["m_ops"], prop_ops=ctx["p_ops"],
                            prop_ctx=ctx["p_ctx"], meas_ctx=ctx["m_ctx"])
        cb.replay()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        for _ in range(args.steps // ctx["params"].n_prop_steps):
            cb.replay()
        torch.cuda.synchronize()
        elapsed_graph = time.perf_counter() - t2
        del cb
    del ctx, state, fields
    torch.cuda.empty_cache()
    return elapsed, ktimes, e_block, elapsed_inst, elapsed_graph


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="n200_o32_g1000_w2048_rhf")
    ap.add_argument("--precision", default="bf16x3", choices=["fp64", "fp32", "bf16x3"])
    ap.add_argument("--no-fp64-compare", action="store_true",
                    help="skip the fp64 run that the mixed throughput is compared with")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true",
                    help="skip the fp32 run (fp32 contractions and Trotter) of the bf16x3 bench")
    ap.add_argument("--accuracy-scale", type=float, default=None,
                    help="Cholesky scale of the common-random-number energy check (default: the "
                         "workload's own; 0 skips it)")
    ap.add_argument("--accuracy-blocks", type=int, default=4)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes 5 times (the known-size "
                         "kernel that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--end-to-end", type=int, nargs=2, default=(5, 20), metavar=("EQL", "SAMPLE"),
                    help="driver run (equilibration and sampling blocks) for the end_to_end field; "
                         "0 0 skips it")
    ap.add_argument("--batched-steps", type=int, default=100,
                    help="steps of the 64-system uneven batch for the batched field; 0 skips it")
    ap.add_argument("--batched-no-baselines", action="store_true",
                    help="run only the bf16x3 batch (no fp32 batch, no one-system-at-a-time run)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint field (gradient of the 96-orbital run vs central differences)")
    ap.add_argument("--no-cadence", action="store_true",
                    help="time afqmc_step alone (no QR / measurement / reconfiguration)")
    ap.add_argument("--fp64-production", default="n120_o20_g600_w1024_uhf", metavar="WORKLOAD",
                    help="workload of the fp64_production field (the plain fp64 method at production "
                         "shape: 120 orbitals, 20 occupied per spin, UHF trial, 600 Cholesky vectors, "
                         "1024 walkers); 'none' skips it")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the walker "
              "step would shard over walkers (one process per GPU), which is not run here.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from ad_afqmc_prototype_amd import synthetic as syn

    device = torch.device("cuda", 0)
    if args.copy_calibration:
        n = args.copy_calibration // 8
        src = torch.ones(n, dtype=torch.float64, device=device)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 8, "launches": 5}))
        return
    cfg = syn.WORKLOADS[args.workload]
    n, o, g, nw = cfg["norb"], cfg["nocc"], cfg["naux"], cfg["n_walkers"]
    nspin = 2 if cfg["walker_kind"] == "unrestricted" else 1

    elapsed, ktimes, e_block, elapsed_inst, elapsed_graph = timed_run(args, args.precision, device)
    table = kernel_table(ktimes, n, o, g, nw, nspin, not cfg["uhf"], args.precision)
    dominant = max((k for k in table if "tflops" in table[k]), key=lambda k: table[k]["total_ms"])
    dom = table[dominant]
    value = nw * args.steps / elapsed
    flops = flops_per_walker_step(n, o, g, nspin, not cfg["uhf"])
    dtype = {"fp64": "f64", "fp32": "f32 contractions + f32 Trotter, f64 LU/QR/weights/energy",
             "bf16x3": "bf16x3 contractions, Trotter and energy products (f32 sums), f64 LU/QR/"
                       "weights/energy sums"}[args.precision]
    out = {
        "metric": "walker-steps/s at N=200,naux=1000: mixed-prec vs FP64 GPU, |dE|<0.1 mHa",
        "value": value,
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        # the same K steps again with HIP events around every library call (the
        # kernel table and roofline come from that pass)
        "ms_per_step_with_kernel_events": 1e3 * elapsed_inst / args.steps,
        # the same cadence as captured blocks (one HIP graph per block, fields
        # and comb uniforms from the device key inside the graph)
        "ms_per_step_graph": None if elapsed_graph is None else 1e3 * elapsed_graph / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": args.workload, "norb": n, "nocc": o, "naux": g, "n_walkers": nw,
                   "walker_kind": cfg["walker_kind"], "dt": cfg["dt"], "chol_scale": cfg["chol_scale"],
                   "precision": args.precision,
                   "qr_every": None if args.no_cadence else cfg["qr_every"],
                   "measure_sr_every": None if args.no_cadence else cfg["sr_every"]},
        "roofline": {"bound": "mfma", "kernel": dominant, "achieved": dom["tflops"],
                     "peak": dom["peak_tflops"], "unit": "TFLOP/s", "frac": dom["frac"],
                     "traffic": measured_traffic(args.workload, dominant)},
        "step_tflops_algorithmic": round(flops["total"] * nw * args.steps / elapsed / 1e12, 2),
        "kernels": table,
        "block_energies": e_block,
    }
    if args.precision != "fp64" and not args.no_fp64_compare:
        el64, kt64, e64, _, el64_graph = timed_run(args, "fp64", device)
        out["fp64_value"] = nw * args.steps / el64
        out["fp64_ms_per_step"] = 1e3 * el64 / args.steps
        out["fp64_ms_per_step_graph"] = None if el64_graph is None else 1e3 * el64_graph / args.steps
        out["speedup_vs_fp64"] = out["value"] / out["fp64_value"]
        out["fp64_kernels"] = kernel_table(kt64, n, o, g, nw, nspin, not cfg["uhf"], "fp64")
        out["block_energy_abs_diff_mHa"] = [1e3 * abs(a - b) for a, b in zip(e_block, e64)]
    if args.precision == "bf16x3" and not args.no_fp32:
        # the workload's other mixed precision: fp32 contractions and Trotter
        el32, kt32, e32, _, el32_graph = timed_run(args, "fp32", device)
        out["fp32_value"] = nw * args.steps / el32
        out["fp32_ms_per_step"] = 1e3 * el32 / args.steps
        out["fp32_ms_per_step_graph"] = None if el32_graph is None else 1e3 * el32_graph / args.steps
        out["fp32_speedup_vs_fp64"] = (out["fp32_value"] / out["fp64_value"]) if "fp64_value" in out else None
        out["fp32_kernels"] = kernel_table(kt32, n, o, g, nw, nspin, not cfg["uhf"], "fp32")
        if "fp64_value" in out:
            out["fp32_block_energy_abs_diff_mHa"] = [1e3 * abs(a - b) for a, b in zip(e32, e64)]
        if args.accuracy_scale != 0:
            scale = cfg["chol_scale"] if args.accuracy_scale is None else args.accuracy_scale
            out["fp32_energy_check"] = accuracy_check(args.workload, "fp32", device, scale,
                                                      args.accuracy_blocks)
    if args.fp64_production not in ("", "none"):
        import copy
        pargs = copy.copy(args)
        pargs.workload = args.fp64_production
        pc = syn.WORKLOADS[pargs.workload]
        pn, po, pg, pnw = pc["norb"], pc["nocc"], pc["naux"], pc["n_walkers"]
        pspin = 2 if pc["walker_kind"] == "unrestricted" else 1
        elp, ktp, ep, _, elp_graph = timed_run(pargs, "fp64", device)
        out["fp64_production"] = {
            "workload": pargs.workload, "norb": pn, "nocc": po, "naux": pg, "n_walkers": pnw,
            "walker_kind": pc["walker_kind"], "dt": pc["dt"],
            "walker_steps_per_s": pnw * args.steps / elp, "ms_per_step": 1e3 * elp / args.steps,
            "ms_per_step_graph": None if elp_graph is None else 1e3 * elp_graph / args.steps,
            "block_energies": ep,
            "kernels": kernel_table(ktp, pn, po, pg, pnw, pspin, not pc["uhf"], "fp64")}
    if args.precision != "fp64" and args.accuracy_scale != 0:
        scale = cfg["chol_scale"] if args.accuracy_scale is None else args.accuracy_scale
        out["energy_check"] = accuracy_check(args.workload, args.precision, device, scale,
                                             args.accuracy_blocks)
    if args.end_to_end[1] > 0:
        e2e = {prec: end_to_end(args.workload, prec, device, args.end_to_end[0], args.end_to_end[1])
               for prec in ([args.precision, "fp64"] if args.precision != "fp64" else ["fp64"])}
        if args.precision != "fp64":
            a, b = e2e[args.precision], e2e["fp64"]
            e2e["mean_diff_mHa"] = 1e3 * (a["mean_Ha"] - b["mean_Ha"])
            e2e["se_fp64_mHa"] = 1e3 * b["se_Ha"]
            # Both runs draw the same fields and comb uniforms, so they follow one
            # trajectory until a comb pick differs (a uniform within the weight
            # perturbation of a cumulative-weight boundary); from then on they are
            # independent chains and only agree statistically.
            diffs = [1e3 * abs(x - y) for x, y in zip(a["block_energies_Ha"], b["block_energies_Ha"])]
            same = next((i for i, d in enumerate(diffs) if d > 0.01), len(diffs))
            e2e["blocks_on_common_trajectory"] = same
            e2e["max_abs_diff_mHa_common_trajectory"] = max(diffs[:same], default=None)
        out["end_to_end"] = e2e
    if args.batched_steps > 0:
        out["batched"] = batched_run(device, args.batched_steps, baselines=not args.batched_no_baselines)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_run_bench(device)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.workload)
        out["cpu_baseline_single_thread"] = cpu_baseline(args.workload, single_thread=True, nw=48)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
