"""Host driver: equilibration blocks then sampling blocks (driver.py:22-176 of
the reference), around the HIP block.

Blocks of one chunk are enqueued back to back without host synchronisation;
the block energies/weights of a chunk come back to the host once per chunk
(the reference's jit + lax.scan over a chunk of blocks, driver.py:22-36).
With use_graph the block is recorded once as a HIP graph (fields and comb
uniforms drawn on the device from the state's key) and each block of a chunk
is one graph launch.
"""
from __future__ import annotations

import time
from functools import partial
from typing import Any, Callable

import numpy as np
import torch

from .core.ops import meas_ops as meas_ops_t
from .core.ops import trial_ops as trial_ops_t
from .core.system import system
from .ham.chol import ham_chol
from .prop.afqmc import init_prop_state
from .prop.blocks import afqmc_block, captured_block
from .prop.chol_afqmc_ops import make_chol_afqmc_ops
from .prop.types import afqmc_params
from .stats import blocking_analysis_ratio, reject_outliers

print = partial(print, flush=True)


def make_run_blocks(block: Callable):
    """Returns run_blocks(state, n_blocks) -> (state, energies, weights) with the
    per-block scalars still on the device."""

    def run_blocks(state, *, n_blocks: int):
        es, ws = [], []
        for _ in range(n_blocks):
            state, obs = block(state)
            es.append(obs.scalars["energy"])
            ws.append(obs.scalars["weight"])
        return state, torch.stack(es), torch.stack(ws)

    return run_blocks


def run_afqmc_energy(*, sys: system, params: afqmc_params, ham_data: ham_chol, trial_data: Any,
                     meas_ops: meas_ops_t, trial_ops: trial_ops_t,
                     block_fn: Callable = afqmc_block, prop_ops=None, verbose: bool = True,
                     use_graph: bool = False):
    """Returns (mean_energy, stderr, block_energies, block_weights) like
    driver.py:39-176: the block arrays hold the initial estimate, the
    equilibration blocks and the sampling blocks left after outlier rejection
    (driver.py:161-176).  use_graph
    replays one captured afqmc_block per block (same results as the eager
    block loop)."""
    if use_graph and block_fn is not afqmc_block:
        raise ValueError("use_graph records afqmc_block itself; a custom block_fn runs eagerly")
    if params.n_eql_blocks < 5 or params.n_blocks < 10:
        raise ValueError("run_afqmc_energy needs n_eql_blocks >= 5 and n_blocks >= 10 "
                         "(blocks are run in chunks of n_eql_blocks//5 and n_blocks//10)")
    log = print if verbose else (lambda *a, **k: None)
    if prop_ops is None:
        prop_ops = make_chol_afqmc_ops(ham_data, sys.walker_kind)
    prop_ctx = prop_ops.build_prop_ctx(trial_ops.get_rdm1(trial_data), params.dt)
    meas_ctx = meas_ops.build_meas_ctx(ham_data, trial_data)
    state = init_prop_state(sys=sys, n_walkers=params.n_walkers, seed=params.seed,
                            ham_data=ham_data, trial_ops_=trial_ops, trial_data=trial_data,
                            meas_ops=meas_ops)

    def block(state_in):
        return block_fn(state_in, sys=sys, params=params, ham_data=ham_data,
                        trial_data=trial_data, meas_ops=meas_ops, prop_ops=prop_ops,
                        prop_ctx=prop_ctx, meas_ctx=meas_ctx)

    if use_graph:
        cb = captured_block(state, sys=sys, params=params, ham_data=ham_data, trial_data=trial_data,
                            meas_ops=meas_ops, prop_ops=prop_ops, prop_ctx=prop_ctx, meas_ctx=meas_ctx)

        def run_blocks(_state, *, n_blocks: int):
            es, ws = [], []
            for _ in range(n_blocks):
                e, w = cb.replay()
                es.append(e.clone())
                ws.append(w.clone())
            return cb.state, torch.stack(es), torch.stack(ws)
    else:
        run_blocks = make_run_blocks(block)
    t0 = time.perf_counter()
    t_mark = t0

    e_eq = [state.e_estimate]
    w_eq = [float(state.weights.sum().item())]
    log("\nEquilibration:\n")
    log(f"{'':4s}{'block':>9s}  {'E_blk':>14s}  {'W':>12s}  {'t[s]':>8s}")
    log(f"[eql {0:4d}/{params.n_eql_blocks}]  {state.e_estimate:14.10f}  {w_eq[0]:12.6e}  {0.0:8.1f}")
    chunk = params.n_eql_blocks // 5
    for start in range(0, params.n_eql_blocks, chunk):
        n = min(chunk, params.n_eql_blocks - start)
        state, e_c, w_c = run_blocks(state, n_blocks=n)
        e_c, w_c = e_c.cpu().numpy(), w_c.cpu().numpy()
        e_eq.extend(e_c.tolist())
        w_eq.extend(w_c.tolist())
        if start > 0:
            log(f"[eql {start:4d}/{params.n_eql_blocks}]  {e_c.mean():14.10f}  {w_c.mean():12.6e}  "
                f"{time.perf_counter() - t0:8.1f}")

    log("\nSampling:\n")
    log(f"{'':4s}{'block':>9s}  {'E_avg':>14s}  {'E_err':>10s}  {'E_block':>14s}  "
        f"{'W':>12s}  {'dt[s/bl]':>9s}  {'t[s]':>8s}")
    e_s, w_s = [], []
    chunk = params.n_blocks // 10
    for start in range(0, params.n_blocks, chunk):
        n = min(chunk, params.n_blocks - start)
        state, e_c, w_c = run_blocks(state, n_blocks=n)
        e_c, w_c = e_c.cpu().numpy(), w_c.cpu().numpy()
        e_s.extend(e_c.tolist())
        w_s.extend(w_c.tolist())
        if start > 0:
            now = time.perf_counter()
            st = blocking_analysis_ratio(np.asarray(e_s), np.asarray(w_s), print_q=False)
            log(f"[blk {start:4d}/{params.n_blocks}]  {st['mu']:14.10f}  {st['se_star']:10.3e}  "
                f"{e_c.mean():14.10f}  {w_c.mean():12.6e}  {(now - t_mark) / n:9.3f}  {now - t0:8.1f}")
            t_mark = now
    e_s = np.asarray(e_s)
    w_s = np.asarray(w_s)
    clean, _ = reject_outliers(np.column_stack((e_s, w_s)), obs=0)
    log(f"\nRejected {e_s.shape[0] - clean.shape[0]} outlier blocks.")
    log("\nFinal blocking analysis:")
    st = blocking_analysis_ratio(clean[:, 0], clean[:, 1], print_q=verbose)
    return (st["mu"], st["se_star"], np.concatenate([np.asarray(e_eq), clean[:, 0]]),
            np.concatenate([np.asarray(w_eq), clean[:, 1]]))
