"""Propagation + measurement block (prop/blocks.py:18-95 of the reference).

Per block: n_prop_steps steps (with an optional QR every params.n_ortho steps),
orthonormalisation, overlaps, clipped local energies, block energy and
population-shift update, comb reconfiguration.  Everything stays on the device;
block scalars are returned as 0-d device tensors.
"""
from __future__ import annotations

from typing import Any, NamedTuple

import torch

from .. import kernels
from .. import walkers as wk
from ..core.ops import k_energy, meas_ops
from ..core.system import system
from ..ham.chol import ham_chol
from .afqmc import _log_overlap_and_greens, afqmc_step, split_key
from .chol_afqmc_ops import chol_afqmc_ctx, chol_afqmc_ops
from .types import afqmc_params, prop_state


class block_obs(NamedTuple):
    scalars: dict


def draw_zeta(key: torch.Tensor):
    """prop/blocks.py:81-82: key, subkey = split(key); zeta = uniform(subkey),
    a 0-d device tensor (the comb reads it on the device)."""
    key_next, sub = split_key(key)
    return kernels.rng_uniform(sub), key_next


def _require_whole(state: prop_state, what: str) -> None:
    if state.deferred:
        raise ValueError(f"{what}: the state's walkers lack their last one-body half "
                         "(afqmc_step(defer_final=True)); only another afqmc_step reads it")


def orthonormalize_state(state: prop_state, *, sys: system, trial_data, meas_ops_: meas_ops) -> prop_state:
    """walkers.py:206-211 + overlap refresh (prop/blocks.py:53-57)."""
    _require_whole(state, "orthonormalize_state")
    w = wk.orthonormalize(state.walkers, sys.walker_kind)
    lo, greens = _log_overlap_and_greens(w, meas_ops_, trial_data)
    return state._replace(walkers=w, log_overlaps=lo, greens=greens)


def measure(state: prop_state, *, params: afqmc_params, ham_data: ham_chol, trial_data: Any,
            meas_ops: meas_ops, meas_ctx: Any):
    """prop/blocks.py:59-79 on orthonormal walkers.  Returns (state with the
    population shift updated, {"energy", "weight"}, per-walker energies)."""
    _require_whole(state, "measure")
    e_kernel = meas_ops.require_kernel(k_energy)
    e = e_kernel(state.walkers, ham_data, meas_ctx, trial_data, greens=state.greens)
    thresh = (2.0 / float(params.dt)) ** 0.5
    out = kernels.block_energy(e.to(torch.complex128).contiguous(), state.weights,
                               state.pop_control_ene_shift, e_ref=state.e_estimate,
                               thresh=thresh, shift_ema=float(params.shift_ema))
    state = state._replace(pop_control_ene_shift=out[2])
    return state, {"energy": out[0], "weight": out[1]}, e


def reconfigure(state: prop_state, *, sys: system, trial_data, meas_ops_: meas_ops,
                zeta=None) -> tuple[prop_state, torch.Tensor]:
    """prop/blocks.py:81-92: comb reconfiguration; the Green's functions and
    log-overlaps of the selected walkers are gathered (the reference recomputes
    the overlaps of the gathered walkers, which gives the same values).
    zeta: a float, a 0-d float64 device tensor, or None (drawn on the device
    from the state's key)."""
    _require_whole(state, "reconfigure")
    if zeta is None:
        zeta, key = draw_zeta(state.rng_key)
    else:
        key = split_key(state.rng_key)[0]
    if not isinstance(zeta, torch.Tensor):
        zeta = float(zeta)
    w_sr, w_new, idx = wk.stochastic_reconfiguration_with_indices(state.walkers, state.weights,
                                                                  zeta, sys.walker_kind)
    lo = kernels.gather_rows(state.log_overlaps, idx)
    greens = state.greens
    if greens is not None:
        split = greens.split
        if split is not None:
            # the bf16 planes (2 nw rows of ldk, ldk a multiple of 8) move as
            # float64 words: a bit copy
            nw_old = greens.logdet.shape[0]
            split = tuple(kernels.gather_rows(x.view(2, nw_old, -1).view(torch.float64), idx, nplanes=2)
                          .view(torch.int16).reshape(2 * idx.shape[0], -1) for x in split)
        rows = None if greens.greens is None else kernels.gather_rows(greens.greens, idx, nplanes=2)
        greens = type(greens)(rows, kernels.gather_rows(greens.logdet, idx), split)
    return state._replace(walkers=w_sr, weights=w_new, log_overlaps=lo, rng_key=key,
                          greens=greens), idx


def afqmc_block(state: prop_state, *, sys: system, params: afqmc_params, ham_data: ham_chol,
                trial_data: Any, meas_ops: meas_ops, prop_ops: chol_afqmc_ops,
                prop_ctx: chol_afqmc_ctx, meas_ctx: Any, fields=None, zeta: float | None = None,
                return_indices: bool = False):
    """prop/blocks.py:22-95.  ``fields`` optionally injects the per-step fields
    (indexable by step), ``zeta`` the comb uniform.  With
    params.merge_one_body_halves, a step followed by another step leaves its
    final one-body half to it (afqmc_step(defer_final=True)); the steps before
    a QR and the block's last step apply theirs."""
    n_ortho = params.n_ortho
    merge = bool(getattr(params, "merge_one_body_halves", False))
    for t in range(params.n_prop_steps):
        f = None if fields is None else fields[t]
        qr_next = bool(n_ortho) and (t + 1) % n_ortho == 0
        state = afqmc_step(state, params=params, ham_data=ham_data, trial_data=trial_data,
                           meas_ops=meas_ops, prop_ops=prop_ops, prop_ctx=prop_ctx,
                           meas_ctx=meas_ctx, fields=f,
                           defer_final=merge and not qr_next and t + 1 < params.n_prop_steps)
        if qr_next and t + 1 < params.n_prop_steps:
            state = orthonormalize_state(state, sys=sys, trial_data=trial_data, meas_ops_=meas_ops)
    state = orthonormalize_state(state, sys=sys, trial_data=trial_data, meas_ops_=meas_ops)
    state, scalars, _ = measure(state, params=params, ham_data=ham_data, trial_data=trial_data,
                                meas_ops=meas_ops, meas_ctx=meas_ctx)
    state, idx = reconfigure(state, sys=sys, trial_data=trial_data, meas_ops_=meas_ops, zeta=zeta)
    obs = block_obs(scalars=scalars)
    if return_indices:
        return state, obs, idx
    return state, obs


def _clone_state(state: prop_state) -> prop_state:
    def c(x):
        if x is None:
            return None
        if isinstance(x, torch.Tensor):
            return x.clone()
        if type(x) is tuple:
            return tuple(c(y) for y in x)
        return type(x)(*(c(y) for y in x))
    return state._replace(walkers=c(state.walkers), weights=c(state.weights),
                          log_overlaps=c(state.log_overlaps), rng_key=c(state.rng_key),
                          pop_control_ene_shift=c(state.pop_control_ene_shift), greens=c(state.greens))


def _copy_state_(dst: prop_state, src: prop_state) -> None:
    def cp(d, s):
        if d is None or s is None:
            if d is not s:
                raise ValueError("state layouts differ (a cached entry is missing on one side)")
        elif isinstance(d, torch.Tensor):
            d.copy_(s)
        else:
            for a, b in zip(d, s):
                cp(a, b)
    for name in ("walkers", "weights", "log_overlaps", "rng_key", "pop_control_ene_shift", "greens"):
        cp(getattr(dst, name), getattr(src, name))


class captured_block:
    """One afqmc_block recorded once as a HIP graph and replayed in place.

    The fields and the comb uniform come from the state's key on the device
    (no injected arrays, no host round trip), so every replay is a new block:
    replay() advances ``self.state`` (the graph's static buffers) by one block
    and returns the block's energy and weight as 0-d device tensors (the
    graph's static outputs, overwritten by the next replay).  Results equal
    those of calling afqmc_block eagerly from the same state.  e_estimate is a
    host constant of the state, fixed at capture (the reference keeps it fixed
    too, prop/afqmc.py:138,146)."""

    def __init__(self, state: prop_state, *, sys: system, params: afqmc_params, ham_data: ham_chol,
                 trial_data: Any, meas_ops: meas_ops, prop_ops: chol_afqmc_ops,
                 prop_ctx: chol_afqmc_ctx, meas_ctx: Any):
        self._kw = dict(sys=sys, params=params, ham_data=ham_data, trial_data=trial_data,
                        meas_ops=meas_ops, prop_ops=prop_ops, prop_ctx=prop_ctx, meas_ctx=meas_ctx)
        if meas_ops.greens is not None and (state.greens is None or state.greens.greens is None):
            # the block ends with the full cache (QR refresh, comb gather): the
            # static state must hold the same layout.  A cache with only the
            # bf16 planes keeps them (the first step reads them) and gains rows.
            full = meas_ops.greens(state.walkers, trial_data)
            state = state._replace(greens=full if state.greens is None
                                   else state.greens._replace(greens=full.greens))
        self.state = _clone_state(state)
        # one eager block on a copy first (lazy operand layouts, kernel
        # attributes), on a side stream as graph capture requires
        torch.cuda.synchronize()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            afqmc_block(_clone_state(self.state), **self._kw)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            out, obs = afqmc_block(self.state, **self._kw)
            _copy_state_(self.state, out)
            self.energy = obs.scalars["energy"]
            self.weight = obs.scalars["weight"]
        del out, obs

    def replay(self):
        self.graph.replay()
        return self.energy, self.weight
