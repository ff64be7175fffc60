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
    """Comb uniform in [0, 1) from the key's subkey (host draw, no device sync)."""
    key_next, sub = split_key(key)
    gen = torch.Generator()
    gen.manual_seed(sub)
    return float(torch.rand((), dtype=torch.float64, generator=gen)), key_next


def orthonormalize_state(state: prop_state, *, sys: system, trial_data, meas_ops_: meas_ops) -> prop_state:
    """walkers.py:206-211 + overlap refresh (prop/blocks.py:53-57)."""
    w = wk.orthonormalize(state.walkers, sys.walker_kind)
    lo, greens = _log_overlap_and_greens(w, meas_ops_, trial_data)
    return state._replace(walkers=w, log_overlaps=lo, greens=greens)


def measure(state: prop_state, *, params: afqmc_params, ham_data: ham_chol, trial_data: Any,
            meas_ops: meas_ops, meas_ctx: Any):
    """prop/blocks.py:59-79 on orthonormal walkers.  Returns (state with the
    population shift updated, {"energy", "weight"}, per-walker energies)."""
    e_kernel = meas_ops.require_kernel(k_energy)
    e = e_kernel(state.walkers, ham_data, meas_ctx, trial_data, greens=state.greens)
    thresh = (2.0 / float(params.dt)) ** 0.5
    out = kernels.block_energy(e.to(torch.complex128).contiguous(), state.weights,
                               state.pop_control_ene_shift, e_ref=state.e_estimate,
                               thresh=thresh, shift_ema=float(params.shift_ema))
    state = state._replace(pop_control_ene_shift=out[2])
    return state, {"energy": out[0], "weight": out[1]}, e


def reconfigure(state: prop_state, *, sys: system, trial_data, meas_ops_: meas_ops,
                zeta: float | None = None) -> tuple[prop_state, torch.Tensor]:
    """prop/blocks.py:81-92: comb reconfiguration; the Green's functions and
    log-overlaps of the selected walkers are gathered (the reference recomputes
    the overlaps of the gathered walkers, which gives the same values)."""
    if zeta is None:
        zeta, key = draw_zeta(state.rng_key)
    else:
        key = split_key(state.rng_key)[0]
    w_sr, w_new, idx = wk.stochastic_reconfiguration_with_indices(state.walkers, state.weights,
                                                                  float(zeta), sys.walker_kind)
    lo = kernels.gather_rows(state.log_overlaps, idx)
    greens = state.greens
    if greens is not None:
        greens = type(greens)(kernels.gather_rows(greens.greens, idx, nplanes=2),
                              kernels.gather_rows(greens.logdet, idx))
    return state._replace(walkers=w_sr, weights=w_new, log_overlaps=lo, rng_key=key,
                          greens=greens), idx


def afqmc_block(state: prop_state, *, sys: system, params: afqmc_params, ham_data: ham_chol,
                trial_data: Any, meas_ops: meas_ops, prop_ops: chol_afqmc_ops,
                prop_ctx: chol_afqmc_ctx, meas_ctx: Any, fields=None, zeta: float | None = None,
                return_indices: bool = False):
    """prop/blocks.py:22-95.  ``fields`` optionally injects the per-step fields
    (indexable by step), ``zeta`` the comb uniform."""
    n_ortho = params.n_ortho
    for t in range(params.n_prop_steps):
        f = None if fields is None else fields[t]
        state = afqmc_step(state, params=params, ham_data=ham_data, trial_data=trial_data,
                           meas_ops=meas_ops, prop_ops=prop_ops, prop_ctx=prop_ctx,
                           meas_ctx=meas_ctx, fields=f)
        if n_ortho and (t + 1) % n_ortho == 0 and t + 1 < params.n_prop_steps:
            state = orthonormalize_state(state, sys=sys, trial_data=trial_data, meas_ops_=meas_ops)
    state = orthonormalize_state(state, sys=sys, trial_data=trial_data, meas_ops_=meas_ops)
    state, scalars, _ = measure(state, params=params, ham_data=ham_data, trial_data=trial_data,
                                meas_ops=meas_ops, meas_ctx=meas_ctx)
    state, idx = reconfigure(state, sys=sys, trial_data=trial_data, meas_ops_=meas_ops, zeta=zeta)
    obs = block_obs(scalars=scalars)
    if return_indices:
        return state, obs, idx
    return state, obs
