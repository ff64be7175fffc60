"""Phaseless AFQMC step (prop/afqmc.py:17-147 of the reference) on HIP kernels.

One step enqueues, on the current stream and without host synchronisation:
force-bias GEMM (from the cached Green's function of the current walkers),
field shifts, VHS GEMM, fused Trotter step, Green's function / log-overlap of
the new walkers (cached for the next step) and the phaseless weight update.
"""
from __future__ import annotations

from typing import Any

import torch

from .. import kernels
from .. import walkers as wk
from ..core.ops import k_energy, k_force_bias, meas_ops, trial_ops
from ..core.system import system
from ..ham.chol import ham_chol
from ..meas.rhf import deferred_view, rhf_meas_ctx
from .chol_afqmc_ops import chol_afqmc_ctx, chol_afqmc_ops
from .types import afqmc_params, prop_state


def make_key(seed: int, device=None) -> torch.Tensor:
    """jax.random.PRNGKey(seed) (prop/afqmc.py:32) as a device uint32[2]."""
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    return kernels.rng_key(seed, device)


def split_key(key: torch.Tensor):
    """(new key, subkey) = jax.random.split(key), on the device."""
    return kernels.rng_split(key)


def draw_fields(key: torch.Tensor, nw: int, n_fields: int, device=None):
    """prop/afqmc.py:86-88: key, subkey = split(key); x = normal(subkey, (nw, n_fields)),
    drawn on the device from the Threefry stream (no host round trip)."""
    key_next, sub = split_key(key)
    return kernels.rng_normal(sub, (nw, n_fields)), key_next


def _log_overlap_and_greens(walkers, meas_ops_: meas_ops, trial_data):
    if meas_ops_.greens is not None:
        g = meas_ops_.greens(walkers, trial_data)
        return g.log_overlap, g
    if meas_ops_.log_overlap is not None:
        return meas_ops_.log_overlap(walkers, trial_data), None
    return torch.log(meas_ops_.overlap(walkers, trial_data).to(torch.complex128)), None


def init_prop_state(*, sys: system, n_walkers: int, seed: int, ham_data: ham_chol,
                    trial_ops_: trial_ops, trial_data: Any, meas_ops: meas_ops,
                    initial_walkers: Any | None = None,
                    initial_e_estimate: float | None = None) -> prop_state:
    """prop/afqmc.py:17-71."""
    device = ham_data.device
    weights = torch.ones((n_walkers,), dtype=torch.float64, device=device)
    if initial_walkers is None:
        initial_walkers = wk.init_walkers(sys=sys, rdm1=trial_ops_.get_rdm1(trial_data),
                                          n_walkers=n_walkers)
    log_ov, greens = _log_overlap_and_greens(initial_walkers, meas_ops, trial_data)
    if initial_e_estimate is not None:
        e_est = float(initial_e_estimate)
    else:
        meas_ctx = meas_ops.build_meas_ctx(ham_data, trial_data)
        e_kernel = meas_ops.require_kernel(k_energy)
        e_samples = torch.real(e_kernel(initial_walkers, ham_data, meas_ctx, trial_data, greens=greens))
        e_est = float(torch.mean(e_samples).item())
    pop_shift = torch.tensor(e_est, dtype=torch.float64, device=device)
    return prop_state(walkers=initial_walkers, weights=weights, log_overlaps=log_ov,
                      rng_key=make_key(seed, device), pop_control_ene_shift=pop_shift,
                      e_estimate=e_est, greens=greens)


def _force_bias_planes(state: prop_state, ham_data, trial_data, meas_ops_: meas_ops, meas_ctx,
                       n_chunks: int):
    greens = state.greens
    if greens is None and meas_ops_.greens is not None:
        greens = meas_ops_.greens(state.walkers, trial_data)
    if meas_ops_.force_bias_planes is not None:
        return meas_ops_.force_bias_planes(state.walkers, meas_ctx, trial_data, greens=greens)
    fb_kernel = meas_ops_.require_kernel(k_force_bias)
    fb = fb_kernel(state.walkers, ham_data, meas_ctx, trial_data, greens=greens)
    if fb.is_complex():
        return torch.stack([fb.real, fb.imag]).to(torch.float64).contiguous()
    return fb


def can_merge_halves(prop_ops: chol_afqmc_ops, meas_ops_: meas_ops, meas_ctx: Any, trial_data: Any) -> bool:
    """Whether afqmc_step can leave a step's final one-body half to the next
    step: the propagator takes its halves separately (the bf16x3 upper-block
    Trotter) and the measurement is the RHF/UHF one with real trial orbitals
    and a step Green's function (restricted or unrestricted walkers)."""
    return (bool(getattr(prop_ops, "merges_halves", False)) and isinstance(meas_ctx, rhf_meas_ctx)
            and not meas_ctx.is_complex and meas_ops_.step_greens is not None
            and not bool(getattr(trial_data, "is_complex", False)))


def afqmc_step(state: prop_state, *, params: afqmc_params, ham_data: ham_chol, trial_data: Any,
               meas_ops: meas_ops, prop_ops: chol_afqmc_ops, prop_ctx: chol_afqmc_ctx,
               meas_ctx: Any, fields: torch.Tensor | None = None, defer_final: bool = False) -> prop_state:
    """prop/afqmc.py:74-147.  ``fields`` (nw, n_fields) float64 may be injected;
    otherwise they are drawn from ``state.rng_key``.  Pure: ``state`` is not
    modified, every output is a new tensor.

    ``defer_final``: for a step that another step follows, leave out the final
    exp(-dt h1_eff / 2) where can_merge_halves holds; the returned state is
    ``deferred`` (walkers E^-1 W, Green's functions for the trial E C, log
    overlaps and weights those of W), and the next step applies exp(-dt h1_eff)
    before its Taylor series: E (T E W) and E E W' in one product.  Elsewhere
    the flag does nothing."""
    nw = wk.n_walkers(state.walkers)
    device = state.weights.device
    merge = can_merge_halves(prop_ops, meas_ops, meas_ctx, trial_data)
    deferred_in = bool(state.deferred)
    if deferred_in and not merge:
        raise ValueError("a deferred state (walkers without their last one-body half) needs the "
                         "propagator that merged the halves")
    defer_out = bool(defer_final) and merge
    if deferred_in or defer_out:
        trial_d, meas_ctx_d = deferred_view(meas_ctx, trial_data, prop_ctx.exp_h1_half)
    if fields is None:
        fields, key = draw_fields(state.rng_key, nw, prop_ops.n_fields(), device)
    else:
        key, _ = split_key(state.rng_key)
        fields = torch.as_tensor(fields, device=device)
        if fields.is_complex():
            fields = fields.real
        fields = fields.to(torch.float64).contiguous()

    fb = _force_bias_planes(state, ham_data, trial_d if deferred_in else trial_data, meas_ops,
                            meas_ctx_d if deferred_in else meas_ctx, params.n_chunks)
    halves = {"merged_first": True} if deferred_in else {}
    if defer_out:
        halves["last_half"] = False
    if getattr(prop_ops, "xbar_split", False) and params.n_chunks == 1:
        # the bf16x3 VHS takes xbar as bf16 planes, made in the same pass
        xbar, shift_term, fb_term, xs = kernels.field_shifts(fields, fb, prop_ctx.mf_shifts, prop_ctx.sqrt_dt,
                                                             split=True)
        walkers_new = prop_ops.apply_trotter_planes(state.walkers, xbar, prop_ctx, params.n_exp_terms,
                                                    params.n_chunks, xbar_split=xs, **halves)
    else:
        xbar, shift_term, fb_term = kernels.field_shifts(fields, fb, prop_ctx.mf_shifts, prop_ctx.sqrt_dt)
        walkers_new = prop_ops.apply_trotter_planes(state.walkers, xbar, prop_ctx, params.n_exp_terms,
                                                    params.n_chunks, **halves)
    if meas_ops.greens is not None:
        make_greens = meas_ops.step_greens or meas_ops.greens
        greens_new = make_greens(walkers_new, trial_d if defer_out else trial_data)
        logdet = greens_new.logdet
    else:
        greens_new = None
        lo = _log_overlap_and_greens(walkers_new, meas_ops, trial_data)[0]
        logdet = lo.reshape(-1, 1) / 2.0  # restricted convention: log ovlp = 2 logdet
    lo_new, w_new, pop_new = kernels.phaseless_update(
        logdet, state.log_overlaps, shift_term, fb_term, state.weights,
        state.pop_control_ene_shift, e_estimate=state.e_estimate, h0_prop=prop_ctx.h0_prop,
        dt=prop_ctx.dt, sqrt_dt=prop_ctx.sqrt_dt,
        weight_floor=float(getattr(params, "weight_floor", 1.0e-3)),
        weight_cap=float(getattr(params, "weight_cap", 100.0)),
        damping=float(getattr(params, "pop_control_damping", 0.1)))
    return prop_state(walkers=walkers_new, weights=w_new, log_overlaps=lo_new, rng_key=key,
                      pop_control_ene_shift=pop_new, e_estimate=state.e_estimate,
                      greens=greens_new, deferred=defer_out)
