"""Propagation state and parameters (prop/types.py:10-32 of the reference)."""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any, NamedTuple, Optional

import torch


class prop_state(NamedTuple):
    """Walker-step state.

    Differences from the reference's prop_state: the overlaps are carried as
    complex log-overlaps (``log_overlaps``; the ``overlaps`` property returns
    exp(log_overlaps), equal to the reference's raw det^2 wherever that is
    finite), ``pop_control_ene_shift`` is a 0-d device tensor (so a step never
    synchronizes with the host), ``e_estimate`` is a Python float, and
    ``greens`` caches the Green's functions / log-determinants of ``walkers``
    (None when they must be recomputed).  ``deferred``: the last step left out
    its final one-body half (afqmc_step(defer_final=True)), so ``walkers`` holds
    E^-1 W (E = exp(-dt h1_eff / 2)) and ``greens`` the matching G E^-1; only
    the next afqmc_step reads such a state (it applies E^2 first)."""

    walkers: Any
    weights: torch.Tensor
    log_overlaps: torch.Tensor
    rng_key: torch.Tensor
    pop_control_ene_shift: torch.Tensor
    e_estimate: float
    greens: Optional[Any] = None
    deferred: bool = False

    @property
    def overlaps(self) -> torch.Tensor:
        return torch.exp(self.log_overlaps)


@dataclass(frozen=True)
class afqmc_params:
    """Same knobs and defaults as the reference, plus ``n_ortho``: orthonormalize
    every n_ortho steps inside a block (None: only at the end of the block, as
    in prop/blocks.py:53), and ``merge_one_body_halves``: inside a block, a step
    followed by another step leaves its final exp(-dt h1_eff / 2) to the next
    one, which applies exp(-dt h1_eff) once (where the propagator supports it:
    the bf16x3 upper-block Trotter; the walkers after the block are the same
    up to rounding)."""

    dt: float = 0.005
    n_chunks: int = 1
    n_exp_terms: int = 6
    pop_control_damping: float = 0.1
    weight_floor: float = 1.0e-3
    weight_cap: float = 100.0
    n_prop_steps: int = 50
    shift_ema: float = 0.1
    n_eql_blocks: int = 50
    n_blocks: int = 500
    n_walkers: int = 200
    seed: int = 42
    n_ortho: Optional[int] = None
    merge_one_body_halves: bool = True
