"""Plugin protocols (core/ops.py:11-90 of the reference), batched over walkers.

Kernels take the whole walker batch -- an array (nw, n, o) or a tuple of two
for unrestricted walkers -- and return one value per walker, instead of the
per-walker functions the reference vmaps (walkers.py:94-166).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Callable, Mapping, NamedTuple, Optional


class trial_ops(NamedTuple):
    """overlap: (walkers, trial_data) -> (nw,) complex; get_rdm1: trial rdm1 (2, n, n);
    log_overlap: (walkers, trial_data) -> (nw,) complex log-overlaps."""

    overlap: Callable[[Any, Any], Any]
    get_rdm1: Callable[[Any], Any]
    greens: Optional[Callable[[Any, Any], Any]] = None
    log_overlap: Optional[Callable[[Any, Any], Any]] = None


class ham_ops(NamedTuple):
    n_fields: Callable[[Any], int]


meas_kernel = Callable[..., Any]

k_energy = "energy"
k_force_bias = "force_bias"


@dataclass(frozen=True)
class meas_ops:
    """Measurement ops: trial overlap, context builder and named kernels.

    ``greens`` (walkers, trial_data) -> greens_cache and ``log_overlap`` are the
    batched device entry points the fused step uses; kernels accept an optional
    ``greens=`` cache so the Green's function of the current walkers is formed
    once per step."""

    overlap: Callable[[Any, Any], Any]
    build_meas_ctx: Callable[[Any, Any], Any] = lambda ham_data, trial_data: None
    kernels: Mapping[str, meas_kernel] = field(default_factory=dict)
    observables: Mapping[str, meas_kernel] = field(default_factory=dict)
    log_overlap: Optional[Callable[[Any, Any], Any]] = None
    greens: Optional[Callable[[Any, Any], Any]] = None
    # (walkers, meas_ctx, trial_data, greens) -> force bias as float64 planes
    # (2, nw, naux): the same values as the force_bias kernel without the
    # complex tensor in between (the step consumes planes)
    force_bias_planes: Optional[Callable[..., Any]] = None
    # (walkers, trial_data) -> greens_cache of the step's new walkers when the
    # step needs less than ``greens`` (the bf16x3 step: the force bias's bf16
    # planes and log det, no fp64 rows); None: the step uses ``greens``
    step_greens: Optional[Callable[[Any, Any], Any]] = None

    def has_kernel(self, name: str) -> bool:
        return name in self.kernels

    def has_observable(self, name: str) -> bool:
        return name in self.observables

    def require_kernel(self, name: str) -> meas_kernel:
        try:
            return self.kernels[name]
        except KeyError as e:
            avail = ", ".join(sorted(self.kernels.keys()))
            raise KeyError(f"missing required kernel '{name}'. available: [{avail}]") from e

    def require_observable(self, name: str) -> meas_kernel:
        try:
            return self.observables[name]
        except KeyError as e:
            avail = ", ".join(sorted(self.observables.keys()))
            raise KeyError(f"missing requested observable '{name}'. available: [{avail}]") from e

    def available_kernels(self) -> tuple[str, ...]:
        return tuple(sorted(self.kernels.keys()))

    def available_observables(self) -> tuple[str, ...]:
        return tuple(sorted(self.observables.keys()))
