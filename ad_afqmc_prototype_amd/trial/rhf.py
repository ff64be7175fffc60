"""RHF / UHF trial determinants (trial/rhf.py:12-71 of the reference).

Overlaps are computed by the HIP Green's-function kernel (LU of C^T w), which
carries the determinant as a complex logarithm; ``overlap_*`` return
exp(log-overlap), identical to the reference's raw det(C^T w)^2 (restricted) or
det(C^T wu) det(C^T wd) (unrestricted) wherever those are finite.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import NamedTuple

import torch

from .. import kernels
from ..core.ops import trial_ops
from ..core.system import system
from ..ham.chol import as_device_f64


def as_device_orbitals(x, device=None) -> torch.Tensor:
    """Trial orbitals on the device: complex128 when any imaginary part is
    nonzero (the reference takes C^H, trial/rhf.py:25-44), else float64."""
    t = torch.as_tensor(x)
    if t.is_complex() and bool(torch.any(t.imag != 0)):
        if device is None:
            device = t.device if t.device.type == "cuda" else torch.device("cuda", 0)
        return t.to(device=device, dtype=torch.complex128).contiguous()
    return as_device_f64(t.real if t.is_complex() else t, device)


@dataclass(frozen=True)
class rhf_trial:
    mo_coeff: torch.Tensor  # (norb, nocc), real or complex

    def __post_init__(self):
        object.__setattr__(self, "mo_coeff", as_device_orbitals(self.mo_coeff))

    @property
    def norb(self) -> int:
        return int(self.mo_coeff.shape[0])

    @property
    def nocc(self) -> int:
        return int(self.mo_coeff.shape[1])

    @property
    def mo_up(self) -> torch.Tensor:
        return self.mo_coeff

    @property
    def mo_dn(self) -> torch.Tensor:
        return self.mo_coeff

    @property
    def restricted(self) -> bool:
        return True

    @property
    def is_complex(self) -> bool:
        return self.mo_coeff.is_complex()


@dataclass(frozen=True)
class uhf_trial:
    """Spin-unrestricted trial (no counterpart in the reference, which only has
    rhf_trial): one (norb, nocc) real orbital block per spin."""

    mo_up: torch.Tensor
    mo_dn: torch.Tensor

    def __post_init__(self):
        up = as_device_orbitals(self.mo_up)
        dn = as_device_orbitals(self.mo_dn, up.device)
        if up.is_complex() or dn.is_complex():
            up, dn = up.to(torch.complex128), dn.to(torch.complex128)
        object.__setattr__(self, "mo_up", up)
        object.__setattr__(self, "mo_dn", dn)
        if self.mo_up.shape != self.mo_dn.shape:
            raise ValueError("uhf_trial: both spins need the same (norb, nocc)")

    @property
    def norb(self) -> int:
        return int(self.mo_up.shape[0])

    @property
    def nocc(self) -> int:
        return int(self.mo_up.shape[1])

    @property
    def restricted(self) -> bool:
        return False

    @property
    def is_complex(self) -> bool:
        return self.mo_up.is_complex()


class greens_cache(NamedTuple):
    """Half-rotated Green's functions (split planes (2, nw, nspin*o*n)) and
    log det(C_s^T w_s) (nw, nspin) of one walker batch; optionally the
    (hi, lo) bf16 planes of the greens rows for the bf16x3 force bias.
    ``greens`` is None when only the bf16 planes were made (the bf16x3 step's
    cache): consumers that need the fp64 rows recompute them."""

    greens: torch.Tensor | None
    logdet: torch.Tensor
    split: tuple | None = None

    @property
    def log_overlap(self) -> torch.Tensor:
        if self.logdet.shape[1] == 1:
            return 2.0 * self.logdet[:, 0]
        return self.logdet[:, 0] + self.logdet[:, 1]


def get_rdm1(trial_data) -> torch.Tensor:
    """trial/rhf.py:29-32: stack of C C^T per spin, (2, n, n)."""
    cu, cd = trial_data.mo_up, trial_data.mo_dn
    return torch.stack([cu @ cu.conj().T, cd @ cd.conj().T], dim=0)


def greens(walkers, trial_data, split: bool = False) -> greens_cache:
    """``split`` also returns the bf16 planes of the rows (same kernel pass)."""
    if isinstance(walkers, tuple):
        out = kernels.overlap_greens(walkers, trial_data.mo_up, trial_data.mo_dn, split=split)
    else:
        if not trial_data.restricted:
            raise ValueError("restricted walkers need an RHF trial")
        out = kernels.overlap_greens(walkers, trial_data.mo_up, split=split)
    return greens_cache(*out)


def greens_split(walkers, trial_data) -> greens_cache:
    """The bf16x3 step's Green's function: log det and the bf16 planes of the
    rows only (o <= 32; larger o takes greens(..., split=True))."""
    w0 = walkers[0] if isinstance(walkers, tuple) else walkers
    if w0.shape[2] > 32 or trial_data.mo_up.is_complex():
        return greens(walkers, trial_data, split=True)
    if isinstance(walkers, tuple):
        ld, sp = kernels.overlap_greens_bf16x3(walkers, trial_data.mo_up, trial_data.mo_dn)
    else:
        if not trial_data.restricted:
            raise ValueError("restricted walkers need an RHF trial")
        ld, sp = kernels.overlap_greens_bf16x3(walkers, trial_data.mo_up)
    return greens_cache(None, ld, sp)


def log_overlap(walkers, trial_data) -> torch.Tensor:
    return greens(walkers, trial_data).log_overlap


def block_trial(trial_data) -> torch.Tensor:
    """diag(C, C) (2n, 2o): the RHF trial in the spin-orbital basis."""
    c = trial_data.mo_coeff
    n, o = c.shape
    out = torch.zeros((2 * n, 2 * o), dtype=c.dtype, device=c.device)
    out[:n, :o] = c
    out[n:, o:] = c
    return out


def greens_g(walkers, trial_data) -> greens_cache:
    """Generalized walkers (nw, 2n, 2o) against an RHF trial (trial/rhf.py:47-53,
    meas/rhf.py:108-130): with the trial diag(C, C) the overlap matrix is the
    stacked [C^T w_top; C^T w_bot] and G = (w m^-1)^T is (2o, 2n).  The overlap
    is det m (not squared): logdet is carried as (nw, 2) = [log det m, 0], so
    the spin sum of log_overlap and of the phaseless update gives log det m."""
    if not trial_data.restricted:
        raise ValueError("generalized walkers need an RHF trial")
    g, ld = kernels.overlap_greens(walkers, block_trial(trial_data))
    return greens_cache(g, torch.cat([ld, torch.zeros_like(ld)], dim=1))


def log_overlap_g(walkers, trial_data) -> torch.Tensor:
    return greens_g(walkers, trial_data).log_overlap


def overlap_g(walkers, trial_data) -> torch.Tensor:
    """det([C^T w_top; C^T w_bot]) per walker (trial/rhf.py:47-53)."""
    return torch.exp(log_overlap_g(walkers, trial_data))


def overlap_r(walkers, trial_data) -> torch.Tensor:
    """det(C^T w)^2 per walker (trial/rhf.py:35-37)."""
    return torch.exp(log_overlap(walkers, trial_data))


def overlap_u(walkers, trial_data) -> torch.Tensor:
    """det(C_up^T wu) det(C_dn^T wd) per walker (trial/rhf.py:40-44)."""
    return torch.exp(log_overlap(walkers, trial_data))


def make_rhf_trial_ops(sys: system) -> trial_ops:
    """trial/rhf.py:56-71."""
    if sys.nup != sys.ndn:
        raise ValueError("RHF requires nelec[0] == nelec[1].")
    wk = sys.walker_kind.lower()
    if wk == "restricted":
        return trial_ops(overlap=overlap_r, get_rdm1=get_rdm1, greens=greens, log_overlap=log_overlap)
    if wk == "unrestricted":
        return trial_ops(overlap=overlap_u, get_rdm1=get_rdm1, greens=greens, log_overlap=log_overlap)
    if wk == "generalized":
        return trial_ops(overlap=overlap_g, get_rdm1=get_rdm1, greens=greens_g, log_overlap=log_overlap_g)
    raise ValueError(f"unknown walker_kind: {sys.walker_kind}")


def make_uhf_trial_ops(sys: system) -> trial_ops:
    """UHF trial with unrestricted walkers."""
    if sys.walker_kind.lower() != "unrestricted":
        raise ValueError("a UHF trial needs walker_kind='unrestricted'")
    return trial_ops(overlap=overlap_u, get_rdm1=get_rdm1, greens=greens, log_overlap=log_overlap)
