"""Cholesky Hamiltonian (ham/chol.py:12-65 of the reference) on torch tensors."""
from __future__ import annotations

from dataclasses import dataclass
from typing import Literal

import torch

ham_basis = Literal["restricted", "generalized"]


def as_device_f64(x, device=None) -> torch.Tensor:
    """float64 tensor on a HIP device (default cuda:0)."""
    if device is None:
        device = x.device if isinstance(x, torch.Tensor) and x.device.type == "cuda" else torch.device("cuda", 0)
    t = torch.as_tensor(x)
    if t.is_complex():
        if torch.any(t.imag != 0):
            raise ValueError("complex-valued integrals are not supported by the HIP path")
        t = t.real
    return t.to(device=device, dtype=torch.float64).contiguous()


@dataclass(frozen=True)
class ham_chol:
    """h0 scalar, h1 (n, n), chol (n_fields, n, n) in the restricted basis, or
    h1 (nso, nso), chol (n_fields, nso, nso) in a generalized (spin-orbital)
    basis, nso = 2 norb (ham/chol.py:12-61)."""

    h0: float
    h1: torch.Tensor
    chol: torch.Tensor
    basis: ham_basis = "restricted"

    def __post_init__(self):
        if self.basis not in ("restricted", "generalized"):
            raise ValueError(f"unknown basis: {self.basis}")
        if getattr(self.h1, "ndim", None) != 2:
            raise ValueError(f"h1 must be rank-2; got ndim={getattr(self.h1, 'ndim', None)}")
        if self.h1.shape[0] != self.h1.shape[1]:
            raise ValueError(f"h1 must be square; got {tuple(self.h1.shape)}")
        if getattr(self.chol, "ndim", None) != 3:
            raise ValueError(f"chol must be rank-3; got ndim={getattr(self.chol, 'ndim', None)}")
        if tuple(self.chol.shape[1:]) != tuple(self.h1.shape):
            raise ValueError(
                f"chol trailing dims must match h1; got chol {tuple(self.chol.shape)}, h1 {tuple(self.h1.shape)}")

    @staticmethod
    def from_arrays(h0, h1, chol, basis: ham_basis = "restricted", device=None) -> "ham_chol":
        return ham_chol(h0=float(h0), h1=as_device_f64(h1, device), chol=as_device_f64(chol, device),
                        basis=basis)

    @property
    def device(self) -> torch.device:
        return self.chol.device


def n_fields(ham: ham_chol) -> int:
    return int(ham.chol.shape[0])
