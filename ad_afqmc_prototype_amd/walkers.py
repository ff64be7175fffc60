"""Walker container operations (walkers.py:13-265 of the reference), batched
on the HIP kernels.  Walkers are complex128 tensors (nw, n, o) (restricted),
(nw, 2n, nup + ndn) (generalized, spin-orbital rows), or a tuple of two
(nw, n, o) tensors (unrestricted)."""
from __future__ import annotations

from typing import Any, Callable

import torch

from . import kernels
from .core.system import system


def _natorbs(dm: torch.Tensor, n_occ: int) -> torch.Tensor:
    """walkers.py:13-16."""
    dm = 0.5 * (dm + dm.conj().T)
    vecs = torch.linalg.eigh(dm)[1].flip(-1)
    return vecs[:, :n_occ]


def init_walkers(sys: system, rdm1: torch.Tensor, n_walkers: int):
    """walkers.py:19-77: natural orbitals of the trial rdm1, identical for every
    walker.  Generalized walkers (2n, nup + ndn): from a (2n, 2n) rdm1, or from
    the spin blocks as [[up, 0], [0, dn]]."""
    wk = sys.walker_kind.lower()

    def bcast(c):
        c = c.to(torch.complex128)
        return c.unsqueeze(0).expand(n_walkers, *c.shape).contiguous()

    if wk == "generalized":
        n, nup, ndn = sys.norb, sys.nup, sys.ndn
        if rdm1.ndim == 2:
            if tuple(rdm1.shape) != (2 * n, 2 * n):
                raise ValueError("For generalized walkers, a 2D rdm1 must have shape (2*norb, 2*norb).")
            return bcast(_natorbs(rdm1, nup + ndn))
        if rdm1.ndim != 3 or rdm1.shape[0] != 2:
            raise ValueError("Expected rdm1 with shape (2, norb, norb) for generalized init from spin blocks.")
        w0 = torch.zeros((2 * n, nup + ndn), dtype=torch.complex128, device=rdm1.device)
        w0[:n, :nup] = _natorbs(rdm1[0], nup)
        w0[n:, nup:] = _natorbs(rdm1[1], ndn)
        return bcast(w0)
    if rdm1.ndim == 2:
        raise ValueError("For walker_kind in {'restricted','unrestricted'}, rdm1 must be spin-block (2, norb, norb).")
    if rdm1.ndim != 3 or rdm1.shape[0] != 2:
        raise ValueError("Expected rdm1 with shape (2, norb, norb).")

    if wk == "restricted":
        return bcast(_natorbs(rdm1[0] + rdm1[1], sys.nup))
    if wk == "unrestricted":
        return (bcast(_natorbs(rdm1[0], sys.nup)), bcast(_natorbs(rdm1[1], sys.ndn)))
    raise ValueError(f"unknown walker_kind: {wk}")


def is_unrestricted(w) -> bool:
    return isinstance(w, tuple) and len(w) == 2


def n_walkers(w) -> int:
    return w[0].shape[0] if is_unrestricted(w) else w.shape[0]


def _chunk_size(nw: int, n_chunks: int) -> int:
    if nw % n_chunks != 0:
        raise ValueError(f"n_walkers={nw} is not divisible into n_chunks={n_chunks}")
    return nw // n_chunks


def _take(w, a, b):
    return (w[0][a:b], w[1][a:b]) if is_unrestricted(w) else w[a:b]


def apply_chunked(w, apply_fn: Callable, n_chunks: int, *args, **kwargs):
    """walkers.py:94-127: apply a batched kernel to sequential chunks of walkers."""
    nw = n_walkers(w)
    if n_chunks == 1:
        return apply_fn(w, *args, **kwargs)
    cs = _chunk_size(nw, n_chunks)
    outs = [apply_fn(_take(w, i * cs, (i + 1) * cs), *args, **kwargs) for i in range(n_chunks)]
    return torch.cat(outs, dim=0)


def apply_chunked_prop(w, fields, prop_fn: Callable, n_chunks: int, *args, **kwargs):
    """walkers.py:130-166: apply a batched propagator to sequential chunks."""
    nw = n_walkers(w)
    if n_chunks == 1:
        return prop_fn(w, fields, *args, **kwargs)
    cs = _chunk_size(nw, n_chunks)
    outs = [prop_fn(_take(w, i * cs, (i + 1) * cs), fields[i * cs:(i + 1) * cs], *args, **kwargs)
            for i in range(n_chunks)]
    if is_unrestricted(w):
        return (torch.cat([o[0] for o in outs]), torch.cat([o[1] for o in outs]))
    return torch.cat(outs, dim=0)


def orthonormalize(w, walker_kind: str):
    """walkers.py:206-211: reduced QR, diag(R) > 0, R discarded."""
    wk = walker_kind.lower()
    if wk == "unrestricted":
        return (kernels.orthonormalize(w[0]), kernels.orthonormalize(w[1]))
    if wk in ("restricted", "generalized"):  # one (n, o) or (2n, 2o) matrix per walker
        return kernels.orthonormalize(w)
    raise ValueError(f"unknown walker_kind: {walker_kind}")


def sr_indices(weights: torch.Tensor, zeta: float) -> torch.Tensor:
    """walkers.py:233-240 (int32 indices; sequential cumulative sum)."""
    return kernels.comb_indices(weights, zeta)[0]


def stochastic_reconfiguration_with_indices(w, weights: torch.Tensor, zeta, walker_kind: str):
    """Comb resampling with the selected indices; zeta a float or a 0-d float64
    device tensor."""
    wk = walker_kind.lower()
    if wk not in ("restricted", "unrestricted", "generalized"):
        raise ValueError(f"unknown walker_kind: {walker_kind}")
    idx, w_new = kernels.comb_indices(weights, zeta)
    if wk == "unrestricted":
        return (kernels.gather_rows(w[0], idx), kernels.gather_rows(w[1], idx)), w_new, idx
    return kernels.gather_rows(w, idx), w_new, idx


def stochastic_reconfiguration(w, weights: torch.Tensor, zeta, walker_kind: str):
    """walkers.py:243-265: comb resampling, weights reset to the mean |w|."""
    w_sr, w_new, _ = stochastic_reconfiguration_with_indices(w, weights, zeta, walker_kind)
    return w_sr, w_new
