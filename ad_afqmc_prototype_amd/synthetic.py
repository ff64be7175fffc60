"""Seeded synthetic Cholesky Hamiltonians and auxiliary-field streams.

The reference has no synthetic generator of its own; its examples build a
Hamiltonian from pyscf integrals (default.py:15-47, integrals.py:3-46).  The
workloads measured here are synthetic, so the generator below fixes the PRNG
and the order of draws once for all tests and benchmarks:

1. ``A`` (naux, n, n) ~ N(0,1) * 0.5 * exp(-4 g / naux);  L_g = (A_g + A_g^T) / 2
2. ``M`` (n, n) ~ N(0,1);  h1 = diag(linspace(-2, 2, n)) + 0.1 * (M + M^T) / 2
3. h0 = 0

all from ``numpy.random.default_rng(seed)``.  The RHF trial holds the lowest
``nocc`` eigenvectors of h1.  The UHF trial takes the lowest ``nocc``
eigenvectors of h1 +/- split * diag(linspace(-1, 1, n)) (a non-uniform
splitting, so that the two spin sectors really differ).

Auxiliary fields for step ``t`` are x ~ N(0,1) of shape (n_walkers, naux)
drawn from ``default_rng([seed, 1, t])``; comb uniforms for reconfiguration
event ``k`` come from ``default_rng([seed, 2, k])``.  The same arrays are fed to
the HIP path and to the CPU oracle.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class synthetic_system:
    h0: float
    h1: np.ndarray          # (n, n) float64
    chol: np.ndarray        # (naux, n, n) float64
    mo_up: np.ndarray       # (n, nocc) float64
    mo_dn: np.ndarray       # (n, nocc) float64
    restricted_trial: bool  # True: RHF trial (mo_up is mo_dn)

    @property
    def norb(self) -> int:
        return int(self.h1.shape[0])

    @property
    def nocc(self) -> int:
        return int(self.mo_up.shape[1])

    @property
    def naux(self) -> int:
        return int(self.chol.shape[0])


def make_hamiltonian(seed: int, norb: int, naux: int, *, chol_scale: float = 1.0):
    """(h0, h1, chol) of the synthetic generator.  ``chol_scale`` != 1 gives the
    explicitly rescaled variant used for energy-accuracy studies."""
    rng = np.random.default_rng(seed)
    decay = 0.5 * np.exp(-4.0 * np.arange(naux) / naux)
    a = rng.standard_normal((naux, norb, norb)) * decay[:, None, None]
    chol = 0.5 * (a + a.transpose(0, 2, 1))
    if chol_scale != 1.0:
        chol = chol * chol_scale
    m = rng.standard_normal((norb, norb))
    h1 = np.diag(np.linspace(-2.0, 2.0, norb)) + 0.1 * 0.5 * (m + m.T)
    return 0.0, h1, chol


def make_system(seed: int, norb: int, nocc: int, naux: int, *, uhf: bool = False,
                split: float = 0.05, chol_scale: float = 1.0) -> synthetic_system:
    h0, h1, chol = make_hamiltonian(seed, norb, naux, chol_scale=chol_scale)
    if not uhf:
        c = np.linalg.eigh(h1)[1][:, :nocc].copy()
        return synthetic_system(h0, h1, chol, c, c, True)
    d = np.diag(np.linspace(-1.0, 1.0, norb))
    cu = np.linalg.eigh(h1 + split * d)[1][:, :nocc].copy()
    cd = np.linalg.eigh(h1 - split * d)[1][:, :nocc].copy()
    return synthetic_system(h0, h1, chol, cu, cd, False)


def fields_for_step(seed: int, step: int, n_walkers: int, naux: int) -> np.ndarray:
    return np.random.default_rng([seed, 1, step]).standard_normal((n_walkers, naux))


def zeta_for_event(seed: int, event: int) -> float:
    return float(np.random.default_rng([seed, 2, event]).random())


# Workloads named by what they are (norb, nocc per spin, naux, walkers, ...).
# ``chol_scale`` puts each workload's walk in the working regime of the
# order-6 Taylor step: sqrt(dt) * ||sum_g x_g L_g||_2 is about 0.2 for a
# typical field draw, as for molecular Hamiltonians of these sizes.  At scale
# 1 the synthetic Hamiltonians give 0.86 (24 orbitals) to 4.6-7.8 (96-200
# orbitals): the larger ones make the truncated exponential meaningless, and
# even the 24-orbital walk leaves the sqrt(2/dt) energy window of the block
# average within a few steps, so every block energy is clipped to the fixed
# estimate.
WORKLOADS = {
    "n24_o6_g96_w128_rhf": dict(seed=0, norb=24, nocc=6, naux=96, n_walkers=128,
                                dt=0.01, n_steps=40, qr_every=5, sr_every=10,
                                uhf=False, walker_kind="restricted", chol_scale=0.25),
    "n120_o20_g600_w1024_uhf": dict(seed=1, norb=120, nocc=20, naux=600, n_walkers=1024,
                                    dt=0.005, n_steps=200, qr_every=10, sr_every=20,
                                    uhf=True, walker_kind="unrestricted", chol_scale=0.05),
    "n200_o32_g1000_w2048_rhf": dict(seed=2, norb=200, nocc=32, naux=1000, n_walkers=2048,
                                     dt=0.005, n_steps=400, qr_every=10, sr_every=20,
                                     uhf=False, walker_kind="restricted", chol_scale=0.03),
    "n96_o16_g480_w512_rhf": dict(seed=3, norb=96, nocc=16, naux=480, n_walkers=512,
                                  dt=0.01, n_steps=50, qr_every=10, sr_every=10,
                                  uhf=False, walker_kind="restricted", chol_scale=0.05),
}


def uneven_batch_specs(n_systems: int = 64, first_seed: int = 100):
    """Shapes of the batched uneven-system workload: N ~ U[40,160],
    nocc ~ U[6,24], naux = round(N * U[4,6]), walkers ~ U[64,512]."""
    specs = []
    for s in range(first_seed, first_seed + n_systems):
        rng = np.random.default_rng([s, 7])
        n = int(rng.integers(40, 161))
        o = int(rng.integers(6, 25))
        o = min(o, n)
        g = int(round(n * rng.uniform(4.0, 6.0)))
        nw = int(rng.integers(64, 513))
        specs.append(dict(seed=s, norb=n, nocc=o, naux=g, n_walkers=nw))
    return specs
