"""Front end: a restricted Hartree-Fock run from molecular integrals
(default.py:15-47 of the reference).

Rhf(mf) mirrors the reference's constructor on a pyscf mean-field object: MO
one-body integrals, 4-fold ERIs, modified Cholesky (on the GPU), an RHF trial
that is the identity block in the MO basis, and the afqmc ops; kernel() runs
the driver.  pyscf is an optional dependency: Rhf(mf) raises ImportError
without it.  Rhf.from_integrals() takes the same integrals directly (h0, the
MO h1 and the 4-fold packed ERIs), with no pyscf involved.
"""
from __future__ import annotations

import numpy as np
import torch

from . import driver
from .core.system import system
from .ham.chol import ham_chol
from .ham.integrals import modified_cholesky
from .meas.rhf import make_rhf_meas_ops
from .prop.blocks import afqmc_block
from .prop.chol_afqmc_ops import make_chol_afqmc_ops
from .prop.types import afqmc_params
from .trial.rhf import make_rhf_trial_ops, rhf_trial


class Rhf:
    def __init__(self, mf=None, *, _integrals=None, device=None, precision: str = "fp64",
                 max_error: float = 1e-6, seed: int | None = None):
        if _integrals is None:
            try:
                from pyscf import ao2mo  # noqa: F401  (optional)
            except ImportError as exc:
                raise ImportError("Rhf(mf) needs pyscf for the integrals; use Rhf.from_integrals "
                                  "with precomputed MO integrals instead") from exc
            mol = mf.mol
            h0 = float(mf.energy_nuc())
            h1 = np.asarray(mf.mo_coeff.T @ mf.get_hcore() @ mf.mo_coeff)
            eri = ao2mo.restore(4, np.asarray(ao2mo.kernel(mol, mf.mo_coeff)), mol.nao)
            nelec = tuple(mol.nelec)
        else:
            h0, h1, eri, nelec = _integrals
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        norb = int(np.asarray(h1).shape[0])
        chol = modified_cholesky(eri, max_error=max_error, device=device)
        self.sys = system(norb=norb, nelec=(int(nelec[0]), int(nelec[1])), walker_kind="restricted")
        self.ham_data = ham_chol.from_arrays(h0, h1, chol, device=device)
        self.trial_data = rhf_trial(torch.eye(norb, self.sys.nup, dtype=torch.float64, device=device))
        self.trial_ops = make_rhf_trial_ops(sys=self.sys)
        self.meas_ops = make_rhf_meas_ops(sys=self.sys, precision=precision)
        self.prop_ops = make_chol_afqmc_ops(self.ham_data, self.sys.walker_kind, precision=precision)
        self.params = afqmc_params(n_eql_blocks=20, n_blocks=200,
                                   seed=int(np.random.randint(0, int(1e6)) if seed is None else seed))
        self.block_fn = afqmc_block

    @classmethod
    def from_integrals(cls, h0: float, h1, eri, nelec, **kw) -> "Rhf":
        """h1 (n, n) and the 4-fold ERIs (size, size) in the MO basis of the RHF
        determinant (its occupied orbitals are the first nelec[0] MOs)."""
        return cls(_integrals=(float(h0), h1, eri, nelec), **kw)

    def kernel(self, **kw):
        return driver.run_afqmc_energy(sys=self.sys, params=self.params, ham_data=self.ham_data,
                                       trial_ops=self.trial_ops, trial_data=self.trial_data,
                                       meas_ops=self.meas_ops, block_fn=self.block_fn, **kw)
