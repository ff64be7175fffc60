"""MI355X-native phaseless-AFQMC walker step (HIP kernels for gfx950).

Host side mirrors the plugin surface of ad_afqmc_prototype (system, ham_chol,
RHF/UHF trial, measurement ops, Cholesky-AFQMC propagator, walker container,
step / block / driver) on torch tensors; the compute runs in libafqmc_hip.so.
"""
from . import kernels  # noqa: F401
from ._lib import HipLibraryError, library_path, load  # noqa: F401

__all__ = ["kernels", "HipLibraryError", "library_path", "load"]
