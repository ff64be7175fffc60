"""Loader for libafqmc_hip.so (the gfx950 kernels behind include/afqmc_hip.h).

There is no CPU fallback: if the library is missing or no HIP device is
present, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libafqmc_hip.so")
_lock = threading.Lock()
_lib = None

c_int = ctypes.c_int
c_ll = ctypes.c_longlong
c_double = ctypes.c_double
c_vp = ctypes.c_void_p

# name -> argtypes (restype is int for every compute entry point)
_SIGNATURES = {
    "afq_dgemm_nt": [c_int, c_int, c_int, c_double, c_vp, c_ll, c_vp, c_ll, c_double, c_vp, c_ll, c_vp],
    "afq_dgemm_nt_lower": [c_int, c_int, c_int, c_double, c_vp, c_ll, c_vp, c_ll, c_vp, c_ll, c_vp],
    "afq_sgemm_nt_lower": [c_int, c_int, c_int, ctypes.c_float, c_vp, c_ll, c_vp, c_ll, c_vp, c_ll, c_int, c_vp],
    "afq_gemm_bf16x3_nt_lower": [c_int, c_int, c_int, ctypes.c_float, c_vp, c_vp, c_ll, c_vp, c_vp, c_ll, c_vp,
                                 c_ll, c_int, c_vp],
    "afq_dgemm_nt_splitk": [c_int, c_int, c_int, c_double, c_vp, c_ll, c_vp, c_ll, c_vp, c_ll, c_int, c_vp, c_vp],
    "afq_sgemm_nt": [c_int, c_int, c_int, ctypes.c_float, c_vp, c_ll, c_vp, c_ll, ctypes.c_float, c_vp,
                     c_ll, c_int, c_vp],
    "afq_gemm_bf16x3_nt": [c_int, c_int, c_int, ctypes.c_float, c_vp, c_vp, c_ll, c_vp, c_vp, c_ll, c_vp,
                           c_ll, c_int, c_vp],
    "afq_gemm_bf16x3_nt_splitk": [c_int, c_int, c_int, ctypes.c_float, c_vp, c_vp, c_ll, c_vp, c_vp, c_ll, c_vp,
                                  c_ll, c_int, c_int, c_vp, c_vp],
    "afq_split_bf16": [c_int, c_int, c_vp, c_ll, c_vp, c_vp, c_ll, c_vp],
    "afq_convert_f32": [c_int, c_int, c_vp, c_ll, c_vp, c_ll, c_vp],
    "afq_trotter_f32": [c_int, c_int, c_int, c_int, c_int, c_int, c_double, c_vp, c_vp, c_vp, c_vp,
                        c_vp, c_vp, c_vp],
    "afq_trotter_bf16x3": [c_int, c_int, c_int, c_int, c_int, c_int, c_double, c_vp, c_vp, c_vp,
                           c_vp, c_vp, c_vp, c_vp],
    "afq_trotter_bf16x3_sym": [c_int, c_int, c_int, c_int, c_int, c_double, c_vp, c_vp, c_vp, c_vp, c_vp,
                               c_vp, c_vp],
    "afq_trotter_bf16x3_sym_halves": [c_int, c_int, c_int, c_int, c_int, c_double, c_vp, c_vp, c_vp, c_vp,
                                      c_vp, c_vp, c_vp, c_vp],
    "afq_overlap_greens": [c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "afq_overlap_greens_split": [c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                 c_ll, c_vp],
    "afq_overlap_greens_ctrial": [c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "afq_overlap_greens_bf16x3": [c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_ll, c_vp],
    "afq_force_bias": [c_int, c_int, c_int, c_double, c_vp, c_vp, c_vp, c_vp],
    "afq_field_shifts": [c_int, c_int, c_double, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "afq_field_shifts_split": [c_int, c_int, c_double, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_ll,
                               c_vp],
    "afq_vhs": [c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp],
    "afq_vhs_sym": [c_int, c_int, c_int, c_int, c_int, c_vp, c_vp, c_ll, c_vp, c_vp, c_ll, c_vp, c_vp],
    "afq_trotter": [c_int, c_int, c_int, c_int, c_int, c_int, c_double, c_vp, c_vp, c_vp, c_vp,
                    c_vp, c_vp, c_vp],
    "afq_trotter_adjoint": [c_int, c_int, c_int, c_int, c_int, c_double] + [c_vp] * 10,
    "afq_greens_adjoint": [c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_vp],
    "afq_phaseless_update": [c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_double, c_double,
                             c_double, c_double, c_double, c_double, c_double, c_double, c_vp,
                             c_vp, c_vp, c_vp],
    "afq_orthonormalize": [c_int, c_int, c_int, c_vp, c_vp, c_vp],
    "afq_energy": [c_int, c_int, c_int, c_int, c_int, c_double, c_vp, c_vp, c_int, c_vp, c_vp, c_vp],
    "afq_energy_f32": [c_int, c_int, c_int, c_int, c_int, c_double, c_vp, c_vp, c_int, c_vp, c_vp, c_vp],
    "afq_energy_bf16x3": [c_int, c_int, c_int, c_int, c_int, c_double, c_vp, c_vp, c_vp, c_int, c_vp,
                          c_vp, c_vp],
    "afq_block_energy": [c_int, c_vp, c_vp, c_double, c_double, c_double, c_vp, c_vp, c_vp],
    "afq_comb_indices": [c_int, c_vp, c_double, c_vp, c_vp, c_vp, c_vp],
    "afq_comb_indices_dev": [c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "afq_comb_indices_noalloc": [c_int, c_vp, c_vp, c_vp, c_vp, c_vp],
    "afq_rng_split": [c_vp, c_vp, c_vp, c_vp],
    "afq_rng_normal": [c_vp, c_ll, c_vp, c_vp],
    "afq_rng_uniform": [c_vp, c_vp, c_vp],
    "afq_threefry2x32": [ctypes.c_uint, ctypes.c_uint, c_ll, c_vp, c_vp, c_vp, c_vp, c_vp],
    "afq_gather_rows": [c_int, c_ll, c_int, c_ll, c_ll, c_vp, c_vp, c_vp, c_vp],
    # grouped launches over many systems (descriptor tables in device memory)
    "afq_gemm_tile_shape": [c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int)],
    "afq_sgemm_nt_grouped": [c_int, c_vp, c_int, c_int, c_vp],
    "afq_gemm_bf16x3_nt_grouped": [c_int, c_vp, c_int, c_int, c_vp],
    "afq_gemm_bf16x3_tile_shape": [c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int)],
    "afq_split_bf16_grouped": [c_int, c_vp, c_ll, c_vp],
    "afq_rng_fields_grouped": [c_int, c_vp, c_int, c_vp],
    "afq_field_shifts_grouped": [c_int, c_vp, c_int, c_vp],
    "afq_trotter_f32_grouped": [c_int, c_vp, c_int, c_int, c_int, c_vp],
    "afq_trotter_bf16x3_grouped": [c_int, c_vp, c_int, c_int, c_int, c_vp],
    "afq_trotter_bf16x3_sym_grouped": [c_int, c_vp, c_int, c_int, c_int, c_vp],
    "afq_trotter_bf16x3_sym_grouped_halves": [c_int, c_vp, c_int, c_int, c_int, c_int, c_int, c_vp],
    "afq_overlap_greens_grouped": [c_int, c_vp, c_int, c_int, c_vp],
    "afq_overlap_greens_bf16x3_grouped": [c_int, c_vp, c_int, c_int, c_vp],
    "afq_phaseless_update_grouped": [c_int, c_vp, c_vp],
    "afq_block_end_grouped": [c_int, c_vp, c_double, c_double, c_vp, c_vp, c_vp, c_vp],
    "afq_reconfigure_gather_grouped": [c_int, c_vp, c_int, c_vp, c_vp, c_vp, c_vp],
    "afq_orthonormalize_grouped": [c_int, c_vp, c_int, c_int, c_int, c_vp],
    "afq_energy_f32_grouped": [c_int, c_vp, c_int, c_int, c_int, c_vp],
    "afq_energy_bf16x3_grouped": [c_int, c_vp, c_int, c_int, c_int, c_int, c_vp],
    "afq_energy_bf16x3_grouped_fits": [c_int, c_int, c_int],
    "afq_energy_assemble": [c_int, c_int, c_int, c_int, c_double, c_double, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "afq_energy_assemble_grouped": [c_int, c_vp, c_int, c_vp],
    # Hamiltonian ingestion (modified Cholesky)
    "afq_mcd_workspace": [c_int, ctypes.POINTER(c_ll), ctypes.POINTER(c_ll)],
    "afq_mcd_start": [c_int, c_vp, c_ll, c_double, c_int, c_vp, c_vp, c_vp, c_vp, c_vp],
    "afq_mcd_iterate": [c_int, c_vp, c_ll, c_double, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp],
    "afq_unpack_pairs": [c_int, c_int, c_vp, c_ll, c_vp, c_vp],
}

ABI_VERSION = 23


class HipLibraryError(RuntimeError):
    pass


def library_path() -> str:
    return _LIB_PATH


def load() -> ctypes.CDLL:
    """Load and type the library (no device work)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(_LIB_PATH):
            raise HipLibraryError(
                f"{_LIB_PATH} is missing: build it with "
                "`python -m ad_afqmc_prototype_amd.csrc.build` (there is no CPU fallback)")
        lib = ctypes.CDLL(_LIB_PATH)
        lib.afq_last_error.restype = ctypes.c_char_p
        lib.afq_last_error.argtypes = []
        lib.afq_abi_version.restype = c_int
        lib.afq_abi_version.argtypes = []
        if lib.afq_abi_version() != ABI_VERSION:
            raise HipLibraryError(f"libafqmc_hip ABI {lib.afq_abi_version()} != {ABI_VERSION}: rebuild")
        for name, args in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = c_int
        _lib = lib
        return lib


def exported_symbols() -> list[str]:
    return ["afq_last_error", "afq_abi_version", *_SIGNATURES.keys()]


def call(name: str, *args) -> None:
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.afq_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(msg)
        raise HipLibraryError(msg)
