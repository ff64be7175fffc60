"""CPU-side checks of the C ABI: the built library loads, exports every
function include/afqmc_hip.h declares, and rejects invalid arguments before
touching any device memory."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "afqmc_hip.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(afq_[a-z0-9_]+)\s*\(", text)))


def _lib():
    from ad_afqmc_prototype_amd import _lib as L
    if not os.path.exists(L.library_path()):
        pytest.skip("libafqmc_hip.so not built")
    return L


def test_header_and_binding_agree():
    from ad_afqmc_prototype_amd import _lib as L
    assert _declared() == sorted(L.exported_symbols())


def test_library_exports_every_declared_symbol():
    L = _lib()
    out = subprocess.run(["nm", "-D", "--defined-only", L.library_path()], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (afq_[a-z0-9_]+)$", out, flags=re.M))
    missing = [s for s in _declared() if s not in exported]
    assert not missing, missing


def test_library_loads_and_reports_abi():
    L = _lib()
    lib = L.load()
    assert lib.afq_abi_version() == L.ABI_VERSION
    assert isinstance(lib.afq_last_error(), bytes)


@pytest.mark.parametrize("name,args,msg", [
    ("afq_dgemm_nt", (-1, 2, 2, 1.0, 1, 2, 1, 2, 0.0, 1, 2, 0), "negative size"),
    ("afq_dgemm_nt", (4, 4, 4, 1.0, 1, 2, 1, 4, 0.0, 1, 4, 0), "leading dimension"),
    ("afq_dgemm_nt_splitk", (-1, 4, 4, 1.0, 16, 4, 16, 4, 16, 4, 2, 16, 0), "negative size"),
    ("afq_dgemm_nt_splitk", (4, 4, 64, 1.0, 16, 64, 16, 64, 16, 4, 0, 16, 0), "splits"),
    ("afq_dgemm_nt_splitk", (4, 4, 64, 1.0, 16, 64, 16, 64, 16, 4, 17, 16, 0), "splits"),
    ("afq_dgemm_nt_splitk", (4, 4, 64, 1.0, 16, 64, 16, 64, 16, 4, 2, 0, 0), "null"),
    ("afq_dgemm_nt_splitk", (4, 4, 64, 1.0, 16, 32, 16, 64, 16, 4, 2, 16, 0), "leading dimension"),
    ("afq_overlap_greens", (4, 8, 9, 1, 1, 0, 1, 0, 1, 1, 0), "bad sizes"),
    ("afq_overlap_greens", (4, 100, 70, 1, 1, 0, 1, 0, 1, 1, 0), "maximum of 64"),
    ("afq_overlap_greens", (4, 8, 2, 2, 1, 0, 1, 0, 1, 1, 0), "spin-down"),
    ("afq_overlap_greens_split", (4, 8, 2, 1, 1, 0, 1, 0, 1, 1, 1, 1, 15, 0), "ldk"),
    ("afq_overlap_greens_split", (4, 8, 2, 1, 1, 0, 1, 0, 1, 1, 1, 1, 80, 0), "ldk"),
    ("afq_overlap_greens_split", (4, 8, 2, 1, 1, 0, 1, 0, 1, 1, 0, 1, 16, 0), "null"),
    ("afq_vhs", (4, 10, 10, 5, 1, 1, 1, 0), "multiple of 4"),
    ("afq_vhs_sym", (4, 10, 10, 5, 0, 16, 0, 5, 16, 0, 5, 16, 0), "multiple of 4"),
    ("afq_field_shifts_split", (4, 10, 0.1, 1, 1, 1, 1, 1, 1, 1, 1, 8, 0), "ldk"),
    ("afq_field_shifts_split", (4, 10, 0.1, 1, 1, 1, 1, 1, 1, 0, 1, 16, 0), "null"),
    ("afq_trotter_bf16x3_sym", (1, 240, 8, 1, 6, 0.1, 1, 1, 2, 0, 3, 0, 0), "224"),
    ("afq_trotter_bf16x3_sym", (1, 8, 2, 1, 6, 0.1, 1, 1, 2, 0, 2, 0, 0), "aliases"),
    ("afq_gemm_bf16x3_nt_splitk", (4, 4, 8, 1.0, 16, 16, 8, 16, 16, 8, 16, 4, 1, 0, 16, 0), "splits"),
    ("afq_gemm_bf16x3_nt_splitk", (4, 4, 8, 1.0, 16, 16, 8, 16, 16, 8, 16, 4, 1, 2, 0, 0), "null"),
    ("afq_gemm_bf16x3_nt_splitk", (4, 4, 12, 1.0, 16, 16, 12, 16, 16, 12, 16, 4, 1, 2, 16, 0), "multiples of 8"),
    ("afq_vhs_sym", (4, 10, 12, 5, 3, 16, 0, 5, 16, 0, 5, 16, 0), "arith"),
    ("afq_vhs_sym", (4, 10, 12, 5, 0, 16, 0, 5, 16, 0, 5, 8, 0), "16-byte"),
    ("afq_vhs_sym", (4, 10, 12, 5, 2, 16, 16, 5, 16, 16, 5, 16, 0), "multiples of 8"),
    ("afq_trotter", (1, 400, 40, 1, 400, 6, 0.1, 1, 1, 2, 0, 3, 0, 0), "exceeds"),
    ("afq_trotter", (1, 8, 2, 1, 8, 6, 0.1, 1, 1, 2, 0, 2, 0, 0), "aliases"),
    ("afq_energy", (4, 50, 40, 1, 5, 0.0, 1, 1, 52, 1, 1, 0), "maximum of 32"),
    ("afq_orthonormalize", (4, 10, 3, 1, 1, 0), "aliased"),
    ("afq_phaseless_update", (4, 1, 1, 1, 1, 1, 1, 1, 0.0, 0.0, 0.0, -1.0, 0.1, 1e-3, 100.0, 0.1,
                              1, 1, 1, 0), "dt must be positive"),
    ("afq_comb_indices", (0, 1, 0.5, 1, 1, 1, 0), "nw=0"),
    ("afq_comb_indices_noalloc", (0, 1, 1, 1, 1, 0), "nw=0"),
    ("afq_comb_indices_noalloc", ((1 << 18) + 1, 1, 1, 1, 1, 0), "exceeds"),
    ("afq_comb_indices_noalloc", (4, 1, 0, 1, 1, 0), "null"),
    ("afq_gather_rows", (4, 0, 1, 0, 0, 1, 1, 2, 0), "bad sizes"),
])
def test_invalid_arguments_rejected_on_host(name, args, msg):
    L = _lib()
    with pytest.raises(ValueError, match=msg):
        L.call(name, *args)


def test_grouped_descriptor_layout_matches_header():
    """The host mirrors of afq_system_step / afq_gemm_problem have the C layout
    of include/afqmc_hip.h (compiled here with gcc)."""
    import ctypes
    import os
    import shutil
    import subprocess
    import tempfile
    from ad_afqmc_prototype_amd.prop import batched as B
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "afqmc_hip.h"
#define O(T, f) printf(#f " %zu\n", offsetof(T, f))
int main(void) {
  printf("size_step %zu\nsize_gemm %zu\n", sizeof(afq_system_step), sizeof(afq_gemm_problem));
  O(afq_system_step, tblock0); O(afq_system_step, sqrt_dt); O(afq_system_step, damping);
  O(afq_system_step, trial); O(afq_system_step, vhs); O(afq_system_step, pop_out);
  O(afq_system_step, key_out); O(afq_system_step, ldr); O(afq_system_step, h0);
  O(afq_system_step, rot_chol); O(afq_system_step, energy); O(afq_system_step, exch_z); O(afq_system_step, exp_h1);
  O(afq_gemm_problem, tile_start); O(afq_gemm_problem, ldc);
  O(afq_gemm_problem, C); O(afq_gemm_problem, alpha); O(afq_gemm_problem, c_fp64);
  O(afq_gemm_problem, b_lower);
  printf("size_gemm3 %zu\n", sizeof(afq_gemm_problem_bf16x3));
  printf("b_lower3 %zu\n", offsetof(afq_gemm_problem_bf16x3, b_lower));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "l.c"), "w") as f:
            f.write(src)
        exe = os.path.join(d, "l")
        subprocess.run([gcc, "-I", os.path.join(root, "include"), os.path.join(d, "l.c"), "-o", exe],
                       check=True)
        out = dict(line.split() for line in subprocess.run([exe], capture_output=True, text=True,
                                                           check=True).stdout.splitlines())
    assert int(out["size_step"]) == ctypes.sizeof(B._system_step)
    assert int(out["size_gemm"]) == ctypes.sizeof(B._gemm_problem)
    for f in ("tblock0", "sqrt_dt", "damping", "trial", "vhs", "pop_out", "key_out", "ldr", "h0",
              "rot_chol", "energy", "exch_z", "exp_h1"):
        assert int(out[f]) == getattr(B._system_step, f).offset, f
    for f in ("tile_start", "ldc", "C", "alpha", "c_fp64", "b_lower"):
        assert int(out[f]) == getattr(B._gemm_problem, f).offset, f
    assert int(out["size_gemm3"]) == ctypes.sizeof(B._gemm_problem_bf16x3)
    assert int(out["b_lower3"]) == B._gemm_problem_bf16x3.b_lower.offset
