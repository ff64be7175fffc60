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
    ("afq_overlap_greens", (4, 8, 9, 1, 1, 0, 1, 0, 1, 1, 0), "bad sizes"),
    ("afq_overlap_greens", (4, 100, 70, 1, 1, 0, 1, 0, 1, 1, 0), "maximum of 64"),
    ("afq_overlap_greens", (4, 8, 2, 2, 1, 0, 1, 0, 1, 1, 0), "spin-down"),
    ("afq_vhs", (4, 10, 10, 5, 1, 1, 1, 0), "multiple of 4"),
    ("afq_trotter", (1, 400, 40, 1, 400, 6, 0.1, 1, 1, 2, 0, 3, 0, 0), "exceeds"),
    ("afq_trotter", (1, 8, 2, 1, 8, 6, 0.1, 1, 1, 2, 0, 2, 0, 0), "aliases"),
    ("afq_energy", (4, 50, 40, 1, 5, 0.0, 1, 1, 52, 1, 1, 0), "maximum of 32"),
    ("afq_orthonormalize", (4, 10, 3, 1, 1, 0), "aliased"),
    ("afq_phaseless_update", (4, 1, 1, 1, 1, 1, 1, 1, 0.0, 0.0, 0.0, -1.0, 0.1, 1e-3, 100.0, 0.1,
                              1, 1, 1, 0), "dt must be positive"),
    ("afq_comb_indices", (0, 1, 0.5, 1, 1, 1, 0), "nw=0"),
    ("afq_gather_rows", (4, 0, 1, 0, 0, 1, 1, 2, 0), "bad sizes"),
])
def test_invalid_arguments_rejected_on_host(name, args, msg):
    L = _lib()
    with pytest.raises(ValueError, match=msg):
        L.call(name, *args)
