"""The C ABI: libdust_gpu.so loads without a GPU and exports every function
declared in include/dust_gpu.h; the Python bindings cover the same set.  No
compute calls (those need a GPU)."""
import ctypes
import os
import re

from dust_amd import _lib as L

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "include", "dust_gpu.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dust_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("dust_fmm_create", "dust_fmm_sort_particles", "dust_fmm_calculate_multipole",
                 "dust_fmm_apply_multipole", "dust_aic_surfpan", "dust_dgetrf", "dust_dgetrs",
                 "dust_dgesv_mixed", "dust_solve_linsys"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_match_header():
    sig = {s[0] for s in L.SIGNATURES}
    assert set(declared()) <= sig | {"dust_last_error", "dust_version"}, set(declared()) - sig


def test_version_and_error_text():
    lib = L.lib()
    assert lib.dust_version().decode()
    # an argument error is reported without touching a GPU
    rc = lib.dust_dgetrf(-1, 1, None, 1, None)
    assert rc == L.DUST_E_ARG
    assert lib.dust_last_error().decode()
