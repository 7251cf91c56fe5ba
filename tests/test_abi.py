"""The C-ABI library loads and exports every entry point include/espp_amd.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

from espressopp_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header="espp_amd.h"):
    src = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int)\s*(espp_\w+)\s*\(", src, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert len(syms) >= 30
    assert "espp_run" in syms and "espp_vl_pairs" in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    syms = declared_symbols() + declared_symbols("espp_amd_prof.h")
    missing = [s for s in syms if not hasattr(lib, s)]
    assert missing == []


def test_python_binding_covers_header():
    assert set(declared_symbols()) == set(_lib.SIGNATURES)
    assert set(declared_symbols("espp_amd_prof.h")) == set(_lib.PROF_SIGNATURES)


def test_version_and_error_strings():
    lib = _lib.load()
    assert lib.espp_version().startswith(b"espp_amd")
    assert isinstance(lib.espp_last_error(), bytes)


def test_create_fails_loudly_without_device_or_with_bad_arguments():
    lib = _lib.load()
    h = ctypes.c_void_p()
    box = (ctypes.c_double * 3)(10.0, 10.0, 10.0)
    for grid, hci in (((0, 3, 3), 1), ((1, 3, 3), 1), ((3, 3, 3), 2), ((4, 4, 4), 0)):
        cells = (ctypes.c_int32 * 3)(*grid)
        rc = lib.espp_create(box, cells, hci, 0.3, 1, 0, ctypes.byref(h))
        assert rc == -1 and not h.value
        msg = lib.espp_last_error()
        assert b"cell grid" in msg or b"halfCellInt" in msg
