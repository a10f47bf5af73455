"""The C ABI library loads and exports every entry point include/platepcg.h
declares; host-side argument checks work without a GPU.  CPU only."""
import ctypes as C
import os
import re

import pytest

from platepcg import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "platepcg.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int32_t|int64_t|double|const char\*)\s+(pp_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_null_arguments_are_rejected():
    lib = _lib.load()
    out = C.c_void_p()
    assert lib.pp_ctx_create(0, None) == _lib.PP_ERR_INVALID
    assert b"NULL" in lib.pp_last_error()
    assert lib.pp_plan_create(None, None, None, None, None, C.byref(out)) == _lib.PP_ERR_INVALID
    assert lib.pp_set_rhs(None, None) == _lib.PP_ERR_INVALID
    assert lib.pp_ctx_destroy(None) == _lib.PP_OK
    assert lib.pp_plan_destroy(None) == _lib.PP_OK


def test_no_device_reports_cleanly():
    if _lib.load().pp_device_count() > 0:
        pytest.skip("a HIP device is present")
    out = C.c_void_p()
    assert _lib.load().pp_ctx_create(0, C.byref(out)) == _lib.PP_ERR_NODEVICE


def test_python_constants_match_the_header():
    """Every #define PP_<NAME> <int> of include/platepcg.h has the same value
    in the Python mirror (as PP_<NAME>, or <NAME> without the prefix)."""
    text = open(os.path.join(ROOT, "include", "platepcg.h")).read()
    defs = re.findall(r"^#define\s+PP_(\w+)\s+\(?(-?\d+)u?\)?", text, re.M)
    assert len(defs) > 20
    for name, value in defs:
        mirror = getattr(_lib, "PP_" + name, getattr(_lib, name, None))
        assert mirror is not None, name
        assert mirror == int(value), (name, mirror, value)
