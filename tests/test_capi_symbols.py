"""The C-ABI library loads without a GPU and exports every function that
include/dts_capi.h declares (no compute calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "dts_capi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b((?:dts|hfbs)_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_declared_symbols():
    from domain_transform_solver_amd import _lib
    lib = _lib.load()
    names = _declared()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_signatures_cover_declarations():
    from domain_transform_solver_amd import _lib
    missing = [n for n in _declared() if n not in _lib.exported_symbols()]
    assert not missing, missing


def test_last_error_without_gpu():
    from domain_transform_solver_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.dts_create(0, 10, 0, ctypes.byref(h)) == 2  # DTS_ERR_INVALID_DIMS
    assert b"max dims" in lib.dts_last_error()
    assert lib.dts_create(10, 10, 0, None) == 1  # DTS_ERR_INVALID_ARG
