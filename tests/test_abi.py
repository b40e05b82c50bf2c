"""The C-ABI library loads and exports every entry point include/atus_hip.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "atus_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(atus_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "atus_run_block" in syms and "atus_sob_iter" in syms
    assert len(syms) >= 25


def test_library_exports_every_declared_symbol():
    from atus2_amd import hip
    path = hip.LIB_PATH
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert set(hip.EXPORTED) == set(declared_symbols())


def test_error_path_without_device_is_an_error_not_a_crash():
    from atus2_amd import hip
    lib = hip.load()
    assert lib.atus_version().decode().startswith("atus_hip")
    # invalid arguments are rejected before any device call
    h = ctypes.c_void_p()
    n = (ctypes.c_longlong * 3)(3, 1, 1)
    d = (ctypes.c_double * 3)(1.0, 0.0, 0.0)
    rc = lib.atus_ctx_create(ctypes.byref(h), 1, n, d, d, 1)
    assert rc == -3 and b"power of two" in lib.atus_last_error()
    rc = lib.atus_ctx_create(ctypes.byref(h), 4, n, d, d, 1)
    assert rc == -1


def test_batch_create_rejects_invalid_arguments_before_any_device_call():
    from atus2_amd import hip
    lib = hip.load()
    h = ctypes.c_void_p()
    n = (ctypes.c_longlong * 3)(64, 1, 1)
    d = (ctypes.c_double * 3)(1.0, 0.0, 0.0)
    rc = lib.atus_batch_create(ctypes.byref(h), 1, n, d, d, 3, 4)
    assert rc == -3 and b"two components" in lib.atus_last_error()
    rc = lib.atus_batch_create(ctypes.byref(h), 1, n, d, d, 2, 0)
    assert rc == -1 and b"nbatch" in lib.atus_last_error()
    rc = lib.atus_batch_create(ctypes.byref(h), 1, n, d, d, 2, 20000)
    assert rc == -1
    assert not h.value
