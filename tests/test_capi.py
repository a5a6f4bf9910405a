"""The C-ABI library loads and exports every symbol include/pfaffian_hip.h
declares; argument checks that return before any GPU work (no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pfaffian_hip.h")
LIB = os.path.join(ROOT, "pfcuda_amd", "libpfaffian_hip.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*|void)\s+(pfh_\w+)\s*\(", text, re.M)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libpfaffian_hip.so is not built (run __graft_entry__.build())")
    return ctypes.CDLL(LIB)


def test_every_declared_symbol_is_exported(lib):
    names = declared()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n


def test_status_strings(lib):
    lib.pfh_status_string.restype = ctypes.c_char_p
    assert lib.pfh_status_string(0) == b"success"
    assert b"invalid" in lib.pfh_status_string(-1)
    assert b"workspace" in lib.pfh_status_string(-2)


def test_argument_rejections_before_gpu_work(lib):
    vp = ctypes.c_void_p
    f = lib.pfh_pfaffian_f64
    f.argtypes = [vp, ctypes.c_int64, ctypes.c_int32, vp, vp]
    assert f(None, -1, 8, None, None) == -1          # negative batch
    assert f(None, 4, -2, None, None) == -1          # negative n
    assert f(None, 4, 66, None, None) == -3          # past the product-form limit
    assert f(None, 0, 8, None, None) == 0            # empty batch is a no-op
    assert lib.pfh_pfaffian_max_n() == 64
    g = lib.pfh_slog_pfaffian_method
    g.argtypes = [ctypes.c_int, vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int, vp, vp, vp, ctypes.c_size_t, vp]
    assert g(1, None, 2, 40, 9, None, None, None, 0, None) == -1   # unknown method
    assert g(7, None, 2, 40, 0, None, None, None, 0, None) == -1   # unknown dtype
    r = lib.pfh_slog_pfaffian_ragged
    r.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_int64, ctypes.c_int32, vp, vp, vp, ctypes.c_size_t, vp]
    assert r(1, None, None, None, 0, 0, None, None, None, 0, None) == 0    # empty
    assert r(1, None, None, None, 5, 64, None, None, None, 0, None) == -1  # null pointers


def test_workspace_queries(lib):
    sz = ctypes.c_size_t(123)
    q = lib.pfh_slog_pfaffian_workspace_size
    q.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int32, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]
    assert q(1, 10, 64, 0, ctypes.byref(sz)) == 0 and sz.value == 0      # register path: none
    assert q(1, 1, 4096, 1, ctypes.byref(sz)) == 0 and sz.value >= 4096 * 4096 * 8
    assert q(3, 1, 1024, 1, ctypes.byref(sz)) == 0 and sz.value >= 1024 * 1024 * 16
    rq = lib.pfh_slog_pfaffian_ragged_workspace_size
    rq.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_size_t)]
    assert rq(1, 16384, 256, ctypes.byref(sz)) == 0 and sz.value >= 16384 * 4


def test_package_imports_and_mirrors_reference_names():
    import pfcuda_amd
    for name in ("pfaffian", "slog_pfaffian", "slog_pfaffian_ragged"):
        assert callable(getattr(pfcuda_amd, name))
    assert pfcuda_amd.PFAFFIAN_MAX_N == 64


def test_host_guards_without_gpu():
    import numpy as np
    import pfcuda_amd
    # shape errors are raised before any device transfer
    with pytest.raises(ValueError):
        pfcuda_amd.pfaffian(np.zeros(3))
    with pytest.raises(ValueError):
        pfcuda_amd.pfaffian(np.zeros((3, 4)))
