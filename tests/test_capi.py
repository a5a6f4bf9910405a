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


def test_tuning_keys_match_header():
    from pfcuda_amd import _lib
    text = open(HEADER).read()
    keys = {m.group(1).lower(): int(m.group(2))
            for m in re.finditer(r"^#define PFH_TUNE_(\w+) (\d+)", text, re.M) if m.group(1) != "NKEYS"}
    assert keys == _lib.TUNE_KEYS
    assert int(re.search(r"^#define PFH_TUNE_NKEYS (\d+)", text, re.M).group(1)) == len(keys)


def test_set_get_tuning(lib):
    # host-only: values are checked against each key's list, -1 restores
    # the built-in choice; nothing here touches a GPU
    v = ctypes.c_int(7)
    try:
        assert lib.pfh_get_tuning(4, ctypes.byref(v)) == 0 and v.value == -1
        assert lib.pfh_set_tuning(4, 1) == 0
        assert lib.pfh_get_tuning(4, ctypes.byref(v)) == 0 and v.value == 1
        assert lib.pfh_set_tuning(4, 2) == -1            # not a value of this key
        assert lib.pfh_set_tuning(0, 1) == -1            # 1 is no kernel of an f64 bucket
        assert lib.pfh_set_tuning(0, 2) == 0
        assert lib.pfh_set_tuning(9, 0) == 0 and lib.pfh_set_tuning(9, 1 << 22) == 0   # panel spin limit
        assert lib.pfh_set_tuning(9, (1 << 22) + 1) == -1 and lib.pfh_set_tuning(9, -2) == -1
        assert lib.pfh_set_tuning(10, 2) == 0 and lib.pfh_set_tuning(10, 3) == -1   # panel exchange form
        assert lib.pfh_set_tuning(22, 32) == 0 and lib.pfh_set_tuning(22, 24) == -1  # gradient pairs per group
        assert lib.pfh_set_tuning(11, 9) == 0 and lib.pfh_set_tuning(11, 4) == -1   # mixed digit count
        nkeys = int(re.search(r"#define PFH_TUNE_NKEYS (\d+)", open(HEADER).read()).group(1))
        assert lib.pfh_set_tuning(nkeys, 0) == -1 and lib.pfh_set_tuning(-1, 0) == -1
        assert lib.pfh_get_tuning(nkeys, ctypes.byref(v)) == -1
        assert lib.pfh_get_tuning(0, None) == -1
    finally:
        for k in range(32):
            lib.pfh_set_tuning(k, -1)


def test_autotune_cache_applies_for_its_architecture_only(tmp_path):
    import json
    from pfcuda_amd import autotune
    path = tmp_path / "tune.json"
    path.write_text(json.dumps({"arch": "gfxTEST", "choices": {"trail_f64": 1, "f64_bucket_192": 2}}))
    try:
        assert autotune.load(str(path), arch="gfxOTHER") == {}
        assert autotune.get_tuning("trail_f64") == -1
        assert autotune.load(str(path), arch="gfxTEST") == {"trail_f64": 1, "f64_bucket_192": 2}
        assert autotune.get_tuning("trail_f64") == 1 and autotune.get_tuning("f64_bucket_192") == 2
        with pytest.raises(KeyError):
            autotune.apply({"no_such_key": 0})
    finally:
        for name in autotune.CANDIDATES:
            autotune.set_tuning(name, -1)


def test_panel_launch_key_takes_only_the_plain_form(lib):
    # the cooperative launch form was removed: key 7 accepts 0 (and -1) only
    lib.pfh_set_tuning.argtypes = [ctypes.c_int, ctypes.c_int]
    assert lib.pfh_set_tuning(7, 1) == -1
    assert lib.pfh_set_tuning(7, 0) == 0
    assert lib.pfh_set_tuning(7, -1) == 0


def test_host_forms_reject_bad_arguments(lib):
    vp = ctypes.c_void_p
    f = lib.pfh_pfaffian_host
    f.argtypes = [ctypes.c_int, vp, ctypes.c_int32, vp]
    buf = (ctypes.c_double * 4)()
    assert f(1, None, 2, buf) == -1                  # null input
    assert f(9, buf, 2, buf) == -1                   # unknown dtype
    assert f(1, buf, -2, buf) == -1                  # negative n
    assert f(1, buf, 66, buf) == -3                  # past the product-form limit
    g = lib.pfh_slog_pfaffian_host
    g.argtypes = [ctypes.c_int, vp, ctypes.c_int32, vp, vp]
    assert g(1, buf, 4, None, buf) == -1             # null log output
    lib.pfh_status_string.restype = ctypes.c_char_p
    assert b"gave up" in lib.pfh_status_string(-4)


def test_ragged_workspace_scales_with_count(lib):
    # one global slab per matrix past the register kernels, not one per
    # resident workgroup: a single n = 4096 f64 matrix needs about one
    # n^2 slab
    q = lib.pfh_slog_pfaffian_ragged_workspace_size
    q.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_size_t)]
    sz = ctypes.c_size_t(0)
    assert q(1, 1, 4096, ctypes.byref(sz)) == 0
    slab = 8 * 4096 * 4096
    assert slab <= sz.value < 2 * slab
