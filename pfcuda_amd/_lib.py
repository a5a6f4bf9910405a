"""ctypes binding of libpfaffian_hip.so (the C ABI in include/pfaffian_hip.h).

The library is loaded from this package directory.  There is no fallback:
if it is missing or fails to load, importing ``pfcuda_amd`` raises.
"""
from __future__ import annotations

import ctypes
import os

# torch first: it brings its own HIP runtime, and the library must bind to
# that one instance (loaded before torch, the library pulls in a second
# runtime, which then sees no device)
import torch  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpfaffian_hip.so")

PFH_F32, PFH_F64, PFH_C64, PFH_C128 = 0, 1, 2, 3
PFH_METHOD_AUTO, PFH_METHOD_UNBLOCKED, PFH_METHOD_BLOCKED, PFH_METHOD_MIXED = 0, 1, 2, 3
# tuning keys (include/pfaffian_hip.h, PFH_TUNE_*)
TUNE_KEYS = {
    "f64_bucket_128": 0, "f64_bucket_192": 1, "f64_bucket_256": 2, "c128_bucket_128": 3,
    "trail_f64": 4, "trail_c128": 5, "trail_sweep": 6, "panel_launch": 7, "grad_blocked": 8,
    "panel_spin": 9, "panel_xcd": 10, "mixed_slices": 11,
    "bat_nb_128": 12, "bat_nb_192": 13, "bat_nb_256": 14,
    "bat_nb_c128_128": 15, "bat_nb_c128_192": 16, "bat_nb_c128_256": 17,
    "blocked_nb": 18, "blocked_nb_c128": 19, "switch_n": 20, "small_kernel": 21, "grad_nb": 22,
}

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_sz = ctypes.c_size_t


class PfaffianLibError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python pfcuda_amd/build.py` (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    lib.pfh_status_string.argtypes = [ctypes.c_int]
    lib.pfh_status_string.restype = ctypes.c_char_p
    lib.pfh_pfaffian_max_n.argtypes = []
    for name in ("pfh_pfaffian_f32", "pfh_pfaffian_f64", "pfh_pfaffian_c64", "pfh_pfaffian_c128"):
        getattr(lib, name).argtypes = [_vp, _i64, _i32, _vp, _vp]
    lib.pfh_slog_pfaffian_workspace_size.argtypes = [ctypes.c_int, _i64, _i32, ctypes.c_int, ctypes.POINTER(_sz)]
    for name in ("pfh_slog_pfaffian_f32", "pfh_slog_pfaffian_f64", "pfh_slog_pfaffian_c64",
                 "pfh_slog_pfaffian_c128"):
        getattr(lib, name).argtypes = [_vp, _i64, _i32, _vp, _vp, _vp, _sz, _vp]
    lib.pfh_slog_pfaffian_method.argtypes = [ctypes.c_int, _vp, _i64, _i32, ctypes.c_int, _vp, _vp, _vp, _sz, _vp]
    lib.pfh_slog_pfaffian_ragged_workspace_size.argtypes = [ctypes.c_int, _i64, _i32, ctypes.POINTER(_sz)]
    lib.pfh_slog_pfaffian_ragged.argtypes = [ctypes.c_int, _vp, _vp, _vp, _i64, _i32, _vp, _vp, _vp, _sz, _vp]
    lib.pfh_slog_pfaffian_grad_workspace_size.argtypes = [ctypes.c_int, _i64, _i32, ctypes.POINTER(_sz)]
    lib.pfh_slog_pfaffian_grad.argtypes = [ctypes.c_int, _vp, _i64, _i32, _vp, _vp, _vp, _vp, _sz, _vp]
    lib.pfh_pfaffian_cpu_f64.argtypes = [_vp, _i32, ctypes.POINTER(ctypes.c_double)]
    lib.pfh_pfaffian_host.argtypes = [ctypes.c_int, _vp, _i32, _vp]
    lib.pfh_slog_pfaffian_host.argtypes = [ctypes.c_int, _vp, _i32, _vp, _vp]
    lib.pfh_stream_status.argtypes = [_vp]
    lib.pfh_slog_pfaffian_ragged_noalloc.argtypes = [ctypes.c_int, _vp, _vp, _vp, _i64, _i32, _vp, _vp, _vp]
    lib.pfh_set_tuning.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.pfh_get_tuning.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    return lib


lib = _load()


def check(status: int, what: str) -> None:
    if status != 0:
        msg = lib.pfh_status_string(status).decode()
        raise PfaffianLibError(f"{what} failed: status {status} ({msg})")


def exported_symbols():
    """Names of the pfh_* entry points, as bound above."""
    return [n for n in dir(lib) if n.startswith("pfh_")]
