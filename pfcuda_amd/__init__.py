"""MI355X-native Pfaffian library: HIP kernels for gfx950 behind a C ABI.

Public surface (mirrors the reference package pfcuda, pfcuda/__init__.py:1-5):
  pfaffian, slog_pfaffian             -- as pfcuda.pfaffian / pfcuda.slog_pfaffian
  pfaffian_cpu, pfaffian_py           -- as pfcuda.pfaffian_cpu / pfcuda.pfaffian_py (host data, CPU)
  slog_pfaffian_batched               -- slog without the n > 32 guard
  slog_pfaffian_ragged, pack_ragged   -- uneven batches in one call
  pfaffian_host, slog_pfaffian_host   -- one matrix in host memory, synchronous (as the reference's
                                         host-pointer templates, src/pfaffian.cu:4-40, src/slog_pfaffian.cu:4-32)
  check_status                        -- raise if a blocked panel gave up waiting for its blocks
  autotune                            -- device tuning of the run-time choices (python -m pfcuda_amd.autotune)
  torch.ops.pfcuda.{pfaffian, slog_pfaffian, slog_pfaffian_grad, slog_pfaffian_ragged}
                                      -- the entry points as torch.library operators (compile, vmap, grad)
Importing this package loads libpfaffian_hip.so and fails if it is absent.
"""
import os as _os

from ._lib import PFH_METHOD_AUTO, PFH_METHOD_BLOCKED, PFH_METHOD_MIXED, PFH_METHOD_UNBLOCKED
from .api import (PFAFFIAN_MAX_N, check_status, pack_ragged, pfaffian, pfaffian_host, slog_pfaffian,
                  slog_pfaffian_batched, slog_pfaffian_host, slog_pfaffian_ragged)
from .adjoint import slog_grad  # noqa: F401  (also registers pfcuda::slog_pfaffian_grad)
from .cpu import pfaffian_cpu, pfaffian_py

# A tuning cache written by pfcuda_amd.autotune, applied when PFH_TUNE_FILE
# names one (and a GPU is present).
if _os.environ.get("PFH_TUNE_FILE"):
    import torch as _torch
    if _torch.cuda.is_available():
        from .autotune import load as _load_tuning
        _load_tuning(_os.environ["PFH_TUNE_FILE"])

__all__ = [
    "pfaffian", "slog_pfaffian", "pfaffian_cpu", "pfaffian_py",
    "slog_pfaffian_batched", "slog_pfaffian_ragged", "pack_ragged",
    "pfaffian_host", "slog_pfaffian_host", "check_status",
    "PFAFFIAN_MAX_N", "PFH_METHOD_AUTO", "PFH_METHOD_UNBLOCKED", "PFH_METHOD_BLOCKED",
    "PFH_METHOD_MIXED",
]
