"""Reverse-mode rule of slog_pfaffian: d log Pf / dA = 1/2 A^-T.

The reference obtains it by differentiating its JVP rule through
jnp.linalg.inv (pfcuda/cuda_api.py:93-107).  Here a HIP block-sweep kernel
recomputes the factorization with the forward pass's pivots and turns the
matrix into A^-1 in the same pass (pfh_slog_pfaffian_grad), so nothing from
the forward pass needs to be stored.
"""
from __future__ import annotations

import torch

from . import _lib
from ._lib import check, lib


def slog_grad(A: torch.Tensor, return_slog: bool = False):
    """1/2 A^-T for a skew matrix or batch (..., n, n) on the GPU."""
    from .api import _DTYPES, _flat, _stream
    n = A.shape[-1]
    Af, batch_shape, B = _flat(A)
    dt = _DTYPES[A.dtype]
    G = torch.empty_like(Af)
    logabs = torch.empty(B, dtype=torch.float64, device=A.device)
    phase = torch.empty(B, dtype=A.dtype, device=A.device)
    need = _lib._sz(0)
    check(lib.pfh_slog_pfaffian_grad_workspace_size(dt, B, n, need), "workspace query")
    ws = torch.empty(max(int(need.value), 1), dtype=torch.uint8, device=A.device)
    check(lib.pfh_slog_pfaffian_grad(dt, Af.data_ptr(), B, n, logabs.data_ptr(), phase.data_ptr(),
                                     G.data_ptr(), ws.data_ptr(), ws.numel(), _stream(A)),
          "slog_pfaffian_grad")
    G = G.reshape(*batch_shape, n, n)
    if return_slog:
        return G, logabs.reshape(batch_shape), phase.reshape(batch_shape)
    return G
