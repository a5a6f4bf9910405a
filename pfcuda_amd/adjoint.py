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


def _grad_kernel(A: torch.Tensor):
    """(1/2 A^-T, log|Pf|, sign/phase) from one call of the adjoint entry point."""
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
    return G.reshape(*batch_shape, n, n), logabs.reshape(batch_shape), phase.reshape(batch_shape)


# The adjoint as an operator, like the forward entry points: a shape-only
# implementation for torch.compile and a batching rule for torch.func.vmap.
@torch.library.custom_op("pfcuda::slog_pfaffian_grad", mutates_args=(), schema="(Tensor A) -> Tensor")
def _grad_op(A):
    return _grad_kernel(A)[0]


@_grad_op.register_fake
def _(A):
    return A.new_empty(A.shape)


def _grad_vmap(info, in_dims, A):
    (d,) = in_dims
    if d is None:
        return _grad_op(A), None
    return _grad_op(A.movedim(d, 0)), 0


_grad_op.register_vmap(_grad_vmap)


class _GradFn(torch.autograd.Function):
    """The adjoint's own derivative, so that second derivatives of log|Pf|
    exist as they do through the reference's jnp.linalg.inv
    (cuda_api.py:93-107).  With G = 1/2 A^-T:
      dG = -2 G dA^T G                      (jvp)
      A_bar = -2 conj(G) G_bar^T conj(G)    (its adjoint under Re tr(X^H Y))
    Both are torch operations on the saved G, so they differentiate again."""
    generate_vmap_rule = True

    @staticmethod
    def forward(A):
        return _grad_op(A)

    @staticmethod
    def setup_context(ctx, inputs, output):
        ctx.save_for_backward(output)
        ctx.save_for_forward(output)

    @staticmethod
    def backward(ctx, g_bar):
        (G,) = ctx.saved_tensors
        Gc = G.conj() if G.is_complex() else G
        return -2.0 * torch.matmul(torch.matmul(Gc, g_bar.mT.contiguous()), Gc)

    @staticmethod
    def jvp(ctx, dA):
        (G,) = ctx.saved_tensors
        return -2.0 * torch.matmul(torch.matmul(G, dA.mT.contiguous()), G)


def slog_grad(A: torch.Tensor, return_slog: bool = False):
    """1/2 A^-T for a skew matrix or batch (..., n, n) on the GPU.

    With return_slog the same call's (log|Pf|, sign/phase) come back too
    (the kernel computes them on the way); that form is not an operator.
    """
    if return_slog:
        return _grad_kernel(A)
    return _GradFn.apply(A)
