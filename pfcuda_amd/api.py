"""Python surface of the library, mirroring the reference's pfcuda.cuda_api.

``pfaffian`` and ``slog_pfaffian`` keep the reference's names, guards,
output dtypes and derivative rules (pfcuda/cuda_api.py:27-107); torch
tensors on the GPU are the device-memory plumbing, and the arithmetic runs
in the HIP kernels of libpfaffian_hip.so.  There is no CPU fallback: inputs
that are not on a GPU are copied to the current GPU first, as JAX does for
NumPy inputs, and a missing library fails at import.

Derivatives follow the reference's JVP rules (cuda_api.py:78-107), exposed
through ``torch.autograd.Function`` with both ``jvp`` (forward mode) and
``backward`` (reverse mode):
  d Pf      = 1/2 Pf tr(A^-1 dA)       ->  A_bar = g * 1/2 Pf A^-T
  d log|Pf| = 1/2 tr(A^-1 dA)          ->  A_bar = g * 1/2 A^-T
For complex inputs the log-magnitude tangent is the real part of the trace
and the phase tangent is i*Im(...)*phase (the reference returns a complex
log tangent and a zero phase tangent, cuda_api.py:103-104); reverse mode
takes the phase's cotangent as well.  A real input's sign is piecewise
constant and carries no derivative.
"""
from __future__ import annotations

import math
from typing import Sequence, Tuple

import numpy as np
import torch

from . import _lib
from ._lib import check, lib

_DTYPES = {
    torch.float32: _lib.PFH_F32,
    torch.float64: _lib.PFH_F64,
    torch.complex64: _lib.PFH_C64,
    torch.complex128: _lib.PFH_C128,
}

PFAFFIAN_MAX_N = int(lib.pfh_pfaffian_max_n())


def _as_tensor(A) -> torch.Tensor:
    if isinstance(A, np.ndarray):
        return torch.from_numpy(A)
    if not isinstance(A, torch.Tensor):
        return torch.as_tensor(A)
    return A


def _to_device(A) -> torch.Tensor:
    A = _as_tensor(A)
    if not A.is_cuda:
        if not torch.cuda.is_available():
            raise RuntimeError("pfcuda_amd needs a ROCm GPU; none is visible")
        A = A.to("cuda")
    if A.dtype not in _DTYPES:
        raise TypeError(f"unsupported dtype {A.dtype}; use float32/64 or complex64/128")
    return A


def _stream(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def _square(A: torch.Tensor) -> int:
    if A.dim() < 2:
        raise ValueError(f"Pfaffian requires at least a 2D matrix (n, n); received rank {A.dim()}")
    n = A.shape[-1]
    if A.shape[-2] != n:
        raise ValueError(f"Pfaffian requires square matrices; received {tuple(A.shape[-2:])}")
    return n


def _flat(A: torch.Tensor) -> Tuple[torch.Tensor, Tuple[int, ...], int]:
    n = A.shape[-1]
    batch_shape = tuple(A.shape[:-2])
    B = int(np.prod(batch_shape)) if batch_shape else 1
    return A.reshape(B, n, n).contiguous(), batch_shape, B


def _real_dtype(dt):
    return torch.float32 if dt in (torch.float32, torch.complex64) else torch.float64


# ---------------------------------------------------------------------------
# raw kernel calls (no autograd)
# ---------------------------------------------------------------------------

def _pf_kernel(A: torch.Tensor) -> torch.Tensor:
    n = A.shape[-1]
    Af, batch_shape, B = _flat(A)
    out = torch.empty(B, dtype=A.dtype, device=A.device)
    fn = {torch.float32: lib.pfh_pfaffian_f32, torch.float64: lib.pfh_pfaffian_f64,
          torch.complex64: lib.pfh_pfaffian_c64, torch.complex128: lib.pfh_pfaffian_c128}[A.dtype]
    check(fn(Af.data_ptr(), B, n, out.data_ptr(), _stream(A)), "pfaffian")
    return out.reshape(batch_shape)


def _slog_kernel(A: torch.Tensor, method: int = _lib.PFH_METHOD_AUTO):
    n = A.shape[-1]
    Af, batch_shape, B = _flat(A)
    dt = _DTYPES[A.dtype]
    logabs = torch.empty(B, dtype=torch.float64, device=A.device)
    phase = torch.empty(B, dtype=A.dtype, device=A.device)
    need = _lib._sz(0)
    check(lib.pfh_slog_pfaffian_workspace_size(dt, B, n, method, need), "workspace query")
    ws = torch.empty(max(int(need.value), 1), dtype=torch.uint8, device=A.device)
    check(lib.pfh_slog_pfaffian_method(dt, Af.data_ptr(), B, n, method, logabs.data_ptr(),
                                       phase.data_ptr(), ws.data_ptr(), ws.numel(), _stream(A)),
          "slog_pfaffian")
    return logabs.reshape(batch_shape), phase.reshape(batch_shape)


def _inv(A: torch.Tensor) -> torch.Tensor:
    return torch.linalg.inv(A)


def _trace_prod(X: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
    """tr(X @ Y) over the last two dims, elementwise (no batched product)."""
    return (X * Y.mT).sum(dim=(-2, -1))


# ---------------------------------------------------------------------------
# operators: the entry points as torch.library primitives
# ---------------------------------------------------------------------------
# The reference registers each FFI handler with its framework
# (pfcuda/cuda_api.py:10-18) and calls it through jax.ffi.ffi_call
# (:47-52, :69-74).  Here each C-ABI entry point is an operator in the
# "pfcuda" namespace with
#   - a shape-only (fake) implementation, so that torch.compile traces
#     through it without a graph break and meta tensors propagate shapes;
#   - a batching rule that folds a vmapped dimension into the kernels' own
#     leading batch, so that torch.func.vmap costs one launch, not a loop.
# Their outputs never alias their inputs and the kernels never write them.

def _front(x: torch.Tensor, d):
    """x with its vmapped dimension d moved to the front (d is not None)."""
    return x.movedim(d, 0)


@torch.library.custom_op("pfcuda::pfaffian", mutates_args=(), schema="(Tensor A) -> Tensor")
def _pfaffian_op(A):
    return _pf_kernel(A)


@_pfaffian_op.register_fake
def _(A):
    return A.new_empty(A.shape[:-2])


def _pfaffian_vmap(info, in_dims, A):
    (d,) = in_dims
    if d is None:
        return _pfaffian_op(A), None
    return _pfaffian_op(_front(A, d)), 0


_pfaffian_op.register_vmap(_pfaffian_vmap)


@torch.library.custom_op("pfcuda::slog_pfaffian", mutates_args=(),
                         schema="(Tensor A, int method) -> (Tensor, Tensor)")
def _slog_op(A, method):
    return _slog_kernel(A, method)


@_slog_op.register_fake
def _(A, method):
    return (A.new_empty(A.shape[:-2], dtype=torch.float64), A.new_empty(A.shape[:-2]))


def _slog_vmap(info, in_dims, A, method):
    d = in_dims[0]
    if d is None:
        return _slog_op(A, method), (None, None)
    return _slog_op(_front(A, d), method), (0, 0)


_slog_op.register_vmap(_slog_vmap)


@torch.library.custom_op("pfcuda::slog_pfaffian_ragged", mutates_args=(),
                         schema="(Tensor data, Tensor offsets, Tensor sizes, int max_n) -> (Tensor, Tensor)")
def _ragged_op(data, offsets, sizes, max_n):
    return _ragged_kernel(data, offsets, sizes, max_n)


@_ragged_op.register_fake
def _(data, offsets, sizes, max_n):
    count = sizes.shape[0]
    return (data.new_empty((count,), dtype=torch.float64), data.new_empty((count,)))


# ---------------------------------------------------------------------------
# pfaffian
# ---------------------------------------------------------------------------

class _PfaffianFn(torch.autograd.Function):
    # under vmap the forward runs batched and reaches the operator's
    # batching rule; backward and jvp are torch operations
    generate_vmap_rule = True

    @staticmethod
    def forward(A):
        return _pfaffian_op(A)

    @staticmethod
    def setup_context(ctx, inputs, output):
        (A,) = inputs
        ctx.save_for_backward(A, output)
        ctx.save_for_forward(A, output)

    @staticmethod
    def backward(ctx, g):
        A, pf = ctx.saved_tensors
        G = 0.5 * pf[..., None, None] * _inv(A).mT
        if A.is_complex():
            G = G.conj()
        return g[..., None, None] * G

    @staticmethod
    def jvp(ctx, dA):
        A, pf = ctx.saved_tensors
        tr = _trace_prod(_inv(A), dA)
        return 0.5 * pf * tr


def pfaffian(A):
    """Pfaffian of a skew-symmetric matrix or batch (..., n, n).

    As pfcuda.pfaffian (cuda_api.py:27-54): odd or zero n gives 0, n = 2 and
    4 use closed forms, larger n runs the HIP kernel.  The kernel takes
    n <= 64 (the reference stops at 32); past that the product overflows
    and ``slog_pfaffian`` is the function to use.
    """
    A = _as_tensor(A)
    n = _square(A)
    A = _to_device(A)
    if n == 0 or (n & 1):
        return torch.zeros(A.shape[:-2], dtype=A.dtype, device=A.device)
    if n == 2:
        return A[..., 0, 1]
    if n == 4:
        return (A[..., 0, 1] * A[..., 2, 3] - A[..., 0, 2] * A[..., 1, 3]
                + A[..., 0, 3] * A[..., 1, 2])
    if n > PFAFFIAN_MAX_N:
        raise ValueError(
            f"Matrix size exceeds the maximum supported size of {PFAFFIAN_MAX_N}x{PFAFFIAN_MAX_N}.")
    return _PfaffianFn.apply(A)


# ---------------------------------------------------------------------------
# slog_pfaffian
# ---------------------------------------------------------------------------

class _SlogFn(torch.autograd.Function):
    generate_vmap_rule = True

    @staticmethod
    def forward(A, method):
        return _slog_op(A, method)

    @staticmethod
    def setup_context(ctx, inputs, output):
        A, _ = inputs
        logabs, phase = output
        ctx.save_for_backward(A, phase)
        ctx.save_for_forward(A, phase)
        if not A.is_complex():
            ctx.mark_non_differentiable(phase)    # a sign: piecewise constant

    @staticmethod
    def backward(ctx, g_log, g_phase):
        # With t = 1/2 tr(A^-1 dA) = sum(G * dA), G = 1/2 A^-T:
        #   d log|Pf| = Re t,  d phase = i Im(t) phase  (complex),
        # so A_bar = (g_log + i c) conj(G) with c = -Im(conj(g_phase) phase).
        A, phase = ctx.saved_tensors
        from .adjoint import slog_grad
        G = slog_grad(A)                      # 1/2 A^-T, from the HIP adjoint
        if not A.is_complex():
            return g_log.to(G.dtype)[..., None, None] * G, None
        coef = g_log.to(G.dtype)
        if g_phase is not None:
            coef = coef + 1j * (-(g_phase.conj() * phase).imag)
        return coef[..., None, None] * G.conj(), None

    @staticmethod
    def jvp(ctx, dA, _dmethod):
        A, phase = ctx.saved_tensors
        tr = 0.5 * _trace_prod(_inv(A), dA)
        if A.is_complex():
            return tr.real, 1j * tr.imag * phase
        return tr, None


def _slog_guarded(A, method: int, min_n: int):
    A = _as_tensor(A)
    n = _square(A)
    if n <= min_n:
        raise ValueError(
            f"Matrix size is less than the minimum supported size of ({min_n + 1}x{min_n + 1}). "
            "Use pfaffian() for smaller matrices.")
    A = _to_device(A)
    if n & 1:
        shape = A.shape[:-2]
        return (torch.full(shape, -math.inf, dtype=torch.float64, device=A.device),
                torch.zeros(shape, dtype=A.dtype, device=A.device))
    return _SlogFn.apply(A, method)


def slog_pfaffian(A, method: int = _lib.PFH_METHOD_AUTO):
    """(log|Pf|, sign or phase) of a skew-symmetric matrix or batch (..., n, n).

    As pfcuda.slog_pfaffian (cuda_api.py:56-76): n <= 32 raises ValueError,
    odd n gives (-inf, 0).  log|Pf| is float64 for every input dtype; the
    sign (real) or unit phase (complex) has the input dtype.
    """
    return _slog_guarded(A, method, 32)


def slog_pfaffian_batched(A, method: int = _lib.PFH_METHOD_AUTO):
    """slog_pfaffian without the n > 32 guard (any even n >= 2)."""
    return _slog_guarded(A, method, 0)


# ---------------------------------------------------------------------------
# ragged batches
# ---------------------------------------------------------------------------

def pack_ragged(mats: Sequence[torch.Tensor]):
    """Concatenate square matrices into (data, offsets, sizes) for the ragged call."""
    sizes = torch.tensor([int(m.shape[-1]) for m in mats], dtype=torch.int32)
    counts = (sizes.to(torch.int64) ** 2)
    offsets = torch.zeros(len(mats), dtype=torch.int64)
    if len(mats) > 1:
        offsets[1:] = torch.cumsum(counts, 0)[:-1]
    data = torch.cat([m.reshape(-1) for m in mats]) if mats else torch.empty(0)
    return data, offsets, sizes


def _ragged_kernel(data: torch.Tensor, offsets: torch.Tensor, sizes: torch.Tensor, max_n: int):
    """The ragged entry point on device tensors: contiguous data, int64
    offsets and int32 sizes, all on one GPU."""
    dev = data.device
    count = int(sizes.numel())
    dt = _DTYPES[data.dtype]
    logabs = torch.empty(count, dtype=torch.float64, device=dev)
    phase = torch.empty(count, dtype=data.dtype, device=dev)
    if count == 0:
        return logabs, phase
    need = _lib._sz(0)
    check(lib.pfh_slog_pfaffian_ragged_workspace_size(dt, count, int(max_n), need), "workspace query")
    ws = torch.empty(max(int(need.value), 1), dtype=torch.uint8, device=dev)
    check(lib.pfh_slog_pfaffian_ragged(dt, data.data_ptr(), offsets.data_ptr(), sizes.data_ptr(), count,
                                       int(max_n), logabs.data_ptr(), phase.data_ptr(), ws.data_ptr(),
                                       ws.numel(), _stream(data)),
          "slog_pfaffian_ragged")
    return logabs, phase


def _on_host(x) -> bool:
    return not (isinstance(x, torch.Tensor) and x.is_cuda)


def slog_pfaffian_ragged(data, offsets, sizes, max_n: int | None = None):
    """(log|Pf|, sign/phase) of every matrix of a packed ragged batch.

    ``data`` holds the matrices back to back, matrix i being the
    sizes[i] x sizes[i] row-major slab at element offsets[i].  One call, one
    persistent launch per size bucket, no host synchronisation once
    ``max_n`` is given (without it, the largest size is read back first).
    Offsets and sizes given in host memory are checked against ``data``
    here (ValueError); device tensors are taken as they are, so that the
    call stays free of host synchronisation.  A matrix larger than
    ``max_n`` gives NaN without disturbing the others.
    """
    data = _to_device(data).contiguous()
    dev = data.device
    if _on_host(offsets) and _on_host(sizes):
        o = np.asarray(torch.as_tensor(offsets).cpu(), dtype=np.int64).reshape(-1)
        z = np.asarray(torch.as_tensor(sizes).cpu(), dtype=np.int64).reshape(-1)
        if o.shape != z.shape:
            raise ValueError(f"offsets ({o.size}) and sizes ({z.size}) differ in length")
        if z.size:
            if (z < 0).any() or (o < 0).any():
                raise ValueError("offsets and sizes must be non-negative")
            if (z > 2**31 - 1).any():
                raise ValueError("a size does not fit int32")
            if int((o + z * z).max()) > data.numel():
                bad = int(np.argmax(o + z * z > data.numel()))
                raise ValueError(f"matrix {bad} (offset {o[bad]}, size {z[bad]}) runs past the end of "
                                 f"data ({data.numel()} elements)")
    offsets = torch.as_tensor(offsets, dtype=torch.int64).to(dev).contiguous()
    sizes_t = torch.as_tensor(sizes, dtype=torch.int32)
    if max_n is None:
        max_n = int(sizes_t.max().item()) if sizes_t.numel() else 0
    sizes_t = sizes_t.to(dev).contiguous()
    return _ragged_op(data, offsets, sizes_t, int(max_n))


# ---------------------------------------------------------------------------
# host-pointer forms and the stream status
# ---------------------------------------------------------------------------

_NP_DTYPES = {np.dtype(np.float32): _lib.PFH_F32, np.dtype(np.float64): _lib.PFH_F64,
              np.dtype(np.complex64): _lib.PFH_C64, np.dtype(np.complex128): _lib.PFH_C128}


def _host_matrix(A) -> Tuple[np.ndarray, int]:
    a = np.ascontiguousarray(A)
    if a.dtype not in _NP_DTYPES:
        raise TypeError(f"unsupported dtype {a.dtype}; use float32/64 or complex64/128")
    if a.ndim != 2 or a.shape[0] != a.shape[1]:
        raise ValueError(f"expected one square matrix; received shape {a.shape}")
    return a, _NP_DTYPES[a.dtype]


def pfaffian_host(A):
    """Pf of one n x n matrix in host memory (NumPy), through the
    synchronous host-pointer entry point (as the reference's
    pfaffian<T>(h_A, n), src/pfaffian.cu:4-40; n <= 64)."""
    a, dt = _host_matrix(A)
    out = np.zeros(1, dtype=a.dtype)
    check(lib.pfh_pfaffian_host(dt, a.ctypes.data, a.shape[0], out.ctypes.data), "pfaffian_host")
    return out[0]


def slog_pfaffian_host(A):
    """(log|Pf|, sign or phase) of one matrix in host memory, through the
    synchronous host-pointer entry point (as the reference's
    slog_pfaffian<T>(h_A, n, ...), src/slog_pfaffian.cu:4-32)."""
    a, dt = _host_matrix(A)
    la = np.zeros(1, dtype=np.float64)
    ph = np.zeros(1, dtype=a.dtype)
    check(lib.pfh_slog_pfaffian_host(dt, a.ctypes.data, a.shape[0], la.ctypes.data, ph.ctypes.data),
          "slog_pfaffian_host")
    return float(la[0]), ph[0]


def check_status(device=None) -> None:
    """Wait for the current stream and raise if a blocked panel on this
    device gave up waiting for its blocks since the last check (the results
    of that call are NaN)."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    check(lib.pfh_stream_status(torch.cuda.current_stream(dev).cuda_stream), "stream status")
