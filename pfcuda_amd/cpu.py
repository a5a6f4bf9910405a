"""The reference's CPU surface, rebuilt: ``pfaffian_cpu`` and ``pfaffian_py``
(exported by the reference's pfcuda/__init__.py:1-5).  Both take host data,
work on the CPU and are never called by the GPU paths.

``pfaffian_cpu(A)``: the C++ entry point pfh_pfaffian_cpu_f64 (FP64, one
thread), with the reference binding's argument handling
(bindings/pybind_bindings.cpp:8-30): the input is taken as a C-contiguous
float64 array (converted, i.e. copied, only when it is not one already),
must be 2-D, square and of even size (RuntimeError otherwise), and is
worked on in place.  Returns a float.

``pfaffian_py(A)``: NumPy, any floating or complex dtype, as the reference's
pfcuda/pfaffian_py.py:3-50: n = 0 or odd gives 0.0, n = 2 and 4 the closed
forms, and A is worked on in place.  Its elimination is right-looking
(a rank-2 update of the trailing block per pair step) with the largest
|A[k][c]| as pivot, lowest c on ties.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import lib


def pfaffian_cpu(A) -> float:
    a = np.asarray(A)
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 2:
        raise RuntimeError("Input must be 2D")
    n = a.shape[0]
    if n != a.shape[1]:
        raise RuntimeError("Matrix must be square")
    if n % 2 != 0:
        raise RuntimeError("n must be even")
    out = ctypes.c_double(0.0)
    status = lib.pfh_pfaffian_cpu_f64(a.ctypes.data if n else None, n, ctypes.byref(out))
    if status != 0:
        raise RuntimeError(f"pfh_pfaffian_cpu_f64 failed: status {status}")
    return out.value


def pfaffian_py(A):
    n = A.shape[0]
    if n == 0 or n & 1:
        return 0.0
    if n == 2:
        return A[0, 1]
    if n == 4:
        return A[0, 1] * A[2, 3] - A[0, 2] * A[1, 3] + A[0, 3] * A[1, 2]
    result = 1.0
    for k in range(0, n - 1, 2):
        row = A[k, k + 1:]
        p = k + 1 + int(np.argmax(np.abs(row)))       # first maximum: lowest index
        if p != k + 1:
            A[[k + 1, p], :] = A[[p, k + 1], :]
            A[:, [k + 1, p]] = A[:, [p, k + 1]]
            result = -result
        a = A[k, k + 1]
        if a == 0:
            return 0.0
        result = result * a
        if k + 2 < n:
            x = A[k, k + 2:]
            y = A[k + 1, k + 2:]
            A[k + 2:, k + 2:] += (np.outer(y, x) - np.outer(x, y)) / a
    return result
