"""CPU oracle for the Pfaffian hot path — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package, and only as the checker or
the timed CPU baseline.  The product package ``pfcuda_amd`` never imports it.

Contents, each citing the reference function it restates:

* ``pf_f64`` / ``pf_c128`` — ctypes calls into ``liboracle.so``, the C++
  restatement of ``pfaffian_cpu`` (src/pfaffian_cpu.cpp:7-106) in
  log/sign form, real and complex.  ``stale=True`` reproduces the
  reference's pivot rule bit for bit (pinned against ``oracle/_ref``).
* ``pf_f64_ld`` / ``pf_c128_ld`` — the same elimination in x87 long double,
  the extended-precision "truth" for the mixed-precision accuracy checks.
* ``pf_f64_mt`` / ``pf_c128_mt`` — the same elimination with each pair
  step's column loops split over host threads, bitwise equal to the above
  (for the N = 4096 / 8192 single-matrix checks).
* ``slog_ragged_f64`` / ``slog_ragged_c128`` — threaded batch driver (one
  matrix per task) used as the CPU baseline.
* ``pfaffian_py`` — NumPy restatement of ``pfaffian_py``
  (pfcuda/pfaffian_py.py:3-50), dtype-generic, returning log/sign too.
* ``pfaffian_brute`` — expansion over perfect matchings (exact on integer
  or Fraction input, n <= 12).
* ``pfaffian_mp`` — mpmath Parlett-Reid at a chosen precision.
* ``ref_pfaffian`` — the reference's own ``pfaffian`` compiled from its
  source by ``oracle/Makefile`` into ``oracle/_ref`` (only when built).

Parity status of each function (what pins it) is stated in the tests
``tests/test_oracle.py``: every function here is checked against closed
forms, brute force, mpmath or the compiled reference.
"""
from __future__ import annotations

import ctypes
import math
import os
from fractions import Fraction

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)


def build() -> None:
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        lib.oracle_pf_f64.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        lib.oracle_pf_f64_ld.argtypes = [_dp, ctypes.c_int, _dp]
        lib.oracle_pf_c128.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        lib.oracle_pf_c128_ld.argtypes = [_dp, ctypes.c_int, _dp]
        lib.oracle_pf_f64_mt.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp]
        lib.oracle_pf_c128_mt.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp]
        for name in ("oracle_slog_ragged_f64", "oracle_slog_ragged_c128"):
            getattr(lib, name).argtypes = [_dp, _i64p, _i32p, ctypes.c_int64, ctypes.c_int, _dp, _dp]
            getattr(lib, name).restype = None
        _LIB = lib
    return _LIB


def ref_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libpfref.so"))


def ref_pfaffian(A: np.ndarray) -> float:
    """The reference's compiled ``pfaffian(double*, int)`` on a copy of A."""
    global _REF
    if _REF is None:
        _REF = ctypes.CDLL(os.path.join(_HERE, "_ref", "libpfref.so"))
        _REF.pfref_pfaffian.argtypes = [_dp, ctypes.c_int]
        _REF.pfref_pfaffian.restype = ctypes.c_double
    a = np.ascontiguousarray(A, dtype=np.float64).copy()
    return float(_REF.pfref_pfaffian(a.ctypes.data_as(_dp), a.shape[0]))


def _as_dp(a):
    return a.ctypes.data_as(_dp)


def pf_f64(A: np.ndarray, stale: bool = False, closed_forms: bool = False):
    """(prod, logabs, sign, info) of a real matrix, FP64, left-looking."""
    a = np.ascontiguousarray(A, dtype=np.float64)
    out = np.zeros(3)
    info = _lib().oracle_pf_f64(_as_dp(a), a.shape[0], int(stale), int(closed_forms), _as_dp(out))
    return float(out[0]), float(out[1]), float(out[2]), int(info)


def pf_f64_ld(A: np.ndarray):
    """(prod, logabs, sign, info) with the elimination in long double."""
    a = np.ascontiguousarray(A, dtype=np.float64)
    out = np.zeros(3)
    info = _lib().oracle_pf_f64_ld(_as_dp(a), a.shape[0], _as_dp(out))
    return float(out[0]), float(out[1]), float(out[2]), int(info)


def pf_c128(A: np.ndarray, stale: bool = False, closed_forms: bool = False):
    """(prod, logabs, phase, info) of a complex matrix, complex128."""
    a = np.ascontiguousarray(A, dtype=np.complex128)
    out = np.zeros(5)
    info = _lib().oracle_pf_c128(_as_dp(a.view(np.float64)), a.shape[0], int(stale),
                                 int(closed_forms), _as_dp(out))
    return complex(out[0], out[1]), float(out[2]), complex(out[3], out[4]), int(info)


def pf_c128_ld(A: np.ndarray):
    a = np.ascontiguousarray(A, dtype=np.complex128)
    out = np.zeros(5)
    info = _lib().oracle_pf_c128_ld(_as_dp(a.view(np.float64)), a.shape[0], _as_dp(out))
    return complex(out[0], out[1]), float(out[2]), complex(out[3], out[4]), int(info)


def default_threads() -> int:
    """Host threads for the threaded oracle: the cores this process may use, at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def pf_f64_mt(A: np.ndarray, threads: int | None = None):
    """pf_f64 with each pair step's column loops split over host threads.

    Bitwise equal to ``pf_f64(A)`` (each column keeps its operation order;
    tests/test_oracle.py checks it), for the N = 4096 / 8192 checks."""
    a = np.ascontiguousarray(A, dtype=np.float64)
    out = np.zeros(3)
    info = _lib().oracle_pf_f64_mt(_as_dp(a), a.shape[0], threads or default_threads(), _as_dp(out))
    return float(out[0]), float(out[1]), float(out[2]), int(info)


def pf_c128_mt(A: np.ndarray, threads: int | None = None):
    """pf_c128 with threaded column loops, bitwise equal to ``pf_c128(A)``."""
    a = np.ascontiguousarray(A, dtype=np.complex128)
    out = np.zeros(5)
    info = _lib().oracle_pf_c128_mt(_as_dp(a.view(np.float64)), a.shape[0], threads or default_threads(),
                                    _as_dp(out))
    return complex(out[0], out[1]), float(out[2]), complex(out[3], out[4]), int(info)


def slog(A: np.ndarray, threads: int | None = None):
    """(logabs, sign-or-phase) for real or complex input via the C++ oracle.

    Matrices of 512 and more run on the threaded form (bitwise the same)."""
    big = A.shape[-1] >= 512
    if np.iscomplexobj(A):
        _, la, ph, _ = pf_c128_mt(A, threads) if big else pf_c128(A)
        return la, ph
    _, la, s, _ = pf_f64_mt(A, threads) if big else pf_f64(A)
    return la, s


def slog_ragged(data: np.ndarray, offsets: np.ndarray, sizes: np.ndarray, threads: int = 1):
    """Batch oracle over a packed ragged batch (offsets in elements)."""
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    sizes = np.ascontiguousarray(sizes, dtype=np.int32)
    count = len(sizes)
    logabs = np.zeros(count)
    if np.iscomplexobj(data):
        d = np.ascontiguousarray(data, dtype=np.complex128)
        ph = np.zeros((count, 2))
        _lib().oracle_slog_ragged_c128(_as_dp(d.view(np.float64)), offsets.ctypes.data_as(_i64p),
                                       sizes.ctypes.data_as(_i32p), count, threads,
                                       _as_dp(logabs), _as_dp(ph))
        return logabs, ph[:, 0] + 1j * ph[:, 1]
    d = np.ascontiguousarray(data, dtype=np.float64)
    sign = np.zeros(count)
    _lib().oracle_slog_ragged_f64(_as_dp(d), offsets.ctypes.data_as(_i64p),
                                  sizes.ctypes.data_as(_i32p), count, threads,
                                  _as_dp(logabs), _as_dp(sign))
    return logabs, sign


# --------------------------------------------------------------------------
# NumPy restatement of pfcuda/pfaffian_py.py:3-50 (the reference's second check)
# --------------------------------------------------------------------------

def pfaffian_py(A: np.ndarray, return_log: bool = False):
    """Left-looking Parlett-Reid in NumPy over the pair-row views.

    Follows pfcuda/pfaffian_py.py: V = rows 0::2 and W = rows 1::2 of the
    working copy (:19-20); row k is updated by the pair-panel product dR
    (:22-26, :30-31); true argmax pivot (:32); exact zero test (:34-35); full
    symmetric swap (:37-43); the new row k+1 is stored negated (:46) and row
    k scaled by its pivot (:48).  Works on a copy.  Closed forms for n = 2, 4
    as :8-16.  With ``return_log`` it returns (logabs, sign/phase) instead.
    """
    A = np.array(A, copy=True)
    n = A.shape[0]
    if n == 0 or (n & 1):
        return (-math.inf, A.dtype.type(0)) if return_log else A.dtype.type(0)
    if not return_log:
        if n == 2:
            return A[0, 1]
        if n == 4:
            return A[0, 1] * A[2, 3] - A[0, 2] * A[1, 3] + A[0, 3] * A[1, 2]
    result = A.dtype.type(1)
    logabs = 0.0
    phase = A.dtype.type(1)
    V = A[0::2, :]
    W = A[1::2, :]

    for i, k in enumerate(range(0, n - 1, 2)):
        def dR(kk):
            return V[:i, kk] @ W[:i, kk + 1:] - W[:i, kk] @ V[:i, kk + 1:]
        Ak = A[k, k + 1:]
        Ak += dR(k)
        km = int(np.abs(Ak).argmax())
        if Ak[km] == 0:
            return (-math.inf, A.dtype.type(0)) if return_log else A.dtype.type(0)
        if km != 0:
            kp = k + 1 + km
            A[[k + 1, kp], :] = A[[kp, k + 1], :]
            A[:, [k + 1, kp]] = A[:, [kp, k + 1]]
            result *= -1
            phase = -phase
        A[k + 1, k + 2:] = -(A[k + 1, k + 2:] + dR(k + 1))
        piv = Ak[0]
        Ak[1:] /= piv
        result *= piv
        logabs += math.log(abs(piv))
        phase = phase * (piv / abs(piv))
    if return_log:
        return logabs, phase
    return result


# --------------------------------------------------------------------------
# Mathematical pins
# --------------------------------------------------------------------------

def pfaffian_brute(A) -> object:
    """Pf by expansion along the first row over perfect matchings.

    Pf(A) = sum_{j>0} (-1)^(j+1) a_{0j} Pf(A with rows/cols 0, j removed).
    Exact when A holds ints or Fractions; n <= 12 (10395 matchings).
    """
    n = len(A)
    if n == 0:
        return 1
    if n & 1:
        return 0
    idx = list(range(n))

    def rec(ids):
        if not ids:
            return 1
        i0 = ids[0]
        total = 0
        for pos in range(1, len(ids)):
            j = ids[pos]
            a = A[i0][j]
            if a == 0:
                continue
            rest = ids[1:pos] + ids[pos + 1:]
            sgn = 1 if (pos % 2 == 1) else -1
            total = total + sgn * a * rec(rest)
        return total

    return rec(idx)


def pfaffian_mp(A, dps: int = 50):
    """mpmath Parlett-Reid (right-looking, partial pivoting) at ``dps`` digits.

    Returns (logabs as float, sign or phase as complex/float, Pf as mpc/mpf).
    """
    import mpmath as mp
    with mp.workdps(dps):
        n = len(A)
        if n & 1:
            return -math.inf, 0.0, mp.mpf(0)
        is_c = np.iscomplexobj(A)
        M = [[(mp.mpc(complex(A[i][j])) if is_c else mp.mpf(float(A[i][j]))) for j in range(n)]
             for i in range(n)]
        pf = mp.mpf(1)
        for k in range(0, n - 1, 2):
            kp = max(range(k + 1, n), key=lambda c: abs(M[k][c]))
            if kp != k + 1:
                M[k + 1], M[kp] = M[kp], M[k + 1]
                for row in M:
                    row[k + 1], row[kp] = row[kp], row[k + 1]
                pf = -pf
            a = M[k][k + 1]
            if a == 0:
                return -math.inf, 0.0, mp.mpf(0)
            pf = pf * a
            for r in range(k + 2, n):
                yr = M[k + 1][r]
                xr = M[k][r]
                for c in range(k + 2, n):
                    M[r][c] += (yr * M[k][c] - xr * M[k + 1][c]) / a
        la = float(mp.log(abs(pf)))
        ph = pf / abs(pf)
        ph = complex(ph) if is_c else float(ph)
        return la, ph, pf


# --------------------------------------------------------------------------
# Seeded generators for the benchmark configurations (shared by tests/bench)
# --------------------------------------------------------------------------

def skew_gaussian(rng: np.random.Generator, n: int, complex_: bool = False, batch=None):
    """A = (B - B^T)/2 with B ~ N(0,1) (complex: re, im ~ N(0,1)/sqrt2)."""
    shape = (n, n) if batch is None else (batch, n, n)
    if complex_:
        B = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / math.sqrt(2.0)
    else:
        B = rng.standard_normal(shape)
    return (B - np.swapaxes(B, -1, -2)) / 2


def orthogonal(rng: np.random.Generator, n: int) -> np.ndarray:
    """Random orthogonal Q from QR with the sign fix of test/pfaffian_test.py:18-26."""
    A = rng.standard_normal((n, n))
    Q, R = np.linalg.qr(A)
    return Q * np.sign(np.diag(R))


def tridiag_identity_case(rng: np.random.Generator, n: int):
    """The reference's identity test (test/pfaffian_test.py:80-110).

    Returns (M, expected_pf) with M = B T B^T, T skew tridiagonal with
    super-diagonal t (odd positions zero), B orthogonal:
    Pf(M) = det(B) * prod(t[0::2]).
    """
    alt = 1.0 + rng.standard_normal(n - 1)
    alt = alt * (1 - (np.arange(n - 1) % 2))
    T = np.diag(alt, k=1) - np.diag(alt, k=-1)
    B = orthogonal(rng, n)
    return B @ T @ B.T, float(np.linalg.det(B) * np.prod(alt[0::2]))
