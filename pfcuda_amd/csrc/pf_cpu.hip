// Host (CPU) Pfaffian, the library's counterpart of the reference's public
// pfaffian_cpu (pfcuda/cpp_api.py:3-4 -> bindings/pybind_bindings.cpp:8-30
// -> src/pfaffian_cpu.cpp:7-106): FP64, one thread, the product form (so it
// overflows past ~1e308 as the reference's does), working in place on the
// caller's row-major n x n buffer.  No GPU code: it is a separate entry
// point for host data, never a fallback of the GPU paths.
//
// Right-looking Parlett-Reid on the full square: per pair step k, the pivot
// is the largest |A[k][c]|, c > k (lowest c on ties, NaN never chosen; the
// reference compares against a stale value, pfaffian_cpu.cpp:29-30), the
// symmetric interchange k+1 <-> p flips the sign, an exact zero pivot gives
// Pf = 0, and the trailing block takes the skew Schur update
//   A[r][c] += (A[k+1][r] A[k][c] - A[k][r] A[k+1][c]) / a,  r, c >= k+2.
#include <cmath>
#include <cstdint>
#include <utility>

#include "../../include/pfaffian_hip.h"

namespace {

double pf_cpu(double* A, int n) {
    double prod = 1.0;
    for (int k = 0; k + 1 < n; k += 2) {
        const double* rk = A + (int64_t)k * n;
        int p = k + 1;
        double best = -1.0;
        for (int c = k + 1; c < n; ++c) {
            const double m = std::fabs(rk[c]);
            if (m > best) { best = m; p = c; }
        }
        if (p != k + 1) {
            double* r1 = A + (int64_t)(k + 1) * n;
            double* rp = A + (int64_t)p * n;
            for (int c = 0; c < n; ++c) std::swap(r1[c], rp[c]);
            for (int r = 0; r < n; ++r) std::swap(A[(int64_t)r * n + k + 1], A[(int64_t)r * n + p]);
            prod = -prod;
        }
        const double a = rk[k + 1];
        if (a == 0.0) return 0.0;
        prod *= a;
        const double inv = 1.0 / a;
        const double* x = rk;                          // row k
        const double* y = A + (int64_t)(k + 1) * n;    // row k+1
        for (int r = k + 2; r < n; ++r) {
            double* ar = A + (int64_t)r * n;
            const double yr = y[r] * inv, xr = x[r] * inv;
            for (int c = k + 2; c < n; ++c) ar[c] += yr * x[c] - xr * y[c];
        }
    }
    return prod;
}

}  // namespace

// As the reference (pfaffian_cpu.cpp:8-20): n = 0 gives 0, n = 2 and 4
// the closed forms; odd n is rejected (its binding raises, :17-18).
extern "C" int pfh_pfaffian_cpu_f64(double* A, int32_t n, double* out) {
    if (!out || n < 0 || (n > 0 && !A)) return PFH_EINVAL;
    if (n & 1) return PFH_EINVAL;
    if (n == 0) *out = 0.0;
    else if (n == 2) *out = A[1];
    else if (n == 4) *out = A[1] * A[11] - A[2] * A[7] + A[3] * A[6];
    else *out = pf_cpu(A, n);
    return 0;
}
