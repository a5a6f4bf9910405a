// Scalar type layer for the HIP Pfaffian kernels (gfx950).
//
// Plays the role of the reference's include/pfaffian_utils.cuh:1-72 (type
// traits, complex arithmetic, magnitudes) with this library's own complex
// type.  Complex division uses Smith's algorithm (R. L. Smith, "Algorithm
// 116: Complex division", CACM 5(8), 1962).  The zero-pivot test is exact
// (pivot == 0), as the reference's CPU and NumPy oracles do
// (src/pfaffian_cpu.cpp:79, pfcuda/pfaffian_py.py:34); the reference GPU
// kernels' absolute 10*eps threshold (pfaffian_utils.cuh:22-29) is not used.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pfh {

template <typename R>
struct cplx {
    R re, im;
};

using c64 = cplx<float>;
using c128 = cplx<double>;

template <typename T> struct real_of { using type = T; };
template <typename R> struct real_of<cplx<R>> { using type = R; };
template <typename T> struct is_cplx { static constexpr bool value = false; };
template <typename R> struct is_cplx<cplx<R>> { static constexpr bool value = true; };

// ---- real ops ----
__host__ __device__ inline float  t_zero(float)  { return 0.f; }
__host__ __device__ inline double t_zero(double) { return 0.0; }
__host__ __device__ inline float  t_one(float)   { return 1.f; }
__host__ __device__ inline double t_one(double)  { return 1.0; }
template <typename R> __host__ __device__ inline cplx<R> t_zero(cplx<R>) { return {R(0), R(0)}; }
template <typename R> __host__ __device__ inline cplx<R> t_one(cplx<R>)  { return {R(1), R(0)}; }

template <typename T> __host__ __device__ inline T zero() { return t_zero(T()); }
__host__ __device__ inline float  t_nan(float)  { return __builtin_nanf(""); }
__host__ __device__ inline double t_nan(double) { return __builtin_nan(""); }
template <typename R> __host__ __device__ inline cplx<R> t_nan(cplx<R>) { return {t_nan(R()), t_nan(R())}; }
template <typename T> __host__ __device__ inline T nan_of() { return t_nan(T()); }
template <typename T> __host__ __device__ inline T one()  { return t_one(T()); }

template <typename R> __host__ __device__ inline cplx<R> operator+(cplx<R> a, cplx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R> __host__ __device__ inline cplx<R> operator-(cplx<R> a, cplx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R> __host__ __device__ inline cplx<R> operator-(cplx<R> a) { return {-a.re, -a.im}; }
template <typename R> __host__ __device__ inline cplx<R> operator*(cplx<R> a, cplx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R> __host__ __device__ inline cplx<R> operator*(cplx<R> a, R s) { return {a.re * s, a.im * s}; }
template <typename R> __host__ __device__ inline bool operator==(cplx<R> a, cplx<R> b) { return a.re == b.re && a.im == b.im; }

// Squared magnitude used to order pivot candidates (monotone in |z|).
__host__ __device__ inline float  mag2(float x)  { return x < 0 ? -x : x; }
__host__ __device__ inline double mag2(double x) { return x < 0 ? -x : x; }
template <typename R> __host__ __device__ inline R mag2(cplx<R> z) { return z.re * z.re + z.im * z.im; }

__host__ __device__ inline bool is_zero(float x)  { return x == 0.f; }
__host__ __device__ inline bool is_zero(double x) { return x == 0.0; }
template <typename R> __host__ __device__ inline bool is_zero(cplx<R> z) { return z.re == R(0) && z.im == R(0); }

// Reciprocal: real 1/x; complex by Smith's algorithm.
__host__ __device__ inline float  recip(float x)  { return 1.f / x; }
__host__ __device__ inline double recip(double x) { return 1.0 / x; }
template <typename R> __host__ __device__ inline cplx<R> recip(cplx<R> z) {
    R a = z.re, b = z.im;
    if ((a < 0 ? -a : a) >= (b < 0 ? -b : b)) {
        R r = b / a, d = a + b * r;
        return {R(1) / d, -r / d};
    }
    R r = a / b, d = a * r + b;
    return {r / d, R(-1) / d};
}

// |x| in double, for the log accumulation.
__device__ inline double absd(float x)  { return fabs((double)x); }
__device__ inline double absd(double x) { return fabs(x); }
template <typename R> __device__ inline double absd(cplx<R> z) { return hypot((double)z.re, (double)z.im); }

// Running phase: real inputs carry a sign in double; complex a unit
// complex<double> renormalised after every product.
template <typename T> struct Phase {
    double s = 1.0;
    __device__ void flip() { s = -s; }
    __device__ void mul(T p) { if (p < T(0)) s = -s; }
    __device__ void kill() { s = 0.0; }
};
template <typename R> struct Phase<cplx<R>> {
    double re = 1.0, im = 0.0;
    __device__ void flip() { re = -re; im = -im; }
    __device__ void mul(cplx<R> p) {
        double m = hypot((double)p.re, (double)p.im);
        double pr = (double)p.re / m, pi = (double)p.im / m;
        double nr = re * pr - im * pi, ni = re * pi + im * pr;
        double nm = hypot(nr, ni);
        re = nr / nm; im = ni / nm;
    }
    __device__ void kill() { re = 0.0; im = 0.0; }
};

// log|prod| kept as mantissa * 2^exponent so that no log() is taken per
// pivot and nothing overflows: log|Pf| = log(m) + e*ln2 at the end.
struct LogAcc {
    double m = 1.0;
    int e = 0;
    __device__ void mul(double a) {
        m *= a;
        int ex;
        m = frexp(m, &ex);
        e += ex;
    }
    // out of line: inlined, the log's constants get hoisted into registers
    // that stay live across the whole elimination
    __device__ __attribute__((noinline)) double value() const {
        return log(m) + (double)e * 0.69314718055994530942;
    }
};

}  // namespace pfh
