// One wave per matrix for n <= 32: the product form and slog of the
// reference's pf_sm_calc (src/pfaffian_sm.cu:3-99), with the rows kept in
// logical order so that every pair step updates only the live part.
//
// Layout: lane l holds physical column c = l % 32 and the rows of parity
// h = l / 32; register i holds logical row 2i + h.  Columns are pivoted
// virtually (lg: the logical index of the lane's column), rows physically:
// the interchange k+1 <-> p moves logical row p into the register of row
// k+1 and row k+1's old values into row p's register (a wave-uniform
// register index; across the two halves through a lane exchange).  So at
// pair step k = 2i the pivot row k is register i of the even half and
// row k+1 register i of the odd half, and the rank-2 update touches
// registers i+1 .. 15 only (the live rows), about half the register work of
// updating the whole square every step.
//
// Per pair step (right-looking, the skew Schur complement
// S = A22 + (y x^T - x y^T)/a of pfaffian_sm.cu:71-89):
//   1. argmax |X[c]| over the even half's lanes with lg > k (the largest
//      magnitude, ties to the lowest logical index; NaN never wins), as
//      src/pfaffian_cpu.cpp:47-52;
//   2. the interchange k+1 <-> p (rows physically, column labels);
//   3. the two pivot rows go to LDS by logical column: X = row k, Y = row k+1;
//   4. R[r][c] += (Y[r] X[c] - X[r] Y[c]) / a over the live rows.
// The input is read from its strict upper triangle only (the lower entries
// are its negatives).  Pf = (-1)^swaps prod(pivots); the exact zero-pivot
// test gives Pf = 0 (log -inf, phase 0).
#include <type_traits>

#include "pf_batch.h"
#include "pf_types.h"
#include "pf_wave.h"

namespace pfh {

namespace {

typedef double c32_v16d __attribute__((ext_vector_type(16)));
typedef float c32_v16f __attribute__((ext_vector_type(16)));

// 16 register rows of one type; get / set take compile-time or wave-uniform
// indices (the latter compile to s_set_gpr_idx moves, no scratch)
template <typename T> struct C32Regs;
template <> struct C32Regs<double> {
    c32_v16d v;
    __device__ __forceinline__ double get(int i) const { return v[i]; }
    __device__ __forceinline__ void set(int i, double x) { v[i] = x; }
};
template <> struct C32Regs<float> {
    c32_v16f v;
    __device__ __forceinline__ float get(int i) const { return v[i]; }
    __device__ __forceinline__ void set(int i, float x) { v[i] = x; }
};
template <typename R> struct C32Regs<cplx<R>> {
    C32Regs<R> re, im;
    __device__ __forceinline__ cplx<R> get(int i) const { return {re.get(i), im.get(i)}; }
    __device__ __forceinline__ void set(int i, cplx<R> x) { re.set(i, x.re); im.set(i, x.im); }
};

// The interchange's exchange between the halves, in one v_permlane32_swap
// per 32-bit part (a VALU operation; the lane shuffle it replaces went
// through the LDS crossbar): with a = row k+1 and b = row p, the upper half
// receives b from the lower half (a_x, row p seen from the odd half) and the
// lower half receives a from the upper half (b_x, row k+1 seen from the even
// half).  Each result is used only in the half that receives it.
__device__ __forceinline__ void c32_swap(unsigned a, unsigned b, unsigned& a_x, unsigned& b_x) {
    const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
    a_x = r[0];   // upper lanes: b's lower half
    b_x = r[1];   // lower lanes: a's upper half
}
__device__ __forceinline__ void c32_swap(float a, float b, float& a_x, float& b_x) {
    unsigned x, y;
    c32_swap(__builtin_bit_cast(unsigned, a), __builtin_bit_cast(unsigned, b), x, y);
    a_x = __builtin_bit_cast(float, x);
    b_x = __builtin_bit_cast(float, y);
}
__device__ __forceinline__ void c32_swap(double a, double b, double& a_x, double& b_x) {
    const unsigned long long ua = __builtin_bit_cast(unsigned long long, a), ub = __builtin_bit_cast(unsigned long long, b);
    unsigned lx, ly, hx, hy;
    c32_swap((unsigned)ua, (unsigned)ub, lx, ly);
    c32_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), hx, hy);
    a_x = __builtin_bit_cast(double, ((unsigned long long)hx << 32) | lx);
    b_x = __builtin_bit_cast(double, ((unsigned long long)hy << 32) | ly);
}
template <typename R>
__device__ __forceinline__ void c32_swap(cplx<R> a, cplx<R> b, cplx<R>& a_x, cplx<R>& b_x) {
    c32_swap(a.re, b.re, a_x.re, b_x.re);
    c32_swap(a.im, b.im, a_x.im, b_x.im);
}

__device__ __forceinline__ double c32_lane(double x, int l) { return readlane_f64(x, l); }
__device__ __forceinline__ float c32_lane(float x, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
}
template <typename R> __device__ __forceinline__ cplx<R> c32_lane(cplx<R> z, int l) { return {c32_lane(z.re, l), c32_lane(z.im, l)}; }

// c ? a : b, complex values part by part
template <typename T>
__device__ __forceinline__ T c32_sel(bool c, T a, T b) {
    if constexpr (is_cplx<T>::value) return T{c ? a.re : b.re, c ? a.im : b.im};
    else return c ? a : b;
}

// f(std::integral_constant<int, I>) for I = I0 .. N-1 while it returns true
template <int I, int N, typename F>
__device__ __forceinline__ void c32_steps(F&& f) {
    if constexpr (I < N) {
        if (f(std::integral_constant<int, I>{})) c32_steps<I + 1, N>(f);
    }
}

}  // namespace

// Waves per SIMD the register allocation aims at, the most without spills
// (float64 product form: 125 VGPRs at 4 waves; 2^20 32 x 32 matrices took
// 5.46 ms there and 6.35 ms at the unconstrained 157 VGPRs / 3 waves)
template <typename T, int MODE>
constexpr int c32_waves() {
    if constexpr (std::is_same<T, c128>::value) return 2;
    else if constexpr (std::is_same<T, float>::value) return 4;
    else return MODE == MODE_PROD ? 4 : 3;
}

template <typename T, int MODE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(c32_waves<T, MODE>(), 8)))
pf_c32_kernel(BatchDesc d) {
    using R = typename real_of<T>::type;
    struct XY { T x, y; };
    __shared__ XY xy[32];              // pivot rows k (x) and k+1 (y), by logical column
    const int lane = threadIdx.x, c = lane & 31, h = lane >> 5;

    int64_t count = d.count;
    if (d.count_dev) {
        const int64_t cd = *d.count_dev;
        count = cd < count ? cd : count;
    }
    for (int64_t j = blockIdx.x; j < count; j += gridDim.x) {
        const int idx = d.order ? d.order[(d.begin_dev ? *d.begin_dev : 0) + j] : (int)j;
        const int n = __builtin_amdgcn_readfirstlane(d.sizes ? d.sizes[idx] : d.dense_n);
        const int64_t off = d.offsets ? d.offsets[idx] : (int64_t)idx * n * n;
        const T* base = (const T*)d.data + off;
        if (n < 0 || n > 32) {   // outside this kernel's capacity: NaN, no access
            if (lane == 0) {
                if (MODE == MODE_PROD) ((T*)d.prod_out)[idx] = nan_of<T>();
                else { d.log_out[idx] = __builtin_nan(""); ((T*)d.phase_out)[idx] = nan_of<T>(); }
            }
            continue;
        }
        if (n == 0 || (n & 1)) {
            if (lane == 0) {
                if (MODE == MODE_PROD) ((T*)d.prod_out)[idx] = zero<T>();
                else { d.log_out[idx] = -INFINITY; ((T*)d.phase_out)[idx] = zero<T>(); }
            }
            continue;
        }

        // rows 2i + h at column c, from the strict upper triangle
        C32Regs<T> A;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int r = 2 * i + h;
            const bool in = r < n && c < n && r != c;
            const int a = r < c ? r : c, b = r < c ? c : r;
            const T v = base[in ? (unsigned)(a * n + b) : 0u];
            A.set(i, !in ? zero<T>() : (r < c ? v : -v));
        }
        int lg = c;                    // logical index of this lane's column
        T prod = one<T>();
        LogAcc la;
        Phase<T> ph;
        bool dead = false;

        // pair steps i = 0 .. 15, each compiled with its own register indices
        c32_steps<0, 16>([&](auto ic) -> bool {
            constexpr int i = decltype(ic)::value;
            const int k = 2 * i;
            if (k >= n) return false;
            // 1. pivot search over row k (the even half's lanes)
            const T xk = A.get(i);
            double best = -1.0;
            if (h == 0 && lg > k && lg < n) {
                const double m = (double)mag2(xk);
                if (m == m) best = m;
            }
            double wb;
            int wk;
            wave_argmax(best, lg * 64 + lane, wb, wk);
            int p, cs;
            if (wk != 0x7fffffff) {
                p = wk >> 6;
                cs = wk & 63;
            } else {                   // only NaN candidates: keep k+1
                p = k + 1;
                const unsigned long long at = __ballot(h == 0 && lg == k + 1);
                cs = (int)__builtin_ctzll(at);
            }
            const T piv = c32_lane(xk, cs);
            if (is_zero(piv)) { dead = true; return false; }
            const bool sw = p != k + 1;
            // 2. the interchange k+1 <-> p, without branches around the
            //    register vector (a merge of two versions of it would copy
            //    all sixteen registers): one read and one write of row p's
            //    register at a wave-uniform index (without an interchange,
            //    another live register is rewritten with its own values),
            //    per-lane selects for the halves
            {
                const int pi = sw ? (p >> 1) : (i < 15 ? i + 1 : i);   // register of row p (> i)
                const T rowp = A.get(pi);                  // row p lives in half p & 1
                const T rowk1 = A.get(i);                  // row k+1 lives in the odd half
                T rowp_x, rowk1_x;                         // row p seen from the odd half, row k+1
                c32_swap(rowk1, rowp, rowp_x, rowk1_x);    // seen from the even half
                const bool podd = p & 1;
                // row p's register takes row k+1's values (odd half when p is
                // odd, else the even half through the exchange)
                A.set(pi, c32_sel(sw && (podd ? h == 1 : h == 0), podd ? rowk1 : rowk1_x, rowp));
                // row k+1 (odd half of register i) takes row p's values
                A.set(i, c32_sel(sw && h == 1, podd ? rowp : rowp_x, rowk1));
                if (sw) {
                    if (lg == k + 1) lg = p;
                    else if (c == cs) lg = k + 1;
                }
            }
            if constexpr (MODE == MODE_PROD) { prod = prod * piv; if (sw) prod = -prod; }
            else { la.mul(absd(piv)); ph.mul(piv); if (sw) ph.flip(); }
            if (k + 2 >= n) return false;
            // 3. rows k (even half) and k+1 (odd half) by logical column
            const T own = A.get(i);
            if (h == 0) xy[lg].x = own;
            else xy[lg].y = own;
            __syncthreads();
            // 4. rank-2 update of the live rows
            const T inv = recip(piv);
            const XY mine = xy[lg];
            const T xs = mine.x * inv, ys = mine.y * inv;
            // four rows' LDS reads at a time (all at once would hold up to 60
            // registers and cost occupancy)
#pragma unroll
            for (int r2 = i + 1; r2 < 16; ++r2) {
                const XY q = xy[2 * r2 + h];
                A.set(r2, A.get(r2) + q.y * xs - q.x * ys);
                if ((r2 - i) % 4 == 0) __builtin_amdgcn_sched_barrier(0);
            }
            __syncthreads();           // xy is rewritten by the next step
            return true;
        });

        if (lane == 0) {
            if constexpr (MODE == MODE_PROD) {
                ((T*)d.prod_out)[idx] = dead ? zero<T>() : prod;
            } else {
                d.log_out[idx] = dead ? -INFINITY : la.value();
                T phv;
                if constexpr (is_cplx<T>::value) { phv.re = (R)ph.re; phv.im = (R)ph.im; }
                else phv = (T)ph.s;
                ((T*)d.phase_out)[idx] = dead ? zero<T>() : phv;
            }
        }
    }
}

template <typename T, int MODE>
static int c32_launch(const BatchDesc& d, hipStream_t s) {
    if (d.count <= 0) return 0;
    auto kern = pf_c32_kernel<T, MODE>;
    static DevCache resident;   // resident workgroups on the device
    int64_t grid = resident.get([&] {
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, 64, 0) != hipSuccess || occ <= 0) occ = 1;
        return occ * device_cu_count();
    });
    if (grid > d.count) grid = d.count;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64), 0, s, d);
    return (int)hipGetLastError();
}

int launch_c32(int dtype, int mode, const BatchDesc& d, hipStream_t s) {
    if (mode == MODE_PROD) {
        switch (dtype) {
            case DT_F32: return c32_launch<float, MODE_PROD>(d, s);
            case DT_F64: return c32_launch<double, MODE_PROD>(d, s);
            case DT_C64: return c32_launch<c64, MODE_PROD>(d, s);
            case DT_C128: return c32_launch<c128, MODE_PROD>(d, s);
        }
    } else if (mode == MODE_SLOG) {
        switch (dtype) {
            case DT_F32: return c32_launch<float, MODE_SLOG>(d, s);
            case DT_F64: return c32_launch<double, MODE_SLOG>(d, s);
            case DT_C64: return c32_launch<c64, MODE_SLOG>(d, s);
            case DT_C128: return c32_launch<c128, MODE_SLOG>(d, s);
        }
    }
    return (int)hipErrorInvalidValue;
}

}  // namespace pfh
