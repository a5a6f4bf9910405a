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

// ---------------------------------------------------------------------------
// Two real matrices per wave.  Half h = lane / 32 holds matrix 2q + h of pair
// q: lane c = lane % 32 its column c and register j its row j, both in the
// input's order for the whole elimination.  The interchanges are virtual and
// symmetric, so one map serves rows and columns: lg, the logical index of
// the lane's column (and of row c).  Per pair step k = 2i, with one
// wave-uniform step count for both halves (the larger matrix's):
//   1. X = row k at the lane's column: register phys(k), read at a
//      wave-uniform index once per half, and a select;
//   2. argmax |X| over the half's lanes with k < lg < n (the largest
//      magnitude, ties to the lowest logical index, NaN never wins), as
//      src/pfaffian_cpu.cpp:47-52, in DPP steps that stay inside 32 lanes;
//   3. the interchange k+1 <-> p in the map; Y = row k+1 likewise;
//   4. X and Y to LDS by column, R[j][c] += (Y[j] X[c] - X[j] Y[c]) / a over
//      all 32 rows (rows already eliminated take values never read again).
// The work of a step besides the update (pivot search, reciprocal,
// bookkeeping) serves two matrices and no register moves, for the price of
// updating the eliminated rows too.  Every live entry takes the arithmetic
// of pf_c32_kernel, in the same order, so both give the same bits.

// 32 register rows as two objects of 16, rows 0..15 and 16..31 (a vector of
// 32 doubles is past what a register index reaches, and one struct holding
// both vectors went to scratch memory).  Row j at a wave-uniform j: a read
// from each object and a select of the two values (the values pass through
// an empty asm statement, or the compiler selects between the whole vectors
// first, 32 moves, and reads once).
template <typename T>
__device__ __forceinline__ T c32_row(const C32Regs<T>& lo, const C32Regs<T>& hi, int j) {
    T a = lo.get(j & 15), b = hi.get(j & 15);
    asm volatile("" : "+v"(a), "+v"(b));
    return j < 16 ? a : b;
}

// maximum / minimum over each 32-lane half: lanes 16..31 end with half 0's,
// lanes 48..63 with half 1's (the last step, row_bcast:15, writes rows 1
// and 3 only)
__device__ __forceinline__ unsigned c32_half_max(unsigned v) {
    v = max(v, dpp_u32<0xB1, 0xF>(v));
    v = max(v, dpp_u32<0x4E, 0xF>(v));
    v = max(v, dpp_u32<0x141, 0xF>(v));
    v = max(v, dpp_u32<0x140, 0xF>(v));
    return max(v, dpp_u32<0x142, 0xA>(v));
}
__device__ __forceinline__ int c32_half_min(int v) {
    v = min(v, dpp_min_i32<0xB1, 0xF>(v));
    v = min(v, dpp_min_i32<0x4E, 0xF>(v));
    v = min(v, dpp_min_i32<0x141, 0xF>(v));
    v = min(v, dpp_min_i32<0x140, 0xF>(v));
    return min(v, dpp_min_i32<0x142, 0xA>(v));
}

// The winning lane of each half (0..31, 32..63) for (m, lg): the largest m
// (-1 marks a lane that is no candidate; the others are >= 0, no NaN), ties
// to the lowest lg; -1 for a half without a candidate.  Values >= 0 order
// as their bit patterns (plus one, so that 0 marks no candidate): the high
// words decide unless a half holds its largest high word twice.
__device__ __forceinline__ void c32_argmax2(double m, int lg, int& w0, int& w1) {
    const int lane = (int)threadIdx.x;
    const unsigned long long kb =
        m >= 0.0 ? (__builtin_bit_cast(unsigned long long, m) & 0x7fffffffffffffffull) + 1ull : 0ull;
    const unsigned hi = (unsigned)(kb >> 32), lo = (unsigned)kb;
    const unsigned hr = c32_half_max(hi);
    const unsigned H0 = (unsigned)__builtin_amdgcn_readlane((int)hr, 31);
    const unsigned H1 = (unsigned)__builtin_amdgcn_readlane((int)hr, 63);
    const unsigned Hh = lane < 32 ? H0 : H1;
    const unsigned long long hit = __ballot(hi == Hh);
    const unsigned b0 = (unsigned)hit, b1 = (unsigned)(hit >> 32);
    if ((b0 & (b0 - 1)) == 0 && (b1 & (b1 - 1)) == 0) {
        // one lane per half holds its largest high word, which is then > 0
        // (a half without candidates has 32 lanes at 0)
        w0 = (int)__builtin_ctz(b0);
        w1 = 32 + (int)__builtin_ctz(b1);
        return;
    }
    const unsigned lr = c32_half_max(hi == Hh ? lo : 0u);
    const unsigned L0 = (unsigned)__builtin_amdgcn_readlane((int)lr, 31);
    const unsigned L1 = (unsigned)__builtin_amdgcn_readlane((int)lr, 63);
    const bool tie = hi == Hh && lo == (lane < 32 ? L0 : L1);
    const int kr = c32_half_min(tie ? lg : 0x7fffffff);
    const int K0 = __builtin_amdgcn_readlane(kr, 31), K1 = __builtin_amdgcn_readlane(kr, 63);
    const unsigned long long win = __ballot(tie && lg == (lane < 32 ? K0 : K1));
    w0 = (H0 | L0) ? (int)__builtin_ctz((unsigned)win) : -1;
    w1 = (H1 | L1) ? 32 + (int)__builtin_ctz((unsigned)(win >> 32)) : -1;
}

// five waves per SIMD (96 VGPRs; a spill or two outside the pair steps): 2^20
// 32 x 32 float64 4.62 -> 4.48 ms against four
template <typename T, int MODE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5, 8)))
pf_c32x2_kernel(BatchDesc d) {
    struct XY { T x, y; };
    __shared__ XY xy[64];              // rows k (x) and k+1 (y) by column, per half
    const int lane = threadIdx.x, c = lane & 31, h = lane >> 5;

    int64_t count = d.count;
    if (d.count_dev) {
        const int64_t cd = *d.count_dev;
        count = cd < count ? cd : count;
    }
    C32Regs<T> Alo, Ahi;               // rows 0..15 and 16..31 at the lane's column
    for (int64_t q = blockIdx.x; 2 * q < count; q += gridDim.x) {
        const int64_t j = 2 * q + h;
        const bool has = j < count;
        int idx = 0, n = 0;
        if (has) {
            idx = d.order ? d.order[(d.begin_dev ? *d.begin_dev : 0) + j] : (int)j;
            n = d.sizes ? d.sizes[idx] : d.dense_n;
        }
        const bool bad = has && (n < 0 || n > 32);            // outside this kernel's capacity: NaN
        const bool nil = has && !bad && (n == 0 || (n & 1));  // Pf = 0
        const int nh = (has && !bad && !nil) ? n : 0;         // this half's size (0: no steps)
        // (a half with no steps reads nothing)
        const T* base = (const T*)d.data + (nh > 0 ? (d.offsets ? d.offsets[idx] : (int64_t)idx * n * n) : 0);
        const int nmax = max(__builtin_amdgcn_readlane(nh, 0), __builtin_amdgcn_readlane(nh, 32));

        // rows j at column c, from the strict upper triangle
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            const bool in = r < nh && c < nh && r != c;
            const int a = r < c ? r : c, b = r < c ? c : r;
            const T v = base[in ? (unsigned)(a * nh + b) : 0u];
            const T e = !in ? T(0) : (r < c ? v : -v);
            if (r < 16) Alo.set(r, e);
            else Ahi.set(r - 16, e);
        }
        int lg = c;                        // logical index of this lane's column
        T prod = T(1);
        LogAcc la;
        Phase<T> ph;
        bool dead = false;

        for (int k = 0; k < nmax; k += 2) {
            const bool act = k < nh && !dead;  // this half takes a pivot at this step
            // 1. row k at the own column
            const unsigned long long bk = __ballot(lg == k);
            const int pk0 = (int)__builtin_ctz((unsigned)bk), pk1 = (int)__builtin_ctz((unsigned)(bk >> 32));
            const T xa = c32_row(Alo, Ahi, pk0), xb = c32_row(Alo, Ahi, pk1);
            const T X = h ? xb : xa;
            // 2. pivot search over the half's live columns
            double m = -1.0;
            if (lg > k && lg < nh) {
                const double mm = (double)mag2(X);
                if (mm == mm) m = mm;
            }
            int w0, w1;
            c32_argmax2(m, lg, w0, w1);
            if (w0 < 0 || w1 < 0) {            // no candidate (a NaN row, a half past its size): keep k+1
                const unsigned long long b1 = __ballot(lg == k + 1);
                if (w0 < 0) w0 = (int)__builtin_ctz((unsigned)b1);
                if (w1 < 0) w1 = 32 + (int)__builtin_ctz((unsigned)(b1 >> 32));
            }
            const int p0 = __builtin_amdgcn_readlane(lg, w0), p1 = __builtin_amdgcn_readlane(lg, w1);
            const T pv0 = c32_lane(X, w0), pv1 = c32_lane(X, w1);
            const int p = h ? p1 : p0, w = h ? w1 : w0;
            const T piv = h ? pv1 : pv0;
            const bool sw = p != k + 1;
            // 3. the interchange k+1 <-> p of this half's map
            if (sw) {
                if (lg == k + 1) lg = p;
                else if (lane == w) lg = k + 1;
            }
            if (act) {
                if (is_zero(piv)) dead = true;
                else if constexpr (MODE == MODE_PROD) { prod = prod * piv; if (sw) prod = -prod; }
                else { la.mul(absd(piv)); ph.mul(piv); if (sw) ph.flip(); }
            }
            if (k + 2 >= nmax) break;
            // row k+1 (physical w) at the own column
            const T ya = c32_row(Alo, Ahi, w0 & 31), yb = c32_row(Alo, Ahi, w1 & 31);
            const T Y = h ? yb : ya;
            xy[lane] = XY{X, Y};
            __syncthreads();
            // 4. rank-2 update of every row
            const T inv = recip(piv);
            const T xs = X * inv, ys = Y * inv;
#pragma unroll
            for (int r = 0; r < 32; ++r) {
                const XY qv = xy[32 * h + r];
                if (r < 16) Alo.set(r, Alo.get(r) + qv.y * xs - qv.x * ys);
                else Ahi.set(r - 16, Ahi.get(r - 16) + qv.y * xs - qv.x * ys);
            }
            __syncthreads();           // xy is rewritten by the next step
        }

        if (c == 0 && has) {
            if (bad) {
                if (MODE == MODE_PROD) ((T*)d.prod_out)[idx] = nan_of<T>();
                else { d.log_out[idx] = __builtin_nan(""); ((T*)d.phase_out)[idx] = nan_of<T>(); }
            } else if (MODE == MODE_PROD) {
                ((T*)d.prod_out)[idx] = (nil || dead) ? T(0) : prod;
            } else {
                d.log_out[idx] = (nil || dead) ? -INFINITY : la.value();
                ((T*)d.phase_out)[idx] = (nil || dead) ? T(0) : (T)ph.s;
            }
        }
    }
}

template <typename T, int MODE>
static int c32x2_launch(const BatchDesc& d, hipStream_t s) {
    if (d.count <= 0) return 0;
    auto kern = pf_c32x2_kernel<T, MODE>;
    static DevCache resident;   // resident workgroups on the device
    int64_t grid = resident.get([&] {
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, 64, 0) != hipSuccess || occ <= 0) occ = 1;
        return occ * device_cu_count();
    });
    const int64_t pairs = (d.count + 1) / 2;
    if (grid > pairs) grid = pairs;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64), 0, s, d);
    return (int)hipGetLastError();
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

// Real batches of at least this many matrices take two per wave; smaller
// ones keep one per wave, whose shorter pair steps win while the grid does
// not fill the device (1024 32 x 32 float64: 0.014 ms against 0.021; 2^16:
// 0.358 against 0.336; 2^20: 5.5 against 4.64)
constexpr int64_t C32X2_MIN_COUNT = 8192;

int launch_c32(int dtype, int mode, const BatchDesc& d, hipStream_t s) {
    if (d.count >= C32X2_MIN_COUNT && (dtype == DT_F32 || dtype == DT_F64)) {
        if (mode == MODE_PROD) return dtype == DT_F64 ? c32x2_launch<double, MODE_PROD>(d, s) : c32x2_launch<float, MODE_PROD>(d, s);
        if (mode == MODE_SLOG) return dtype == DT_F64 ? c32x2_launch<double, MODE_SLOG>(d, s) : c32x2_launch<float, MODE_SLOG>(d, s);
    }
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
