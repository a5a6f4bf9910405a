// Batched blocked Parlett-Reid, slog in FP64 and complex128: one workgroup
// per matrix, with several workgroups resident per CU.
//
// The LDS-resident kernel (pf_lds.hip) keeps the whole triangle in LDS, so
// above n = 128 one matrix occupies a CU and its serial pair-step chain is
// exposed.  Here the matrix lives in global memory (the input for the first
// panel, then two per-workgroup slabs that alternate, both L2 / MALL
// traffic) and only the panel vectors live in LDS (2 nb n doubles), so
// several matrices share a CU and hide each other's latency:
//
//   panel (nb pair steps, left-looking inside the panel): thread t owns
//     column ("slot") k0 + t of the panel's starting matrix and keeps its
//     U_j, V_j in LDS.  Pair step i (k = k0 + 2i) reads row k+1 of the stale
//     matrix at the own slot and the pair's corrections
//       U_i[x] = (W[y1][x] + sum_{j<i} U_j[y1] V_j[x] - V_j[y1] U_j[x]) / a
//     and forms the next pivot row the same way (V_{i+1} = row k+2); the
//     pivot is a workgroup argmax (DPP per wave, then across the waves).
//     An interchange only changes maps: logical -> slot for the panel's own
//     positions (Qpan, LDS, double-buffered by step parity) and slot ->
//     logical per thread (a register).
//   trailing update (once per panel): C = stale[Q r][Q c] + A B with
//     A[r][2j] = U_j, A[r][2j+1] = V_j, B[2j][c] = V_j, B[2j+1][c] = -U_j
//     (the two terms of a pair adjacent in K), one v_mfma_f64_16x16x4f64
//     per 16 x 16 x 4 step (four per complex step), operands from LDS,
//     written compacted into the other slab.
//
// The arithmetic is the skew Schur complement S = A22 + (y x^T - x y^T)/a of
// the reference's pf_sm_calc (src/pfaffian_sm.cu:71-89) and the lazy row
// update of its large driver (src/slog_pfaffian_lg.cu:75-112), a pair step
// at a time with the pivot search of src/pfaffian_cpu.cpp:47-52 (largest
// magnitude, lowest index on ties, NaN never wins).
#include <cstdlib>
#include <type_traits>

#include "../../include/pfaffian_hip.h"
#include "pf_batch.h"
#include "pf_types.h"
#include "pf_wave.h"

// Pair steps per panel, per capacity: 2 nb NP doubles of LDS per workgroup
// (fewer workgroups per CU as nb grows) against one trailing pass per nb
// pairs.  Measured on the 16384-matrix batch (ms per bucket, nb 4 / 6 / 8 /
// 12 / 16 / 32): 64-128 0.59 / 0.47 / 0.39 / 0.43 / 0.58 / 1.05, 128-192
// 1.94 / 1.56 / 1.17 / 1.24 / 1.43 / 3.66, 192-256 5.03 / 3.56 / 2.76 / 2.60
// / 3.30 / 5.94.  Capping the workgroups per CU below the occupancy limit
// was slower at every width tried.
#ifndef PFH_BAT_NB_128
#define PFH_BAT_NB_128 8
#endif
#ifndef PFH_BAT_NB_192
#define PFH_BAT_NB_192 12   // the autotuner's choice (2048 matrices: 0.57 ms at 12, 0.66 at 8; 16384-batch 3.96 / 4.02 ms)
#endif
#ifndef PFH_BAT_NB_256
#define PFH_BAT_NB_256 12
#endif
// complex128: twice the LDS per pair.  Measured (ms per bucket, nb 4 / 5-6
// / 8 / 12): 64-128 0.96 / 0.83 (6) / 1.11 / 1.20; 128-192 3.88 / 2.74 (6) /
// 2.80; 192-256 7.75 / 7.04 (5) / 5.89 (8); nb = 8 is the widest that keeps
// two workgroups of NP = 256 on a CU.  The autotuner's 2048-matrix buckets
// later put 128-192 at 8 (ms, nb 4 / 6 / 8: 1.77 / 1.47 / 1.24).
#ifndef PFH_BAT_NB_C128_128
#define PFH_BAT_NB_C128_128 6
#endif
#ifndef PFH_BAT_NB_C128_192
#define PFH_BAT_NB_C128_192 8
#endif
#ifndef PFH_BAT_NB_C128_256
#define PFH_BAT_NB_C128_256 8
#endif

namespace pfh {

namespace {

typedef double bat_d4 __attribute__((ext_vector_type(4)));

// Element e of a matrix or slab whose base is wave-uniform, by a 32-bit byte
// offset (every offset within one matrix or slab stays far below 2^32), so
// that the access takes the scalar-base form with one 32-bit vector offset.
template <typename T>
__device__ __forceinline__ const T& bat_at(const T* W, unsigned e) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(W) + e * (unsigned)sizeof(T));
}

// Entry (r, c) of a matrix whose upper triangle is stored with row stride
// ld; the load is unconditional (the diagonal when r == c).
template <typename T>
__device__ __forceinline__ T bat_w(const T* W, int ld, int r, int c) {
    const int a = r < c ? r : c, b = r < c ? c : r;
    const T v = bat_at(W, (unsigned)(a * ld + b));
    return r < c ? v : (r > c ? -v : zero<T>());
}

template <typename T>
__device__ __forceinline__ T bat_sel(bool on, T v) { return on ? v : zero<T>(); }

// The same entry split in two, so that a run of loads can be issued before
// any of them is used: the element index, and whether the entry lies below
// the diagonal (its value is then the negated stored one).  On the diagonal
// (r == c) the stored value comes back as it is: the trailing update only
// reads an entry there for a result nothing reads.
__device__ __forceinline__ unsigned bat_ix(int ld, int r, int c, bool& below) {
    below = r > c;
    const int a = r < c ? r : c, b = r < c ? c : r;
    return (unsigned)(a * ld + b);
}
// v negated when `below`, as a flip of the sign bit(s)
__device__ __forceinline__ double bat_flip(bool below, double v) {
    return __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, v) ^
                                          ((unsigned long long)below << 63));
}
__device__ __forceinline__ c128 bat_flip(bool below, c128 v) { return c128{bat_flip(below, v.re), bat_flip(below, v.im)}; }

// Pairs of panel values the LDS holds: NB pairs over NP slots, except at
// the built-in widths of the largest buckets, where the capacity is the most
// that still lets one more workgroup onto a CU (160 KB of LDS per CU, the
// kernel's static LDS beside it): float64 NP = 256 four workgroups (2432
// pairs, 38.9 KB, against 3072), NP = 192 five (1920 against 2304),
// complex128 NP = 256 three (1600 against 2048).  The panels of the wider
// matrices start a pair or two narrower; the extra matrix per CU hides more
// of the pair steps' latency.  Measured on the 16384-matrix batch (ms per
// bucket, full capacity / this): f64 128-192 1.043 / 0.989, 192-256 2.183
// / 2.107 (4 workgroups still fit at 2464 pairs: 2.102; 5 do not at 1968 for
// NP = 192: 1.150); c128 192-256 5.305 / 5.127-5.186 (3 do not fit at 1648:
// 5.816).  For c128 NP = 192 four workgroups (1216 pairs) measured no faster
// (2.515 against 2.457-2.498 ms).
template <typename T, int NP, int NB>
__host__ __device__ constexpr int bat_capp() {
    if (!is_cplx<T>::value && NP == 256 && NB == 12) return 2432;
    if (!is_cplx<T>::value && NP == 192 && NB == 12) return 1920;
    if (is_cplx<T>::value && NP == 256 && NB == 8) return 1600;
    return NB * NP;
}

// Pairs per panel.  The LDS holds NB NP pairs of panel values (the
// instantiation's width NB over NP slots); a panel of live width w = n - k0
// stores its vectors with slot stride w and takes min(cap, NB NP / w) pairs,
// so the panels widen as the matrix shrinks (and a matrix narrower than NP
// starts wider): fewer trailing passes at the same occupancy, against longer
// chains of corrections in the pair steps, hence the cap.  Measured on the
// 16384-matrix batch (ms, f64 buckets 64-128 / 128-192 / 192-256 and the
// whole call; the fixed widths before: 0.380 / 1.112 / 2.256, 3.775):
//   cap 12: 0.344 / 1.042 / 2.180, 3.673;  16: 0.375 / 1.048 / 2.166, 3.700;
//   20: 0.380 / 1.095 / 2.209, 3.756;  32: 0.377 / 1.146 / 2.322, 3.928;
// complex128 (fixed widths 0.755 / 2.499 / 5.450, 8.867) at a cap set at run
// time (about 1% slower than a compile-time cap for float64): 8: 0.744 /
// 2.459 / 5.384, 8.764;  10: 0.764 / 2.466 / 5.301, 8.705;  12: 0.786 / 2.514
// / 5.363, 8.828;  16: 0.811 / 2.600 / 5.503, 9.080.
template <typename T, int NB>
__host__ __device__ constexpr int bat_cap() {
    constexpr int c = is_cplx<T>::value ? 10 : 12;
    return NB > c ? NB : c;
}

// Tiles in flight per wave in the trailing update (the slots of its
// software pipeline).  Measured on the 16384-matrix f64 batch: 2 slots 3.28
// ms, 3 slots 3.31, 4 slots 3.32 (complex128 holds two: a third would cost a
// wave per SIMD).
template <typename T>
constexpr int bat_slots() { return 2; }

// pivot-search magnitude: |x| for real, |z|^2 for complex (NaN stays NaN)
__device__ __forceinline__ double bat_mag(double x) { return fabs(x); }
__device__ __forceinline__ double bat_mag(c128 z) { return z.re * z.re + z.im * z.im; }

// workgroup argmax, part 1: the wave's best (|value|, key) into LDS
__device__ __forceinline__ void bat_wave_best(double best, int key, double* r_best, int* r_key) {
    double b;
    int k;
    wave_argmax(best, key, b, k);                                               // ties: lowest key
    if ((threadIdx.x & 63) == 0) { r_best[threadIdx.x >> 6] = b; r_key[threadIdx.x >> 6] = k; }
}

}  // namespace

// Waves per SIMD the register allocation must allow: complex128 needs three
// for its occupancy-sized LDS capacity at NP = 256 (three workgroups of four
// waves per CU); float64 fits five unconstrained.
template <typename T>
constexpr int bat_waves() { return is_cplx<T>::value ? 3 : 1; }

template <typename T, int NP, int NB>
__global__ void __launch_bounds__(NP) __attribute__((amdgpu_waves_per_eu(bat_waves<T>(), 8)))
pf_bat_kernel(BatchDesc d) {
    constexpr int NW = NP / 64;
    constexpr int CAPP = bat_capp<T, NP, NB>();   // pairs of panel values the LDS holds
    constexpr int BCAP = bat_cap<T, NB>();   // pairs per panel at most
    extern __shared__ double bat_lds[];     // U_j, V_j of local slot s at [2 (j w + s)] and [2 (j w + s) + 1]
    T* Pp = reinterpret_cast<T*>(bat_lds);
    __shared__ int Qpan[2][2 * BCAP + 2];   // slot of logical k0 + q, by pair-step parity
    __shared__ int Qs[NP];                  // trailing update: local slot of logical lo + r
    __shared__ double r_best[2][NW];        // per-wave argmax, by pair-step parity
    __shared__ int r_key[2][NW];

    // (the wave index through readfirstlane: the compiler does not know it
    // is uniform, and the trailing update's tile walk, which starts from
    // it, then ran as a per-lane loop on the vector ALU; 16384-matrix
    // batch: complex128 8.03 -> 7.95 ms, float64 unchanged)
    const int t = threadIdx.x, lane = t & 63, wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const int lr = lane & 15, lq = lane >> 4;
    T* slab0 = reinterpret_cast<T*>(d.bslab) + (int64_t)blockIdx.x * 2 * NP * NP;   // this workgroup's slab pair
    T* slab1 = slab0 + (int64_t)NP * NP;
    T* phase_out = reinterpret_cast<T*>(d.phase_out);

    int64_t count = d.count;
    if (d.count_dev) {
        const int64_t cd = *d.count_dev;
        count = cd < count ? cd : count;
    }
    // Matrices in the order given (a bucket's list is largest first): the
    // first by block index, then, with a queue, the next unclaimed one (the
    // longest-first greedy schedule), else a fixed stride.
    __shared__ int64_t s_next;
    auto next = [&](int64_t jm) -> int64_t {
        if (!d.queue) return jm + gridDim.x;
        __syncthreads();
        if (t == 0) s_next = (int64_t)gridDim.x + atomicAdd(d.queue, 1);
        __syncthreads();
        return s_next;
    };
    for (int64_t jm = blockIdx.x; jm < count; jm = next(jm)) {
        const int idx = d.order ? d.order[(d.begin_dev ? *d.begin_dev : 0) + jm] : (int)jm;
        const int n = __builtin_amdgcn_readfirstlane(d.sizes ? d.sizes[idx] : d.dense_n);
        const int64_t off = d.offsets ? d.offsets[idx] : (int64_t)idx * n * n;
        if (n < 0 || n > NP) {
            if (t == 0) { d.log_out[idx] = __builtin_nan(""); phase_out[idx] = nan_of<T>(); }
            continue;
        }
        if (n == 0 || (n & 1)) {
            if (t == 0) { d.log_out[idx] = -INFINITY; phase_out[idx] = zero<T>(); }
            continue;
        }
        const T* src = reinterpret_cast<const T*>(d.data) + off;
        int ld = n;
        T* dst = slab0;
        LogAcc la;
        Phase<T> ph;
        bool dead = false;

        for (int k0 = 0, np = 0; k0 < n; k0 += 2 * np) {
            const int w = n - k0;            // live width: the slot stride of the panel vectors
            np = min(min(BCAP, CAPP / w), w / 2);      // >= 1: w >= 2, CAPP > w
            const int x = k0 + t;            // slot: index in src
            const bool own = x < n;
            const int s = own ? t : 0;       // local slot (a valid one for masked reads)
            const int xs = own ? x : k0;     // a valid slot for masked reads
            int lg = x;                      // logical index of the own slot
            __syncthreads();                 // the previous panel's LDS reads are done
            for (int q = t; q < 2 * np + 2; q += NP) Qpan[0][q] = k0 + q;
            if (own) {
#pragma unroll 4
                for (int j = 0; j < np; ++j) { Pp[2 * (j * w + s)] = zero<T>(); Pp[2 * (j * w + s) + 1] = zero<T>(); }
            }
            // first pivot row: row k0 of src (nothing pending)
            {
                double best = -1.0;
                const T r0v = bat_w(src, ld, k0, xs);
                const T r = bat_sel(own && x > k0, r0v);
                if (own) Pp[2 * s + 1] = r;                          // V_0
                const double m = bat_mag(r);
                if (own && x > k0 && m == m) best = m;
                bat_wave_best(best, lg * 512 + s, r_best[0], r_key[0]);
                __syncthreads();
            }
            for (int i = 0; i < np; ++i) {
                const int k = k0 + 2 * i;
                const int cb = i & 1;
                const bool has_next = i + 1 < np;
                double gb = r_best[cb][0];
                int gk = r_key[cb][0];
#pragma unroll
                for (int u = 1; u < NW; ++u)
                    if (r_best[cb][u] > gb || (r_best[cb][u] == gb && r_key[cb][u] < gk)) { gb = r_best[cb][u]; gk = r_key[cb][u]; }
                const int sa = Qpan[cb][k + 1 - k0];
                const int sb = has_next ? Qpan[cb][k + 2 - k0] : sa;
                int p, sp;
                if (gb >= 0.0) { p = gk >> 9; sp = k0 + (gk & 511); }
                else { p = k + 1; sp = sa; }                         // no finite candidate (NaN row)
                const bool sw = p != k + 1;
                const int y1 = sp, y2 = (p == k + 2) ? sa : sb;
                // the stale rows k+1 and k+2 at the own slot go out first (the
                // addresses are valid whatever the pivot's value, and on the
                // last pair of the matrix too); the pivot's value and the
                // zero test follow, in the shadow of these loads
                const T w1 = bat_w(src, ld, y1, xs);
                const T w2 = bat_w(src, ld, y2, xs);
                const T w12 = bat_w(src, ld, y1, y2);
                const int s1 = y1 - k0, s2 = y2 - k0;
                const T piv = Pp[2 * (i * w + s1) + 1];              // row k at column p
                const T vi = Pp[2 * (i * w + s) + 1];                // V_i at the own slot
                const T v_kl = Pp[2 * (i * w + s2) + 1];             // V_i at row k+2
                // the next step's map: interchange k+1 <-> p
                for (int q = t; q < 2 * np + 2; q += NP) {
                    int v = Qpan[cb][q];
                    if (q == k + 1 - k0) v = sp;
                    else if (sw && q == p - k0) v = sa;
                    Qpan[cb ^ 1][q] = v;
                }
                if (x == sp) lg = k + 1;
                else if (sw && x == sa) lg = p;
                const T inv = recip(piv);                            // inf for a zero pivot, never used then
                T r1 = zero<T>(), rl = zero<T>(), rk = zero<T>();
                for (int j = 0; j < i; ++j) {
                    const T uj = Pp[2 * (j * w + s)], vj = Pp[2 * (j * w + s) + 1];
                    const T b1u = Pp[2 * (j * w + s1)], b1v = Pp[2 * (j * w + s1) + 1];
                    const T b2u = Pp[2 * (j * w + s2)], b2v = Pp[2 * (j * w + s2) + 1];
                    r1 = r1 + b1u * vj - b1v * uj;
                    rl = rl + b2u * vj - b2v * uj;
                    rk = rk + b1u * b2v - b1v * b2u;
                }
                const T u = bat_sel(own && lg >= k + 2, (w1 + r1) * inv);
                if (is_zero(piv)) { dead = true; break; }            // the same decision in every thread
                la.mul(absd(piv));
                ph.mul(piv);
                if (sw) ph.flip();
                if (k + 2 >= n) break;                               // last pair of the matrix
                if (own) Pp[2 * (i * w + s)] = u;                    // U_i
                if (!has_next) break;
                const T u_kl = (w12 + rk) * inv;                     // U_i at row k+2
                double best = -1.0;
                T r = zero<T>();
                if (own && lg >= k + 3) {
                    r = w2 + rl + u_kl * vi - v_kl * u;
                    const double m = bat_mag(r);
                    if (m == m) best = m;
                }
                if (own) Pp[2 * ((i + 1) * w + s) + 1] = r;          // V_{i+1} = row k+2
                bat_wave_best(best, lg * 512 + s, r_best[cb ^ 1], r_key[cb ^ 1]);
                __syncthreads();
            }
            if (dead) break;
            const int lo = k0 + 2 * np;
            if (lo >= n) break;

            // ---- trailing update of [lo, n) into dst, 16 x 16 upper tiles ----
            const int m = n - lo;
            const int mt = (m + 15) >> 4;
            if (own && lg >= lo) Qs[lg - lo] = s;
            // positions m .. 16 mt - 1 (rows and columns past the matrix in
            // its last tile row and column) point at slot 0: valid reads
            // for results nothing reads, so that no index needs a clamp
            if (t >= m && t < 16 * mt) Qs[t] = 0;
            __syncthreads();
            const int ntile = mt * (mt + 1) / 2;
            auto tile_ij = [&](int tt, int& I, int& J) {
                int rem = tt;
                I = 0;
                while (rem >= mt - I) { rem -= mt - I; ++I; }
                J = I + rem;
            };
            // raw loads now, signs applied where the values are used
            auto gather = [&](int I, int J, T (&w4)[4], bool (&bl4)[4]) {
                const int scc = Qs[16 * J + lr] + k0;
                unsigned ix[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) ix[e] = bat_ix(ld, Qs[16 * I + lq + 4 * e] + k0, scc, bl4[e]);
#pragma unroll
                for (int e = 0; e < 4; ++e) w4[e] = bat_at(src, ix[e]);
            };
            // K runs over the panel's pairs in groups of two (one MFMA K
            // range of four: U_j, V_j, U_{j+1}, V_{j+1}), GU groups at a
            // time (four real, two complex: the complex form's registers
            // would otherwise cost a wave per SIMD), every operand read of
            // the GU issued ahead of their products; an odd pair count ends
            // with a masked group.  (Splitting the product in two, U V^T and
            // V U^T, one accumulation each, so that a lane reads its pair with
            // one LDS access per side and negates nothing, measured slower:
            // groups of four pairs leave a panel of ten pairs a half-masked
            // third group, f64 192-256 bucket 1.91 -> 2.05 ms.)
            const int gfull = np >> 1;
            constexpr int GU = is_cplx<T>::value ? 2 : 4;
            // one tile: the products, then the stale entries added and the
            // results stored
            auto tile = [&](int I, int J, const T (&w4)[4], const bool (&bl4)[4]) {
                const int cc = 16 * J + lr;
                const int sA = Qs[16 * I + lr], sB = Qs[cc];
                bat_d4 acc = {0.0, 0.0, 0.0, 0.0};
                bat_d4 aci = {0.0, 0.0, 0.0, 0.0};     // complex: the imaginary part
                // lane (lr, lq) takes K index lq of each group: pair
                // 2 g + lq / 2, its U (lq even) or V; complex products as
                // four real ones:
                //   Re += Re a Re b - Im a Im b,  Im += Re a Im b + Im a Re b
                const int par = lq & 1;
                const T* pa = Pp + 2 * ((lq >> 1) * w + sA) + par;          // U_j, V_j at the row
                const T* pb = Pp + 2 * ((lq >> 1) * w + sB) + (par ^ 1);    // V_j, U_j at the column
                auto group = [&](T a, T b0) {
                    const T b = par ? -b0 : b0;
                    if constexpr (is_cplx<T>::value) {
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.re, acc, 0, 0, 0);
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.im, b.im, acc, 0, 0, 0);
                        aci = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.im, aci, 0, 0, 0);
                        aci = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, b.re, aci, 0, 0, 0);
                    } else {
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
                    }
                };
                const int gs = 4 * w;                 // elements of T per group of two pairs
                int g = 0;
                for (; g + GU <= gfull; g += GU) {
                    T a4[GU], b4[GU];
#pragma unroll
                    for (int u = 0; u < GU; ++u) { a4[u] = pa[(g + u) * gs]; b4[u] = pb[(g + u) * gs]; }
#pragma unroll
                    for (int u = 0; u < GU; ++u) group(a4[u], b4[u]);
                }
                for (; g < gfull; ++g) group(pa[g * gs], pb[g * gs]);
                if (np & 1) {                         // the last pair alone: lanes lq >= 2 masked
                    const bool on = lq < 2;
                    const int o = g * gs - (on ? 0 : 2 * w);   // masked lanes read pair np - 1
                    group(bat_sel(on, pa[o]), bat_sel(on, pb[o]));
                }
                // unconditional stores (a store under a branch made the
                // compiler wait for every memory operation in flight, this
                // wave's next gathers included, before reusing its data
                // registers); the four values are formed first so that they
                // sit in four register pairs and the stores issue back to
                // back.  Rows past the matrix (rr < 16 mt <= NP) and the
                // trash slot lie inside the slab.
                T ov[4];
                unsigned ix[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const T stale = bat_flip(bl4[e], w4[e]);
                    if constexpr (is_cplx<T>::value) ov[e] = T{stale.re + acc[e], stale.im + aci[e]};
                    else ov[e] = stale + acc[e];
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) ix[e] = (unsigned)((16 * I + lq + 4 * e) * NP + cc);
                // a tile on the diagonal or past the matrix's last row or
                // column sends its entries outside the new upper triangle to
                // one slot nothing reads (row NP - 1, column 0: below the
                // diagonal), so that their lines are not written
                if (!(I < J && 16 * J + 16 <= m)) {
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int rr = 16 * I + lq + 4 * e;
                        if (!(cc > rr && cc < m)) ix[e] = (unsigned)((NP - 1) * NP);
                    }
                }
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    *reinterpret_cast<T*>(reinterpret_cast<char*>(dst) + ix[e] * (unsigned)sizeof(T)) = ov[e];
            };
            // Wave wv takes tiles wv, wv + NW, ... (q = 0 .. ntw - 1) in two
            // slots, each tile's stale entries gathered one tile ahead: the
            // steady-state loop handles two tiles with no branch around any
            // memory access (the gathers past the last tile read the last
            // tile again), so the compiler's waits are for the tile at hand
            // only, the other slot's gathers stay in flight.
            const int ntw = wv < ntile ? (ntile - wv + NW - 1) / NW : 0;
            if (ntw > 0) {
                constexpr int S = bat_slots<T>();
                auto tix = [&](int q) { return min(wv + q * NW, ntile - 1); };
                int TI[S], TJ[S];
                T WW[S][4];
                bool SS[S][4];
#pragma unroll
                for (int u = 0; u < S; ++u) {
                    tile_ij(tix(u), TI[u], TJ[u]);
                    gather(TI[u], TJ[u], WW[u], SS[u]);
                }
                int q = 0;
                for (; q + S <= ntw; q += S) {
#pragma unroll
                    for (int u = 0; u < S; ++u) {
                        tile(TI[u], TJ[u], WW[u], SS[u]);
                        tile_ij(tix(q + S + u), TI[u], TJ[u]);
                        gather(TI[u], TJ[u], WW[u], SS[u]);
                    }
                }
#pragma unroll
                for (int u = 0; u < S; ++u)
                    if (q + u < ntw) tile(TI[u], TJ[u], WW[u], SS[u]);
            }
            // the trailing matrix sits at the slab's origin, so that a tile's
            // 16-entry row segments are whole 128-byte lines; src is offset
            // to read it by matrix indices (all >= lo from here on)
            src = dst - (lo * NP + lo);
            ld = NP;
            dst = (dst == slab0) ? slab1 : slab0;
        }
        if (t == 0) {
            d.log_out[idx] = dead ? -INFINITY : la.value();
            T pv;
            if constexpr (is_cplx<T>::value) pv = T{ph.re, ph.im};
            else pv = ph.s;
            phase_out[idx] = dead ? zero<T>() : pv;
        }
    }
}

namespace {

int bat_num_cu() { return device_cu_count(); }

template <typename T, int NP, int NB>
constexpr size_t bat_lds_bytes() { return sizeof(T) * 2 * bat_capp<T, NP, NB>(); }

// resident grid (workgroups per CU x CUs); also sets the kernel's LDS limit,
// which the occupancy query needs
template <typename T, int NP, int NB>
int bat_grid() {
    static DevCache resident;
    return resident.get([] {
        (void)hipFuncSetAttribute((const void*)pf_bat_kernel<T, NP, NB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)bat_lds_bytes<T, NP, NB>());
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, pf_bat_kernel<T, NP, NB>, NP,
                                                         bat_lds_bytes<T, NP, NB>()) != hipSuccess ||
            occ <= 0)
            occ = 1;
        return occ * bat_num_cu();
    });
}

// doubles of one workgroup's slab pair
template <typename T, int NP>
constexpr int64_t bat_per() { return 2 * (int64_t)NP * NP * (int64_t)(sizeof(T) / sizeof(double)); }

template <typename T, int NP, int NB>
int bat_launch(const BatchDesc& d, hipStream_t s) {
    if (d.count <= 0) return 0;
    auto kern = pf_bat_kernel<T, NP, NB>;
    constexpr size_t bytes = bat_lds_bytes<T, NP, NB>();
    static DevCache attr;   // the LDS limit is set on the device
    hipError_t e = hipSuccess;
    if (!attr.get([&] {
            e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
            return e == hipSuccess ? 1 : 0;
        }))
        return (int)e;
    int64_t grid = bat_grid<T, NP, NB>();
    constexpr int64_t per = bat_per<T, NP>();   // one slab pair per workgroup
    if (!d.bslab || d.bslab_elems < per) return (int)hipErrorInvalidValue;
    if (grid > d.bslab_elems / per) grid = d.bslab_elems / per;
    if (grid > d.count) grid = d.count;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(NP), bytes, s, d);
    return (int)hipGetLastError();
}

template <typename T, int NP, int NB>
int64_t bat_need(int64_t count) {
    const int64_t g = bat_grid<T, NP, NB>();
    return (count < g ? count : g) * bat_per<T, NP>();
}

}  // namespace

// Pairs per panel of each bucket's instantiation: the built-in width
// (PFH_BAT_NB_*) unless its tuning key names another instantiated one
// (float64: 6, 8 or 12 up to 192 and 8, 10 or 12 above; complex128: 4, 6
// or 8, and 6 or 8 above 192).
static int bat_nb(int dtype, int np) {
    if (dtype == DT_C128) {
        const int key = np == 128 ? PFH_TUNE_BAT_NB_C128_128 : np == 192 ? PFH_TUNE_BAT_NB_C128_192
                                                                          : PFH_TUNE_BAT_NB_C128_256;
        const int dflt = np == 128 ? PFH_BAT_NB_C128_128 : np == 192 ? PFH_BAT_NB_C128_192 : PFH_BAT_NB_C128_256;
        const int tv = tune_value(key);
        if (np < 256 && (tv == 4 || tv == 6 || tv == 8)) return tv;
        if (np == 256 && (tv == 6 || tv == 8)) return tv;
        return dflt;
    }
    const int key = np == 128 ? PFH_TUNE_BAT_NB_128 : np == 192 ? PFH_TUNE_BAT_NB_192 : PFH_TUNE_BAT_NB_256;
    const int dflt = np == 128 ? PFH_BAT_NB_128 : np == 192 ? PFH_BAT_NB_192 : PFH_BAT_NB_256;
    const int tv = tune_value(key);
    if (np < 256 && (tv == 6 || tv == 8 || tv == 12)) return tv;
    if (np == 256 && (tv == 8 || tv == 10 || tv == 12)) return tv;
    return dflt;
}

// f(std::integral_constant<int, NB>) for the bucket's width nb, over the
// widths instantiated for (T, NP) (those bat_nb can return)
template <typename T, int NP, typename F>
static auto bat_nb_dispatch(int nb, F f) {
    using I4 = std::integral_constant<int, 4>;
    using I6 = std::integral_constant<int, 6>;
    using I8 = std::integral_constant<int, 8>;
    using I10 = std::integral_constant<int, 10>;
    using I12 = std::integral_constant<int, 12>;
    if constexpr (std::is_same<T, double>::value) {
        if constexpr (NP < 256) {
            if (nb == 6) return f(I6{});
            if (nb == 12) return f(I12{});
            return f(I8{});
        } else {
            if (nb == 8) return f(I8{});
            if (nb == 10) return f(I10{});
            return f(I12{});
        }
    } else {
        if constexpr (NP < 256) {
            if (nb == 4) return f(I4{});
            if (nb == 8) return f(I8{});
            return f(I6{});
        } else {
            if (nb == 6) return f(I6{});
            return f(I8{});
        }
    }
}

template <typename T, int NP>
static int64_t bat_need_rt(int dtype, int64_t count) {
    return bat_nb_dispatch<T, NP>(bat_nb(dtype, NP), [&](auto nbc) -> int64_t {
        return bat_need<T, NP, decltype(nbc)::value>(count);
    });
}

template <typename T, int NP>
static int bat_launch_rt(int dtype, const BatchDesc& d, hipStream_t s) {
    return bat_nb_dispatch<T, NP>(bat_nb(dtype, NP), [&](auto nbc) -> int {
        return bat_launch<T, NP, decltype(nbc)::value>(d, s);
    });
}

// doubles of workspace for batches of `count` matrices with n <= maxn: one
// slab pair (2 NP^2 elements) per resident workgroup of the widest
// instantiation a bucket can need (the buckets run one after another), at
// the panel widths in force when the size is asked for
int64_t bat_ws_elems(int dtype, int maxn, int64_t count) {
    if (maxn <= 64 || count <= 0) return 0;
    if (dtype == DT_C128) {
        int64_t v = bat_need_rt<c128, 128>(dtype, count);
        if (maxn > 128) { const int64_t b = bat_need_rt<c128, 192>(dtype, count); v = b > v ? b : v; }
        if (maxn > 192) { const int64_t c = bat_need_rt<c128, 256>(dtype, count); v = c > v ? c : v; }
        return v;
    }
    if (dtype != DT_F64) return 0;
    int64_t v = bat_need_rt<double, 128>(dtype, count);
    if (maxn > 128) { const int64_t b = bat_need_rt<double, 192>(dtype, count); v = b > v ? b : v; }
    if (maxn > 192) { const int64_t c = bat_need_rt<double, 256>(dtype, count); v = c > v ? c : v; }
    return v;
}

int launch_bat(int dtype, const BatchDesc& d, int maxn, hipStream_t s) {
    if (dtype == DT_C128) {
        if (maxn <= 128) return bat_launch_rt<c128, 128>(dtype, d, s);
        if (maxn <= 192) return bat_launch_rt<c128, 192>(dtype, d, s);
        if (maxn <= 256) return bat_launch_rt<c128, 256>(dtype, d, s);
        return (int)hipErrorInvalidValue;
    }
    if (dtype != DT_F64) return (int)hipErrorInvalidValue;
    if (maxn <= 128) return bat_launch_rt<double, 128>(dtype, d, s);
    if (maxn <= 192) return bat_launch_rt<double, 192>(dtype, d, s);
    if (maxn <= 256) return bat_launch_rt<double, 256>(dtype, d, s);
    return (int)hipErrorInvalidValue;
}

}  // namespace pfh
