// Workgroup-per-matrix Parlett-Reid kernels with the matrix resident in
// VGPRs (and, for the largest f64 bucket, its first rows in LDS).
//
// Replaces the reference's one-matrix-per-32-thread-block kernel
// pf_sm_calc (src/pfaffian_sm.cu:3-99), launched once per matrix from a
// serial host loop (bindings/jax_bindings.cu:107-109), with one persistent
// launch per size bucket that walks a list of matrices.
//
// Layout (NCOL column lanes x NRB row blocks = threads):
//   thread t holds physical column c = t % NCOL, row block rb = t / NCOL;
//   logical row r >= LR lives in register (r-LR)/NRB of block (r-LR)%NRB
//   (rows interleaved so that the register index of a row is wave-uniform);
//   rows r < LR live in LDS, row r owned by block r % NRB.  Rows are swapped
//   physically (each by its owning threads); columns carry a
//   logical label L(c) instead, so a column swap costs nothing.
//
// Per pair step k (right-looking, as pf_sm_calc :29-91):
//   1. the owners of row k write it to X[L(c)]             (barrier B1)
//   2. every wave reduces argmax_{j>k} |X[j]| (ties: lowest j) -> p, pivot
//   3. rows k+1 and p swap storage slots, labels k+1 <-> p swap; the new
//      row k+1 goes to Y[L(c)]                             (barrier B2)
//   4. rank-2 Schur update of rows r >= k+2 in every column:
//        A[r][c] += (Y[r] * X[c] - X[r] * Y[c]) / pivot
//      (the same update as pfaffian_sm.cu:71-89, written for the skew
//      Schur complement S = A22 + (y x^T - x y^T)/a).
// Lanes whose column is already eliminated compute values nobody reads, so
// the update needs no column mask.  Pf = (-1)^swaps * prod(pivots); the
// exact zero-pivot test returns Pf = 0 (log -inf, phase 0).
#include "pf_batch.h"
#include "pf_regstore.h"

namespace pfh {

template <typename T>
__device__ __forceinline__ T mul_add_update(T a, T Yr, T xs, T Xr, T ys) {
    return a + Yr * xs - Xr * ys;
}

template <typename T, int NCOL, int NRB, int LR, int MODE>
__global__ void __launch_bounds__(NCOL * NRB)
pf_reg_kernel(BatchDesc d) {
    static_assert((NCOL - LR) % NRB == 0, "register rows must tile the column count exactly");
    constexpr int RB = (NCOL - LR) / NRB;
    using R = typename real_of<T>::type;
    struct XY { T x, y; };
    __shared__ T Xs[NCOL];      // row k by old labels (pivot search)
    __shared__ XY XYs[NCOL];    // rows k, k+1 by new labels (update)
    __shared__ T EXs[NCOL];
    __shared__ T Ls[LR > 0 ? LR : 1][NCOL];

    const int t = threadIdx.x;
    const int c = t % NCOL;
    const int rb = (NCOL % 64 == 0) ? __builtin_amdgcn_readfirstlane(t / NCOL) : t / NCOL;
    const int lane = t & 63;

    int64_t count = d.count;
    if (d.count_dev) {
        int64_t cd = *d.count_dev;
        count = cd < count ? cd : count;
    }
    for (int64_t j = blockIdx.x; j < count; j += gridDim.x) {
        const int idx = d.order ? d.order[(d.begin_dev ? *d.begin_dev : 0) + j] : (int)j;
        const int n = __builtin_amdgcn_readfirstlane(d.sizes ? d.sizes[idx] : d.dense_n);
        const T* base = (const T*)d.data + (d.offsets ? d.offsets[idx] : (int64_t)idx * n * n);
        __syncthreads();   // LDS of the previous matrix is free

        if (n < 0 || n > NCOL) {   // size outside this kernel's capacity: NaN, no access
            if (t == 0) {
                const T bad = nan_of<T>();
                if (MODE == MODE_PROD) ((T*)d.prod_out)[idx] = bad;
                else { d.log_out[idx] = __builtin_nan(""); ((T*)d.phase_out)[idx] = bad; }
            }
            continue;
        }
        if (n == 0 || (n & 1)) {
            if (t == 0) {
                if (MODE == MODE_PROD) ((T*)d.prod_out)[idx] = zero<T>();
                else { d.log_out[idx] = -INFINITY; ((T*)d.phase_out)[idx] = zero<T>(); }
            }
            continue;
        }

        Regs<T, RB> a;
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            const int r = LR + i * NRB + rb;
            a.set(i, (r < n && c < n) ? base[(int64_t)r * n + c] : zero<T>());
        }
        if (LR > 0) {
            for (int r = rb; r < LR; r += NRB)
                Ls[r][c] = (r < n && c < n) ? base[(int64_t)r * n + c] : zero<T>();
        }

        int L = c;
        T prod = one<T>();
        LogAcc la;
        Phase<T> ph;
        bool dead = false;

        for (int k = 0; k < n; k += 2) {
            // 1. row k -> X
            T x = zero<T>();
            bool ownk;
            if (k < LR) {
                ownk = (rb == k % NRB);
                if (ownk) { x = Ls[k][c]; Xs[L] = x; }
            } else {
                const int kk = k - LR;
                ownk = (rb == kk % NRB);
                if (ownk) { x = a.dget(kk / NRB); Xs[L] = x; }
            }
            __syncthreads();   // B1

            // 2. pivot search over logical columns k+1 .. n-1
            double best = -1.0;
            int bi = 0x7fffffff;
            for (int jj = k + 1 + lane; jj < n; jj += 64) {
                const double m = (double)mag2(Xs[jj]);
                if (m > best) { best = m; bi = jj; }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double ob = __shfl_xor(best, off);
                const int oi = __shfl_xor(bi, off);
                if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
            }
            int p = __builtin_amdgcn_readfirstlane(bi);
            if (p >= n) p = k + 1;              // only NaN candidates: keep k+1
            const T piv = Xs[p];
            const T xc = Xs[L];
            if (is_zero(piv)) { dead = true; break; }
            const bool sw = (p != k + 1);

            // 3. swap rows k+1 <-> p, relabel columns
            const int r1 = k + 1;
            T v1 = zero<T>(), vp = zero<T>();
            bool own1, ownp;
            if (r1 < LR) { own1 = (rb == r1 % NRB); if (own1) v1 = Ls[r1][c]; }
            else { own1 = (rb == (r1 - LR) % NRB); if (own1) v1 = a.dget((r1 - LR) / NRB); }
            if (p < LR) { ownp = (rb == p % NRB); if (ownp) vp = Ls[p][c]; }
            else { ownp = (rb == (p - LR) % NRB); if (ownp) vp = a.dget((p - LR) / NRB); }
            int Lnew = L;
            if (sw) { if (L == r1) Lnew = p; else if (L == p) Lnew = r1; }
            if (ownk) XYs[Lnew].x = x;
            if (ownp) XYs[Lnew].y = vp;
            if (sw && own1) EXs[c] = v1;
            __syncthreads();   // B2
            if (sw) {
                if (own1) {
                    const T nv = XYs[Lnew].y;
                    if (r1 < LR) Ls[r1][c] = nv; else a.dset((r1 - LR) / NRB, nv);
                }
                if (ownp) {
                    const T ov = EXs[c];
                    if (p < LR) Ls[p][c] = ov; else a.dset((p - LR) / NRB, ov);
                }
            }
            L = Lnew;
            const T yc = XYs[L].y;

            if (MODE == MODE_PROD) { prod = prod * piv; if (sw) prod = -prod; }
            else { la.mul(absd(piv)); ph.mul(piv); if (sw) ph.flip(); }

            // 4. rank-2 update of rows >= k+2
            const T inv = recip(piv);
            const T xs = xc * inv, ys = yc * inv;
            // LDS row r always belongs to thread block r % NRB, the same thread
            // that wrote it in the interchange above, so no barrier is needed.
            if (LR > 0 && k + 2 < LR) {
                const int r_first = k + 2 + ((rb - (k + 2)) % NRB + NRB) % NRB;
                for (int r = r_first; r < LR && r < n; r += NRB) {
                    const XY q = XYs[r];
                    Ls[r][c] = mul_add_update(Ls[r][c], q.y, xs, q.x, ys);
                }
            }
            // Register rows in groups of G with one uniform guard per group:
            // rows already eliminated or past n inside a live group take a
            // harmless update (nothing reads them again), and the group's LDS
            // reads issue back to back.
            constexpr int G = 8;
#pragma unroll
            for (int g = 0; g < RB; g += G) {
                const int lo = LR + g * NRB;                          // lowest row of the group
                const int hi = LR + (g + G - 1) * NRB + NRB - 1;      // highest row of the group
                if (hi >= k + 2 && lo < n) {
#pragma unroll
                    for (int i = g; i < g + G && i < RB; ++i) {
                        const XY q = XYs[LR + i * NRB + rb];
                        a.set(i, mul_add_update(a.get(i), q.y, xs, q.x, ys));
                    }
                }
            }
        }

        if (t == 0) {
            if (MODE == MODE_PROD) {
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
// launch configuration tables
// ---------------------------------------------------------------------------

static int g_num_cu = 0;

static int num_cu() {
    if (g_num_cu == 0) {
        int dev = 0;
        hipGetDevice(&dev);
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
        g_num_cu = v;
    }
    return g_num_cu;
}

template <typename T, int NCOL, int NRB, int LR, int MODE>
static int launch_one(const BatchDesc& d, hipStream_t stream) {
    if (d.count <= 0) return 0;
    auto kern = pf_reg_kernel<T, NCOL, NRB, LR, MODE>;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, NCOL * NRB, 0) != hipSuccess || occ <= 0) occ = 1;
    int64_t grid = (int64_t)occ * num_cu();
    if (grid > d.count) grid = d.count;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(NCOL * NRB), 0, stream, d);
    return (int)hipGetLastError();
}

int reg_max_n(int dtype) {
    switch (dtype) {
        case DT_F64: return 256;
        case DT_F32: return 256;
        case DT_C64: return 192;
        case DT_C128: return 128;
    }
    return 0;
}

template <typename T, int MODE>
static int launch_typed(const BatchDesc& d, int maxn, hipStream_t s);

#define PFH_CFG(NC, NB, LRR) if (maxn <= NC) return launch_one<T, NC, NB, LRR, MODE>(d, s);

template <> int launch_typed<double, MODE_SLOG>(const BatchDesc& d, int maxn, hipStream_t s) {
    using T = double; constexpr int MODE = MODE_SLOG;
    PFH_CFG(32, 2, 0) PFH_CFG(64, 4, 0) PFH_CFG(128, 4, 0) PFH_CFG(192, 4, 0) PFH_CFG(256, 4, 64)
    return (int)hipErrorInvalidValue;
}
template <> int launch_typed<float, MODE_SLOG>(const BatchDesc& d, int maxn, hipStream_t s) {
    using T = float; constexpr int MODE = MODE_SLOG;
    PFH_CFG(32, 2, 0) PFH_CFG(64, 4, 0) PFH_CFG(128, 4, 0) PFH_CFG(256, 4, 0)
    return (int)hipErrorInvalidValue;
}
template <> int launch_typed<c128, MODE_SLOG>(const BatchDesc& d, int maxn, hipStream_t s) {
    using T = c128; constexpr int MODE = MODE_SLOG;
    PFH_CFG(32, 2, 0) PFH_CFG(64, 4, 0) PFH_CFG(128, 8, 0)
    return (int)hipErrorInvalidValue;
}
template <> int launch_typed<c64, MODE_SLOG>(const BatchDesc& d, int maxn, hipStream_t s) {
    using T = c64; constexpr int MODE = MODE_SLOG;
    PFH_CFG(32, 2, 0) PFH_CFG(64, 4, 0) PFH_CFG(128, 4, 0) PFH_CFG(192, 4, 0)
    return (int)hipErrorInvalidValue;
}
// Product form (pfaffian()): the reference caps it at n <= 32
// (pfcuda/cuda_api.py:44-45); this library takes n <= 64, one wave per
// 32-column matrix and four waves per 64-column matrix.
template <> int launch_typed<double, MODE_PROD>(const BatchDesc& d, int maxn, hipStream_t s) {
    using T = double; constexpr int MODE = MODE_PROD;
    PFH_CFG(32, 2, 0) PFH_CFG(64, 4, 0)
    return (int)hipErrorInvalidValue;
}
template <> int launch_typed<float, MODE_PROD>(const BatchDesc& d, int maxn, hipStream_t s) {
    using T = float; constexpr int MODE = MODE_PROD;
    PFH_CFG(32, 2, 0) PFH_CFG(64, 4, 0)
    return (int)hipErrorInvalidValue;
}
template <> int launch_typed<c128, MODE_PROD>(const BatchDesc& d, int maxn, hipStream_t s) {
    using T = c128; constexpr int MODE = MODE_PROD;
    PFH_CFG(32, 2, 0) PFH_CFG(64, 4, 0)
    return (int)hipErrorInvalidValue;
}
template <> int launch_typed<c64, MODE_PROD>(const BatchDesc& d, int maxn, hipStream_t s) {
    using T = c64; constexpr int MODE = MODE_PROD;
    PFH_CFG(32, 2, 0) PFH_CFG(64, 4, 0)
    return (int)hipErrorInvalidValue;
}
#undef PFH_CFG

int launch_reg(int dtype, int mode, const BatchDesc& d, int maxn, hipStream_t s) {
    if (mode == MODE_PROD) {
        switch (dtype) {
            case DT_F32: return launch_typed<float, MODE_PROD>(d, maxn, s);
            case DT_F64: return launch_typed<double, MODE_PROD>(d, maxn, s);
            case DT_C64: return launch_typed<c64, MODE_PROD>(d, maxn, s);
            case DT_C128: return launch_typed<c128, MODE_PROD>(d, maxn, s);
        }
    } else {
        switch (dtype) {
            case DT_F32: return launch_typed<float, MODE_SLOG>(d, maxn, s);
            case DT_F64: return launch_typed<double, MODE_SLOG>(d, maxn, s);
            case DT_C64: return launch_typed<c64, MODE_SLOG>(d, maxn, s);
            case DT_C128: return launch_typed<c128, MODE_SLOG>(d, maxn, s);
        }
    }
    return (int)hipErrorInvalidValue;
}

}  // namespace pfh
