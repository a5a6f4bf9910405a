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
//   physical row phi >= LR lives in register (phi-LR)/NRB of block
//   (phi-LR) % NRB (interleaved, so the register index of a row is
//   wave-uniform); rows phi < LR live in LDS, owned by block phi % NRB.
// Pivoting is virtual: nothing moves.  lab[phi] is the logical position of
// physical index phi (rows and columns alike, the permutation being
// symmetric) and pos[] its inverse, as pf_sm_calc keeps P[] (:11-12, :48-55).
// Registers are only ever written with compile-time indices; the two rows a
// step needs are read with wave-uniform indices (scalar branches).
//
// Per pair step k (right-looking, as pf_sm_calc :29-91):
//   1. f = pos[k]; the owners of row f write it to X[c]          (barrier B1)
//   2. every wave reduces argmax |X[c]| over columns with lab[c] in (k, n),
//      ties to the lowest logical index -> q, pivot a = X[q]
//   3. the owners of row q write (X[c], Y[c] = A[q][c]) to XY[c]  (barrier B2);
//      thread 0 then swaps logical positions k+1 <-> lab[q]
//   4. rank-2 Schur update of every row phi, every column c:
//        A[phi][c] += (Y[phi] X[c] - X[phi] Y[c]) / a
//      (the update of pfaffian_sm.cu:71-89, written for the skew Schur
//      complement S = A22 + (y x^T - x y^T)/a).
// Rows and columns already eliminated take values nobody reads again, so the
// update has no masks or branches.  Pf = (-1)^swaps * prod(pivots); the exact
// zero-pivot test gives Pf = 0 (log -inf, phase 0).
//
// MODE_GRAD turns the same pass into a block sweep (Gauss-Jordan with the
// 2x2 Parlett-Reid pivots).  Sweeping S = {f, q}, B = [[0, a], [-a, 0]],
// maps A[S][j] -> B^-1 A[S][j], A[i][S] -> A[i][S] B^-1, A[S][S] -> -B^-1 and
// A[i][j] -> A[i][j] - A[i][S] B^-1 A[S][j]; sweeping every pair leaves -A^-1.
// Swept rows are kept negated, which keeps the whole working matrix skew, so
// the one update above serves every row once two entries are changed:
// XY[f] = (-1, -a), XY[q] = (a, -1) (rows f, q become y/a and -x/a) and
// lanes f, q use (X, Y) coefficients (-1/a, -1) and (1, -1/a) (columns f, q
// become -y/a and x/a).  At the end the matrix is A^-1 in the caller's
// order, and the gradient of log|Pf| (the reference's JVP,
// pfcuda/cuda_api.py:93-107, transposed) is 1/2 A^-T = -1/2 A^-1.
#include <cstdlib>
#include <cstring>

#include "../../include/pfaffian_hip.h"
#include "pf_batch.h"
#include "pf_regstore.h"
#include "pf_wave.h"

namespace pfh {

template <typename T>
__device__ __forceinline__ T mul_add_update(T a, T Yr, T xs, T Xr, T ys) {
    return a + Yr * xs - Xr * ys;
}

template <typename T, int NCOL, int NRB, int LR, int MODE>
__global__ void __launch_bounds__(NCOL * NRB)
pf_reg_kernel(BatchDesc d) {
    static_assert((NCOL - LR) % NRB == 0, "register rows must tile the column count exactly");
    static_assert(NCOL <= 1024, "the pivot key packs the physical column in ten bits");
    constexpr int RB = (NCOL - LR) / NRB;
    using R = typename real_of<T>::type;
    struct XY { T x, y; };
    __shared__ T Xs[NCOL];       // pivot row f, by physical column
    __shared__ XY XYs[NCOL];     // rows f and q, by physical index
    __shared__ int lab[NCOL];    // physical -> logical
    __shared__ int pos[NCOL];    // logical -> physical
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
        const int64_t off = d.offsets ? d.offsets[idx] : (int64_t)idx * n * n;
        const T* base = (const T*)d.data + off;
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
            if constexpr (MODE == MODE_GRAD) {   // singular: no gradient
                T* G = (T*)d.grad_out + off;
                for (int e = t; e < n * n; e += NCOL * NRB) G[e] = nan_of<T>();
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
        if (rb == 0) { lab[c] = c; pos[c] = c; }
        __syncthreads();

        T prod = one<T>();
        LogAcc la;
        Phase<T> ph;
        bool dead = false;
        int f = 0;   // physical index of logical row k

        for (int k = 0; k < n; k += 2) {
            // 1. row f -> X
            if (f < LR) {
                if (rb == f % NRB) Xs[c] = Ls[f][c];
            } else {
                if (rb == (f - LR) % NRB) Xs[c] = a.dget((f - LR) / NRB);
            }
            __syncthreads();   // B1
            // logical positions k+1, k+2 (written only after B2): read now, so
            // that their latency overlaps the search
            const int p1 = pos[k + 1];
            const int p2 = (k + 2 < n) ? pos[k + 2] : 0;

            // 2. pivot search over logical columns k+1 .. n-1: every wave
            //    reduces all columns with DPP (the largest magnitude, then
            //    the smallest key among the lanes holding it); the key is
            //    logical index * 1024 + physical column, so ties go to the
            //    lowest logical index and the column needs no lookup
            double best = -1.0;
            int bk = 0x7fffffff;
            for (int cc = lane; cc < NCOL; cc += 64) {
                const int l = lab[cc];
                if (l > k && l < n) {
                    const double m = (double)mag2(Xs[cc]);
                    const int key = l * 1024 + cc;
                    if (m > best || (m == best && key < bk)) { best = m; bk = key; }
                }
            }
            double wb;
            int wk;
            wave_argmax(best, bk, wb, wk);                                     // inputs >= -1, no NaN
            int lq, q;
            if (wk == 0x7fffffff) {                // only NaN candidates: keep k+1
                lq = k + 1;
                q = __builtin_amdgcn_readfirstlane(p1);
            } else {
                lq = wk >> 10;
                q = wk & 1023;
            }
            const T piv = Xs[q];
            if (is_zero(piv)) { dead = true; break; }
            const bool sw = (lq != k + 1);
            const int f_next = (k + 2 < n) ? ((lq == k + 2) ? p1 : p2) : 0;

            // 3. row q -> (X, Y)
            bool ownq;
            T y = zero<T>();
            if (q < LR) { ownq = (rb == q % NRB); if (ownq) y = Ls[q][c]; }
            else { ownq = (rb == (q - LR) % NRB); if (ownq) y = a.dget((q - LR) / NRB); }
            if (ownq) {
                XY v{Xs[c], y};
                if constexpr (MODE == MODE_GRAD) {
                    if (c == f) { v.x = -one<T>(); v.y = -piv; }
                    else if (c == q) { v.x = piv; v.y = -one<T>(); }
                }
                XYs[c] = v;
            }
            __syncthreads();   // B2
            if (t == 0 && sw) {
                pos[k + 1] = q; pos[lq] = p1;
                lab[q] = k + 1; lab[p1] = lq;
            }

            if constexpr (MODE == MODE_PROD) { prod = prod * piv; if (sw) prod = -prod; }
            else { la.mul(absd(piv)); ph.mul(piv); if (sw) ph.flip(); }

            // 4. rank-2 update of every row
            const T inv = recip(piv);
            const XY own = XYs[c];
            T xs = own.x * inv, ys = own.y * inv;
            if constexpr (MODE == MODE_GRAD) {
                if (c == f) { xs = -inv; ys = -one<T>(); }
                else if (c == q) { xs = one<T>(); ys = -inv; }
            }
            if (LR > 0) {
                for (int r = rb; r < LR && r < n; r += NRB) {
                    const XY qv = XYs[r];
                    Ls[r][c] = mul_add_update(Ls[r][c], qv.y, xs, qv.x, ys);
                }
            }
            constexpr int G = (RB * (int)sizeof(T) >= 384) ? 2 : 4;   // rows whose LDS reads issue together
#pragma unroll
            for (int g = 0; g < RB; g += G) {
                if (LR + g * NRB < n) {
#pragma unroll
                    for (int i = g; i < g + G && i < RB; ++i) {
                        const XY qv = XYs[LR + i * NRB + rb];
                        a.set(i, mul_add_update(a.get(i), qv.y, xs, qv.x, ys));
                    }
                }
            }
            f = f_next;
        }

        if (t == 0) {
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
        if constexpr (MODE == MODE_GRAD) {
            // physical order is the caller's order: G = -1/2 A^-1 = 1/2 A^-T
            T* G = (T*)d.grad_out + off;
            const R h = (R)(-0.5);
            if (c < n) {
#pragma unroll
                for (int i = 0; i < RB; ++i) {
                    const int r = LR + i * NRB + rb;
                    if (r < n) G[(int64_t)r * n + c] = dead ? nan_of<T>() : a.get(i) * h;
                }
                if (LR > 0) {
                    for (int r = rb; r < LR && r < n; r += NRB)
                        G[(int64_t)r * n + c] = dead ? nan_of<T>() : Ls[r][c] * h;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// launch configuration tables
// ---------------------------------------------------------------------------

static int num_cu() { return device_cu_count(); }

template <typename T, int NCOL, int NRB, int LR, int MODE>
static int launch_one(const BatchDesc& d, hipStream_t stream) {
    if (d.count <= 0) return 0;
    auto kern = pf_reg_kernel<T, NCOL, NRB, LR, MODE>;
    static DevCache resident;   // resident workgroups on the device
    int64_t grid = resident.get([&] {
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, NCOL * NRB, 0) != hipSuccess || occ <= 0) occ = 1;
        return occ * num_cu();
    });
    if (grid > d.count) grid = d.count;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(NCOL * NRB), 0, stream, d);
    return (int)hipGetLastError();
}

int reg_max_n(int dtype) {
    switch (dtype) {
        case DT_F64: return 256;
        case DT_F32: return 256;
        case DT_C64: return 192;
        case DT_C128: return 256;   // slog above 128 only through the batched blocked kernel (d.bslab)
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
// Gradient (block sweep): register rows only (no LDS rows).
template <> int launch_typed<double, MODE_GRAD>(const BatchDesc& d, int maxn, hipStream_t s) {
    using T = double; constexpr int MODE = MODE_GRAD;
    PFH_CFG(32, 2, 0) PFH_CFG(64, 4, 0) PFH_CFG(128, 4, 0) PFH_CFG(192, 4, 0)
    return (int)hipErrorInvalidValue;
}
template <> int launch_typed<c128, MODE_GRAD>(const BatchDesc& d, int maxn, hipStream_t s) {
    using T = c128; constexpr int MODE = MODE_GRAD;
    PFH_CFG(32, 2, 0) PFH_CFG(64, 4, 0) PFH_CFG(128, 8, 0)
    return (int)hipErrorInvalidValue;
}
template <> int launch_typed<float, MODE_GRAD>(const BatchDesc& d, int maxn, hipStream_t s) {
    using T = float; constexpr int MODE = MODE_GRAD;
    PFH_CFG(32, 2, 0) PFH_CFG(64, 4, 0) PFH_CFG(128, 4, 0)
    return (int)hipErrorInvalidValue;
}
template <> int launch_typed<c64, MODE_GRAD>(const BatchDesc& d, int maxn, hipStream_t s) {
    using T = c64; constexpr int MODE = MODE_GRAD;
    PFH_CFG(32, 2, 0) PFH_CFG(64, 4, 0) PFH_CFG(128, 4, 0)
    return (int)hipErrorInvalidValue;
}
#undef PFH_CFG

int reg_grad_max_n(int dtype) {
    switch (dtype) {
        case DT_F64: return 192;
        case DT_F32: return 128;
        case DT_C64: return 128;
        case DT_C128: return 128;
    }
    return 0;
}

// f64 kernel choice for slog and the product form: "lanes" (column per lane,
// above), "lds" (matrix in LDS, MFMA tile updates, pf_lds.hip; above
// n = 192 only when the caller supplied spill slabs) or "bat" (batched
// blocked, matrix in global slabs, several matrices per CU, pf_bat.hip; only
// when the caller supplied its slabs).  Default, from per-bucket timings of
// the 16384-matrix ragged batch on MI355X (ms, lds | bat: 64-128 0.95 | 0.61,
// 128-192 3.22 | 1.46, 192-256 8.32 | 3.31): slog takes bat for n > 64 when
// its slabs are there, else lds, and lanes up to 64; the product form keeps
// lanes, whose rank-2 updates round as the reference's (rel 1e-12 on
// 32 x 32).  PFH_F64_KERNEL=lanes|lds|bat forces one where it applies.
static int f64_kernel_override() {
    static const int v = [] {
        const char* e = getenv("PFH_F64_KERNEL");
        return !e ? -1 : strcmp(e, "lanes") == 0 ? 0 : strcmp(e, "lds") == 0 ? 2 : strcmp(e, "bat") == 0 ? 3 : -1;
    }();
    return v;
}

// 0 lanes, 2 lds, 3 bat (batched blocked, pf_bat.hip; needs d.bslab)
static int f64_kernel(int mode, int maxn, const BatchDesc& d) {
    if (mode == MODE_GRAD) return 0;
    if (mode == MODE_PROD && maxn > 64) return 0;
    if (maxn > 256) return 0;
    const int key = maxn <= 64 ? -1 : maxn <= 128 ? PFH_TUNE_F64_BUCKET_128 : maxn <= 192 ? PFH_TUNE_F64_BUCKET_192
                                                                            : PFH_TUNE_F64_BUCKET_256;
    const int tv = (mode == MODE_SLOG && key >= 0) ? tune_value(key) : -1;
    const int o = tv >= 0 ? tv : f64_kernel_override();
    const bool bat_ok = mode == MODE_SLOG && maxn > 64 && d.bslab;
    if (o == 3) return bat_ok ? 3 : 0;
    if (o < 0 && bat_ok) return 3;
    if (maxn > lds_max_n() && !d.slab) return 0;
    if (o >= 0) return o;
    if (mode == MODE_PROD) return 0;
    return maxn > 64 ? 2 : 0;
}

// n <= 32, product form and slog: the one-wave live-row kernel
// (pf_c32.hip) unless PFH_TUNE_SMALL_KERNEL (or PFH_SMALL_KERNEL) is 0,
// which keeps the column-per-lane kernel above.
static bool c32_enabled() {
    const int tv = tune_value(PFH_TUNE_SMALL_KERNEL);
    if (tv >= 0) return tv == 1;
    static const int env = [] {
        const char* e = getenv("PFH_SMALL_KERNEL");
        return (e && e[0] == '0') ? 0 : 1;
    }();
    return env == 1;
}

int launch_reg(int dtype, int mode, const BatchDesc& d, int maxn, hipStream_t s) {
    if (maxn <= 32 && (mode == MODE_PROD || mode == MODE_SLOG) && c32_enabled() &&
        !(dtype == DT_F64 && f64_kernel_override() == 2))
        return launch_c32(dtype, mode, d, s);
    if (dtype == DT_F64) {
        const int kf = f64_kernel(mode, maxn, d);
        if (kf == 2) return launch_lds_f64(mode, d, maxn, s);
        if (kf == 3) return launch_bat(DT_F64, d, maxn, s);
    }
    if (dtype == DT_C128 && mode == MODE_SLOG && maxn > 64) {
        // complex128 slog: the batched blocked kernel above 64 (the lanes
        // kernel holds at most 128; PFH_C128_KERNEL=lanes keeps it there)
        static const int lanes_env = [] {
            const char* e = getenv("PFH_C128_KERNEL");
            return (e && strcmp(e, "lanes") == 0) ? 1 : 0;
        }();
        const int tv = maxn <= 128 ? tune_value(PFH_TUNE_C128_BUCKET_128) : -1;
        const bool lanes = tv >= 0 ? tv == 0 : lanes_env == 1;
        if (d.bslab && !(lanes && maxn <= 128)) return launch_bat(DT_C128, d, maxn, s);
        if (maxn > 128) return (int)hipErrorInvalidValue;
    }
    if (mode == MODE_GRAD) {
        switch (dtype) {
            case DT_F32: return launch_typed<float, MODE_GRAD>(d, maxn, s);
            case DT_F64: return launch_typed<double, MODE_GRAD>(d, maxn, s);
            case DT_C64: return launch_typed<c64, MODE_GRAD>(d, maxn, s);
            case DT_C128: return launch_typed<c128, MODE_GRAD>(d, maxn, s);
        }
        return (int)hipErrorInvalidValue;
    }
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
