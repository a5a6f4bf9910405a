// Unblocked all-FP64 ("plain") Parlett-Reid for matrices too large for the
// register-resident workgroup kernels, with the pivot kept on the device.
//
// Replaces the reference's host-driven left-looking driver
// slog_pfaffian_lg (src/slog_pfaffian_lg.cu:151-211), which reads every
// pivot back to the host with a blocking copy (:178) and launches ~7 small
// kernels per pair step.  Here each pair step is three stream-ordered
// launches with no host round trip (so the whole factorization can be
// captured in a hipGraph):
//   lg_pivot  : argmax_{c>k} |W[k][c]| over the current row k (ties: lowest c),
//               sign/log/phase accumulation, zero-pivot stop     (1 workgroup)
//   lg_swap   : symmetric swap of index k+1 with the pivot index p over rows
//               and columns k..n-1                               (1 workgroup)
//   lg_update : right-looking rank-2 Schur update of the trailing block
//               W[r][c] += (W[k+1][r] W[k][c] - W[k][r] W[k+1][c]) / pivot
//               for r, c >= k+2                                  (2-D grid)
// W is a full-square row-major working copy, so every access is to rows
// (coalesced) except the column half of the swap.
//
// pf_glb_kernel: the same algorithm inside one workgroup per matrix, over a
// global-memory working slab, for ragged batches whose sizes exceed the
// register kernels.
#include "pf_batch.h"
#include "pf_large.h"
#include "pf_types.h"

namespace pfh {

template <typename T>
__global__ void __launch_bounds__(1024) lg_init(LargeState<T>* st) {
    if (threadIdx.x == 0) {
        st->p = 0; st->dead = 0; st->swaps = 0;
        st->piv = one<T>();
        st->lm = 1.0; st->le = 0;
        st->ph_re = 1.0; st->ph_im = 0.0;
    }
}

// Block-wide argmax of |row[c]| for c in [lo, hi); ties keep the lowest c.
template <typename T, int NT>
__device__ void block_argmax(const T* row, int lo, int hi, double* s_val, int* s_idx, double& best_out, int& idx_out) {
    double best = -1.0;
    int bi = 0x7fffffff;
    for (int c = lo + (int)threadIdx.x; c < hi; c += NT) {
        const double m = (double)mag2(row[c]);
        if (m > best) { best = m; bi = c; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { s_val[w] = best; s_idx[w] = bi; }
    __syncthreads();
    if (threadIdx.x < 64) {
        best = (threadIdx.x < NT / 64) ? s_val[threadIdx.x] : -1.0;
        bi = (threadIdx.x < NT / 64) ? s_idx[threadIdx.x] : 0x7fffffff;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ob = __shfl_xor(best, off);
            const int oi = __shfl_xor(bi, off);
            if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
        }
        if (threadIdx.x == 0) { s_val[0] = best; s_idx[0] = bi; }
    }
    __syncthreads();
    best_out = s_val[0];
    idx_out = s_idx[0];
}

template <typename T>
__device__ void accumulate_pivot(LargeState<T>* st, T piv, bool sw) {
    // log|prod| as mantissa * 2^exponent; phase renormalised each step.
    double m = st->lm * absd(piv);
    int ex;
    m = frexp(m, &ex);
    st->lm = m;
    st->le += ex;
    if constexpr (is_cplx<T>::value) {
        const double a = hypot((double)piv.re, (double)piv.im);
        const double pr = (double)piv.re / a, pi = (double)piv.im / a;
        double nr = st->ph_re * pr - st->ph_im * pi, ni = st->ph_re * pi + st->ph_im * pr;
        const double nm = hypot(nr, ni);
        nr /= nm; ni /= nm;
        if (sw) { nr = -nr; ni = -ni; }
        st->ph_re = nr; st->ph_im = ni;
    } else {
        double s = st->ph_re;
        if (piv < T(0)) s = -s;
        if (sw) s = -s;
        st->ph_re = s;
    }
}

template <typename T>
__global__ void __launch_bounds__(1024) lg_pivot(const T* __restrict__ W, int64_t n, int k, LargeState<T>* st) {
    __shared__ double s_val[16];
    __shared__ int s_idx[16];
    if (st->dead) return;
    const T* row = W + (int64_t)k * n;
    double best; int p;
    block_argmax<T, 1024>(row, k + 1, (int)n, s_val, s_idx, best, p);
    if (threadIdx.x == 0) {
        if (p >= n) p = k + 1;
        const T piv = row[p];
        st->p = p;
        st->piv = piv;
        if (is_zero(piv)) { st->dead = 1; return; }
        const bool sw = (p != k + 1);
        if (sw) st->swaps += 1;
        accumulate_pivot(st, piv, sw);
    }
}

template <typename T>
__global__ void __launch_bounds__(1024) lg_swap(T* __restrict__ W, int64_t n, int k, const LargeState<T>* st) {
    if (st->dead) return;
    const int p = st->p, i = k + 1;
    if (p == i) return;
    // rows i <-> p over columns k..n-1
    for (int64_t c = k + threadIdx.x; c < n; c += blockDim.x) {
        T a = W[i * n + c], b = W[(int64_t)p * n + c];
        W[i * n + c] = b; W[(int64_t)p * n + c] = a;
    }
    __syncthreads();
    // columns i <-> p over rows k..n-1
    for (int64_t r = k + threadIdx.x; r < n; r += blockDim.x) {
        T a = W[r * n + i], b = W[r * n + p];
        W[r * n + i] = b; W[r * n + p] = a;
    }
}

// Tile: 64 columns x (4 row-threads x RPT rows).
template <typename T, int RPT>
__global__ void __launch_bounds__(256) lg_update(T* __restrict__ W, int64_t n, int k, const LargeState<T>* st) {
    if (st->dead) return;
    const int c = k + 2 + blockIdx.x * 64 + (threadIdx.x & 63);
    const int r0 = k + 2 + blockIdx.y * (4 * RPT) + (threadIdx.x >> 6);
    if (c >= n) return;
    const T inv = recip(st->piv);
    const T* X = W + (int64_t)k * n;
    const T* Y = W + (int64_t)(k + 1) * n;
    const T xs = X[c] * inv, ys = Y[c] * inv;
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
        const int r = r0 + 4 * i;
        if (r < n) {
            T* e = W + (int64_t)r * n + c;
            *e = *e + Y[r] * xs - X[r] * ys;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(64) lg_finish(const LargeState<T>* st, double* log_out, T* phase_out) {
    if (threadIdx.x != 0) return;
    if (st->dead) {
        *log_out = -INFINITY;
        *phase_out = zero<T>();
        return;
    }
    *log_out = log(st->lm) + (double)st->le * 0.69314718055994530942;
    T ph;
    if constexpr (is_cplx<T>::value) { ph.re = st->ph_re; ph.im = st->ph_im; }
    else ph = (T)st->ph_re;
    *phase_out = ph;
}

// ---------------------------------------------------------------------------
// Gradient of log|Pf| for matrices past the register kernels: the same
// block sweep as the register kernel's MODE_GRAD (swept rows stored negated,
// so the working matrix stays skew and one rank-2 update serves every row),
// over a full-square working copy with physical interchanges of whole rows
// and columns.  Per pair step: lg_pivot, gs_swap, gs_update (every row and
// column outside the pair), gs_rows (the pair's two rows).  At the end the
// working matrix is A^-1 in pivot order; gs_out writes -1/2 A^-1 = 1/2 A^-T
// back in the caller's order.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(1024) gs_swap(T* __restrict__ W, int* __restrict__ perm, int64_t n, int k,
                                                const LargeState<T>* st) {
    if (st->dead) return;
    const int p = st->p, i = k + 1;
    if (p == i) return;
    for (int64_t c = threadIdx.x; c < n; c += blockDim.x) {
        T a = W[i * n + c], b = W[(int64_t)p * n + c];
        W[i * n + c] = b; W[(int64_t)p * n + c] = a;
    }
    __syncthreads();
    for (int64_t r = threadIdx.x; r < n; r += blockDim.x) {
        T a = W[r * n + i], b = W[r * n + p];
        W[r * n + i] = b; W[r * n + p] = a;
    }
    if (threadIdx.x == 0) { int t = perm[i]; perm[i] = perm[p]; perm[p] = t; }
}

// rows r, columns c outside the pair {k, k+1}; columns k, k+1 get -y/a, x/a
template <typename T, int RPT>
__global__ void __launch_bounds__(256) gs_update(T* __restrict__ W, int64_t n, int k, const LargeState<T>* st) {
    if (st->dead) return;
    const int c = blockIdx.x * 64 + (threadIdx.x & 63);
    const int r0 = blockIdx.y * (4 * RPT) + (threadIdx.x >> 6);
    if (c >= n) return;
    const T inv = recip(st->piv);
    const T* X = W + (int64_t)k * n;
    const T* Y = W + (int64_t)(k + 1) * n;
    T xs, ys;
    if (c == k) { xs = -inv; ys = zero<T>(); }
    else if (c == k + 1) { xs = zero<T>(); ys = -inv; }
    else { xs = X[c] * inv; ys = Y[c] * inv; }
    const bool special = (c == k || c == k + 1);
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
        const int r = r0 + 4 * i;
        if (r < n && r != k && r != k + 1) {
            T* e = W + (int64_t)r * n + c;
            const T base = special ? zero<T>() : *e;
            *e = base + Y[r] * xs - X[r] * ys;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(256) gs_rows(T* __restrict__ W, int64_t n, int k, const LargeState<T>* st) {
    if (st->dead) return;
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= n) return;
    const T inv = recip(st->piv);
    T* X = W + (int64_t)k * n;
    T* Y = W + (int64_t)(k + 1) * n;
    const T x = X[c], y = Y[c];
    if (c == k) { X[c] = zero<T>(); Y[c] = inv; }
    else if (c == k + 1) { X[c] = -inv; Y[c] = zero<T>(); }
    else { X[c] = y * inv; Y[c] = -(x * inv); }
}

template <typename T>
__global__ void __launch_bounds__(256) gs_out(const T* __restrict__ W, const int* __restrict__ perm, int64_t n,
                                               const LargeState<T>* st, T* __restrict__ G) {
    using R = typename real_of<T>::type;
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= n * n) return;
    const int64_t r = e / n, c = e % n;
    const T v = st->dead ? nan_of<T>() : W[e] * (R)(-0.5);
    G[(int64_t)perm[r] * n + perm[c]] = v;
}

template <typename T>
__global__ void gs_init_perm(int* perm, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        perm[i] = (int)i;
}

template <typename T>
int large_grad(const T* A, int64_t n, T* W, int* perm, LargeState<T>* st, double* log_out, T* phase_out, T* G,
               hipStream_t s) {
    hipError_t e = hipMemcpyAsync(W, A, sizeof(T) * n * n, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(lg_init<T>, dim3(1), dim3(64), 0, s, st);
    hipLaunchKernelGGL(gs_init_perm<T>, dim3(64), dim3(256), 0, s, perm, n);
    constexpr int RPT = 8;
    const dim3 grid((unsigned)((n + 63) / 64), (unsigned)((n + 4 * RPT - 1) / (4 * RPT)));
    for (int k = 0; k + 1 < n; k += 2) {
        hipLaunchKernelGGL(lg_pivot<T>, dim3(1), dim3(1024), 0, s, W, n, k, st);
        hipLaunchKernelGGL(gs_swap<T>, dim3(1), dim3(1024), 0, s, W, perm, n, k, st);
        hipLaunchKernelGGL((gs_update<T, RPT>), grid, dim3(256), 0, s, W, n, k, st);
        hipLaunchKernelGGL(gs_rows<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, W, n, k, st);
    }
    hipLaunchKernelGGL(lg_finish<T>, dim3(1), dim3(64), 0, s, st, log_out, phase_out);
    hipLaunchKernelGGL(gs_out<T>, dim3((unsigned)((n * n + 255) / 256)), dim3(256), 0, s, W, perm, n, st, G);
    return (int)hipGetLastError();
}

template <typename T>
int large_slog(const T* A, int64_t n, T* W, LargeState<T>* st, double* log_out, T* phase_out, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(W, A, sizeof(T) * n * n, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(lg_init<T>, dim3(1), dim3(64), 0, s, st);
    constexpr int RPT = 8;
    for (int k = 0; k + 1 < n; k += 2) {
        hipLaunchKernelGGL(lg_pivot<T>, dim3(1), dim3(1024), 0, s, W, n, k, st);
        hipLaunchKernelGGL(lg_swap<T>, dim3(1), dim3(1024), 0, s, W, n, k, st);
        const int64_t m = n - (k + 2);
        if (m > 0) {
            dim3 grid((unsigned)((m + 63) / 64), (unsigned)((m + 4 * RPT - 1) / (4 * RPT)));
            hipLaunchKernelGGL((lg_update<T, RPT>), grid, dim3(256), 0, s, W, n, k, st);
        }
    }
    hipLaunchKernelGGL(lg_finish<T>, dim3(1), dim3(64), 0, s, st, log_out, phase_out);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// one workgroup per matrix over a global slab (ragged sizes past the
// register kernels)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(1024) pf_glb_kernel(BatchDesc d, T* __restrict__ slabs, int64_t slab_elems) {
    __shared__ double s_val[16];
    __shared__ int s_idx[16];
    using R = typename real_of<T>::type;
    int64_t count = d.count;
    if (d.count_dev) { int64_t cd = *d.count_dev; count = cd < count ? cd : count; }
    T* W = slabs + (int64_t)blockIdx.x * slab_elems;
    for (int64_t j = blockIdx.x; j < count; j += gridDim.x) {
        const int idx = d.order ? d.order[(d.begin_dev ? *d.begin_dev : 0) + j] : (int)j;
        const int n = __builtin_amdgcn_readfirstlane(d.sizes ? d.sizes[idx] : d.dense_n);
        const T* A = (const T*)d.data + (d.offsets ? d.offsets[idx] : (int64_t)idx * n * n);
        __syncthreads();
        if (n < 0 || (int64_t)n * n > slab_elems) {   // outside the slab: NaN, no access
            if (threadIdx.x == 0) { d.log_out[idx] = __builtin_nan(""); ((T*)d.phase_out)[idx] = nan_of<T>(); }
            continue;
        }
        if (n == 0 || (n & 1)) {
            if (threadIdx.x == 0) { d.log_out[idx] = -INFINITY; ((T*)d.phase_out)[idx] = zero<T>(); }
            continue;
        }
        for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += blockDim.x) W[e] = A[e];
        __syncthreads();
        LogAcc la;
        Phase<T> ph;
        bool dead = false;
        for (int k = 0; k + 1 < n; k += 2) {
            const T* row = W + (int64_t)k * n;
            double best; int p;
            block_argmax<T, 1024>(row, k + 1, n, s_val, s_idx, best, p);
            if (p >= n) p = k + 1;
            const T piv = row[p];
            if (is_zero(piv)) { dead = true; break; }
            const int i = k + 1;
            const bool sw = (p != i);
            if (sw) {
                for (int c = k + threadIdx.x; c < n; c += blockDim.x) {
                    T a = W[(int64_t)i * n + c], b = W[(int64_t)p * n + c];
                    W[(int64_t)i * n + c] = b; W[(int64_t)p * n + c] = a;
                }
                __syncthreads();
                for (int r = k + threadIdx.x; r < n; r += blockDim.x) {
                    T a = W[(int64_t)r * n + i], b = W[(int64_t)r * n + p];
                    W[(int64_t)r * n + i] = b; W[(int64_t)r * n + p] = a;
                }
                __syncthreads();
            }
            la.mul(absd(piv)); ph.mul(piv); if (sw) ph.flip();
            const T inv = recip(piv);
            const T* X = W + (int64_t)k * n;
            const T* Y = W + (int64_t)(k + 1) * n;
            const int m = n - (k + 2);
            for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
                const int r = k + 2 + e / m, c = k + 2 + e % m;
                T* w = W + (int64_t)r * n + c;
                *w = *w + Y[r] * (X[c] * inv) - X[r] * (Y[c] * inv);
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            d.log_out[idx] = dead ? -INFINITY : la.value();
            T phv;
            if constexpr (is_cplx<T>::value) { phv.re = (R)ph.re; phv.im = (R)ph.im; }
            else phv = (T)ph.s;
            ((T*)d.phase_out)[idx] = dead ? zero<T>() : phv;
        }
    }
}

template <typename T>
int launch_glb(const BatchDesc& d, T* slabs, int64_t slab_elems, int grid, hipStream_t s) {
    if (d.count <= 0) return 0;
    hipLaunchKernelGGL(pf_glb_kernel<T>, dim3(grid), dim3(1024), 0, s, d, slabs, slab_elems);
    return (int)hipGetLastError();
}

#define PFH_INST(T)                                                                          \
    template int large_slog<T>(const T*, int64_t, T*, LargeState<T>*, double*, T*, hipStream_t); \
    template int large_grad<T>(const T*, int64_t, T*, int*, LargeState<T>*, double*, T*, T*, hipStream_t); \
    template int launch_glb<T>(const BatchDesc&, T*, int64_t, int, hipStream_t);
PFH_INST(float)
PFH_INST(double)
PFH_INST(c64)
PFH_INST(c128)
#undef PFH_INST

}  // namespace pfh
