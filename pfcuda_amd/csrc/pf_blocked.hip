// Blocked Parlett-Reid for one large matrix (N in the thousands).
//
// The reference factors a large matrix left-looking from the host
// (src/slog_pfaffian_lg.cu:151-211): every pair step re-reads the whole
// factor history (row_update, :75-112), and every pivot goes to the host and
// back (:178).  This form is right-looking and blocked, with all pivoting on
// the device:
//
//   panel (nb pair steps, left-looking inside the panel)
//     the current row k is the stale row of W plus the corrections of the
//     panel's pairs so far:  row_k[c] = W[k][c] + sum_j (U_j[k] V_j[c] - V_j[k] U_j[c])
//     (U_j = y_j / a_j, V_j = x_j: the skew Schur complement
//      S = A + (y x^T - x y^T)/a of each pair, kept as two vectors);
//     pivot = argmax_{c>k} |row_k[c]| (ties: lowest c);
//     the interchange k+1 <-> p is applied to a logical->physical map `perm`
//     over W, to the panel vectors and to the row buffer -- W itself is not
//     touched during the panel, so no strided column swap sits on the
//     critical path.
//   trailing update (once per panel)
//     W_new[r][c] = W_old[perm r][perm c] + sum_j (U_j[r] V_j[c] - V_j[r] U_j[c])
//     over the upper triangle r < c of the trailing block (skew symmetry
//     gives the rest); W_old and W_new alternate.
//
// One launch per pair step: launch i computes row k+1 of pair i (storing
// U_i, V_i) and, in the same pass, the next pivot row k+2 (its two entries
// from pair i are recomputed redundantly by every block); the argmax is
// finished by the last block to arrive (a device-scope counter), which also
// applies the interchange and accumulates log|pivot| and the sign/phase.
// Only the upper triangle of W is ever read or written.
#include "pf_batch.h"
#include "pf_blocked.h"
#include "pf_types.h"

namespace pfh {

constexpr int BP_THREADS = 128;    // columns per block in the row kernels
constexpr int BP_TILE = 64;        // trailing-update tile edge

template <typename T>
struct BPState {
    int p, dead, counter, pad;
    T piv;
    double lm;
    int le, pad2;
    double ph_re, ph_im;
};

template <typename T>
struct BPArgs {
    T* W;            // W_old: stale matrix (upper triangle, physical indices)
    T* P;            // (2*nb) x n: U_j at [j*n + c], V_j at [(nb + j)*n + c] (pair-major: coalesced in c)
    T* R;            // 2 x n row buffers
    int* perm;       // logical -> physical index into W
    double* part_v;  // per-block argmax partials
    int* part_i;
    BPState<T>* st;
    int64_t n;
    int nb;
};

// Element (r, c) of the stale matrix seen through perm, from the upper triangle.
template <typename T>
__device__ __forceinline__ T wp(const T* W, const int* perm, int64_t n, int r, int c) {
    const int pr = perm[r], pc = perm[c];
    if (pr < pc) return W[(int64_t)pr * n + pc];
    if (pr > pc) return -W[(int64_t)pc * n + pr];
    return zero<T>();
}

template <typename T>
__device__ __forceinline__ void accumulate(BPState<T>* st, T piv, bool sw) {
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
__global__ void bp_init(BPArgs<T> a) {
    const int64_t n = a.n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a.perm[i] = (int)i;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        BPState<T>* st = a.st;
        st->p = 0; st->dead = 0; st->counter = 0;
        st->piv = one<T>();
        st->lm = 1.0; st->le = 0;
        st->ph_re = 1.0; st->ph_im = 0.0;
    }
}

template <typename T>
__global__ void bp_reset_perm(int* perm, int64_t lo, int64_t n) {
    for (int64_t i = lo + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        perm[i] = (int)i;
}

// DO_M: finish pair i (row k+1, store U_i, V_i for c >= k+2).
// DO_L: compute pivot row kl (= k+2 with DO_M, else k) and choose its pivot.
// Grid: blocks of BP_THREADS columns covering c in [c_lo, n).
template <typename T, bool DO_M, bool DO_L>
__global__ void __launch_bounds__(BP_THREADS) bp_row(BPArgs<T> a, int k, int i, int c_lo) {
    __shared__ double s_v[BP_THREADS / 64];
    __shared__ int s_i[BP_THREADS / 64];
    __shared__ int s_last;
    BPState<T>* st = a.st;
    if (st->dead) return;
    const int64_t n = a.n;
    const int nb = a.nb;
    const int c = c_lo + blockIdx.x * BP_THREADS + threadIdx.x;
    const T* PU = a.P;                    // U_j[r] at PU[j*n + r]
    const T* PV = a.P + (int64_t)nb * n;  // V_j[r] at PV[j*n + r]
    const bool valid = c < n;
    const T* Rk = a.R + (int64_t)((k >> 1) & 1) * n;          // row k (current, post-swap)
    T* Rl = a.R + (int64_t)(((k >> 1) + (DO_M ? 1 : 0)) & 1) * n;  // row kl output

    // One pass over the panel's completed pairs j < i, eight at a time
    // (nb is a multiple of 8, so the batch never leaves P; lanes j >= i are
    // masked to zero):
    //   r1 = sum U_j[k+1] V_j[c]  - V_j[k+1] U_j[c]    (row k+1 at column c)
    //   rk = sum U_j[k+1] V_j[kl] - V_j[k+1] U_j[kl]   (row k+1 at column kl)
    //   rl = sum U_j[kl]  V_j[c]  - V_j[kl]  U_j[c]    (row kl at column c)
    // The [k+1] and [kl] operands are the same for every thread.
    const int kl = DO_M ? k + 2 : k;
    const bool live = valid && c > kl;        // column of the new pivot row
    const int cc = valid ? c : kl;            // a safe column for masked threads
    T r1 = zero<T>(), rk = zero<T>(), rl = zero<T>();
    for (int j0 = 0; j0 < i; j0 += 8) {
        T uc8[8], vc8[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            uc8[t] = PU[(int64_t)(j0 + t) * n + cc];
            vc8[t] = PV[(int64_t)(j0 + t) * n + cc];
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const bool on = j0 + t < i;
            const T z = zero<T>();
            const T uc_ = on ? uc8[t] : z, vc_ = on ? vc8[t] : z;
            const T ul = on ? PU[(int64_t)(j0 + t) * n + kl] : z;
            const T vl = on ? PV[(int64_t)(j0 + t) * n + kl] : z;
            rl = rl + ul * vc_ - vl * uc_;
            if (DO_M) {
                const T u1 = on ? PU[(int64_t)(j0 + t) * n + k + 1] : z;
                const T v1 = on ? PV[(int64_t)(j0 + t) * n + k + 1] : z;
                r1 = r1 + u1 * vc_ - v1 * uc_;
                rk = rk + u1 * vl - v1 * ul;
            }
        }
    }

    T uc = zero<T>(), vc = zero<T>();
    T u_kl = zero<T>(), v_kl = zero<T>();
    if (DO_M) {
        const T inv = recip(st->piv);
        if (valid && c >= k + 2) {
            uc = (wp(a.W, a.perm, n, k + 1, c) + r1) * inv;
            vc = Rk[c];
            a.P[(int64_t)i * n + c] = uc;
            a.P[(int64_t)(nb + i) * n + c] = vc;
        }
        if (DO_L && kl < n) {
            // U_i[kl], V_i[kl], recomputed by every block
            u_kl = (wp(a.W, a.perm, n, k + 1, kl) + rk) * inv;
            v_kl = Rk[kl];
        }
    }
    if (!DO_L) return;

    // row kl at column c > kl
    double best = -1.0;
    int bi = 0x7fffffff;
    if (live) {
        T r = wp(a.W, a.perm, n, kl, c) + rl;
        if (DO_M) r = r + u_kl * vc - v_kl * uc;
        Rl[c] = r;
        best = (double)mag2(r);
        bi = c;
    }
    // block argmax
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    if ((threadIdx.x & 63) == 0) { s_v[threadIdx.x >> 6] = best; s_i[threadIdx.x >> 6] = bi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BP_THREADS / 64; ++w) {
            if (s_v[w] > best || (s_v[w] == best && s_i[w] < bi)) { best = s_v[w]; bi = s_i[w]; }
        }
        a.part_v[blockIdx.x] = best;
        a.part_i[blockIdx.x] = bi;
        __threadfence();
        const int prev = atomicAdd(&st->counter, 1);
        s_last = (prev == (int)gridDim.x - 1);
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();

    // last block: final argmax over the partials (all threads load in
    // parallel), then the interchange
    __shared__ int s_p;
    __shared__ int s_sw;
    {
        double b = -1.0;
        int p = 0x7fffffff;
        for (unsigned g = threadIdx.x; g < gridDim.x; g += BP_THREADS) {
            const double v = __hip_atomic_load(&a.part_v[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int ix = __hip_atomic_load(&a.part_i[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (v > b || (v == b && ix < p)) { b = v; p = ix; }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ob = __shfl_xor(b, off);
            const int oi = __shfl_xor(p, off);
            if (ob > b || (ob == b && oi < p)) { b = ob; p = oi; }
        }
        __syncthreads();   // s_v / s_i free again
        if ((threadIdx.x & 63) == 0) { s_v[threadIdx.x >> 6] = b; s_i[threadIdx.x >> 6] = p; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double b = s_v[0];
        int p = s_i[0];
        for (int w = 1; w < BP_THREADS / 64; ++w) {
            if (s_v[w] > b || (s_v[w] == b && s_i[w] < p)) { b = s_v[w]; p = s_i[w]; }
        }
        if (p >= n) p = kl + 1;   // only NaN candidates
        const T piv = Rl[p];
        st->counter = 0;
        if (is_zero(piv)) {
            st->dead = 1;
            s_p = -1;
        } else {
            const bool sw = (p != kl + 1);
            st->p = p;
            st->piv = piv;
            accumulate(st, piv, sw);
            s_p = p;
            s_sw = sw ? 1 : 0;
            if (sw) {
                const int t0 = a.perm[kl + 1];
                a.perm[kl + 1] = a.perm[p];
                a.perm[p] = t0;
                const T t1 = Rl[kl + 1];
                Rl[kl + 1] = Rl[p];
                Rl[p] = t1;
            }
        }
    }
    __syncthreads();
    if (s_p >= 0 && s_sw) {
        // panel vectors of the pairs so far follow the interchange
        const int p = s_p;
        const int cnt = DO_M ? i + 1 : i;
        for (int j = threadIdx.x; j < cnt; j += BP_THREADS) {
            T* U = a.P + (int64_t)j * n;
            T* V = a.P + (int64_t)(nb + j) * n;
            T x = U[kl + 1]; U[kl + 1] = U[p]; U[p] = x;
            T y = V[kl + 1]; V[kl + 1] = V[p]; V[p] = y;
        }
    }
}

// Trailing update over upper tiles of the block [lo, n)^2:
// Wn[r][c] = Wo[perm r][perm c] + sum_{j<np} (U_j[r] V_j[c] - V_j[r] U_j[c]),  r < c.
// 256 threads; thread (tx, ty) owns rows r0+ty+16a, cols c0+tx+16b (a, b < 4).
template <typename T>
__global__ void __launch_bounds__(256) bp_trailing(const T* __restrict__ Wo, T* __restrict__ Wn,
                                                    const T* __restrict__ P, const int* __restrict__ perm,
                                                    int64_t n, int lo, int nb, int np, int ntile) {
    extern __shared__ unsigned char smem_raw[];
    T* sm = reinterpret_cast<T*>(smem_raw);
    const int ld = np + 1;                       // padded row length
    T* Ur = sm;
    T* Vr = Ur + BP_TILE * ld;
    T* Uc = Vr + BP_TILE * ld;
    T* Vc = Uc + BP_TILE * ld;
    // upper-triangular tile index -> (ti, tj), tj >= ti
    int t = blockIdx.x, ti = 0;
    while (t >= ntile - ti) { t -= ntile - ti; ++ti; }
    const int tj = ti + t;
    const int r0 = lo + ti * BP_TILE, c0 = lo + tj * BP_TILE;
    for (int e = threadIdx.x; e < BP_TILE * np; e += 256) {
        const int q = e % BP_TILE, j = e / BP_TILE;   // consecutive threads: consecutive rows (coalesced)
        const int r = r0 + q, c = c0 + q;
        Ur[q * ld + j] = (r < n) ? P[(int64_t)j * n + r] : zero<T>();
        Vr[q * ld + j] = (r < n) ? P[(int64_t)(nb + j) * n + r] : zero<T>();
        Uc[q * ld + j] = (c < n) ? P[(int64_t)j * n + c] : zero<T>();
        Vc[q * ld + j] = (c < n) ? P[(int64_t)(nb + j) * n + c] : zero<T>();
    }
    __syncthreads();
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    T acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = zero<T>();
    for (int j = 0; j < np; ++j) {
        T ur[4], vr[4], uc[4], vc[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) { ur[x] = Ur[(ty + 16 * x) * ld + j]; vr[x] = Vr[(ty + 16 * x) * ld + j]; }
#pragma unroll
        for (int y = 0; y < 4; ++y) { uc[y] = Uc[(tx + 16 * y) * ld + j]; vc[y] = Vc[(tx + 16 * y) * ld + j]; }
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) acc[x][y] = acc[x][y] + ur[x] * vc[y] - vr[x] * uc[y];
    }
#pragma unroll
    for (int x = 0; x < 4; ++x) {
        const int r = r0 + ty + 16 * x;
        if (r >= n) continue;
        const int pr = perm[r];
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const int c = c0 + tx + 16 * y;
            if (c >= n || c <= r) continue;
            const int pc = perm[c];
            T w = (pr < pc) ? Wo[(int64_t)pr * n + pc] : -Wo[(int64_t)pc * n + pr];
            Wn[(int64_t)r * n + c] = w + acc[x][y];
        }
    }
}

template <typename T>
__global__ void bp_finish(const BPState<T>* st, double* log_out, T* phase_out) {
    if (threadIdx.x != 0) return;
    if (st->dead) { *log_out = -INFINITY; *phase_out = zero<T>(); return; }
    *log_out = log(st->lm) + (double)st->le * 0.69314718055994530942;
    T ph;
    if constexpr (is_cplx<T>::value) { ph.re = st->ph_re; ph.im = st->ph_im; }
    else ph = (T)st->ph_re;
    *phase_out = ph;
}

static size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

template <typename T>
struct BPLayout {
    size_t w0, w1, p, r, perm, pv, pi, st, total;
};

template <typename T>
static BPLayout<T> bp_layout(int64_t n, int nb) {
    BPLayout<T> L{};
    const int64_t nblk = (n + BP_THREADS - 1) / BP_THREADS + 1;
    size_t o = 0;
    L.w0 = o; o = al(o + sizeof(T) * (size_t)n * n);
    L.w1 = o; o = al(o + sizeof(T) * (size_t)n * n);
    L.p = o; o = al(o + sizeof(T) * (size_t)n * 2 * nb);
    L.r = o; o = al(o + sizeof(T) * (size_t)n * 2);
    L.perm = o; o = al(o + sizeof(int) * (size_t)n);
    L.pv = o; o = al(o + sizeof(double) * nblk);
    L.pi = o; o = al(o + sizeof(int) * nblk);
    L.st = o; o = al(o + sizeof(BPState<T>));
    L.total = o;
    return L;
}

int blocked_panel_width(int dtype) {
    (void)dtype;
    return 32;
}

size_t blocked_workspace_bytes(int dtype, int64_t n, bool mixed) {
    (void)mixed;
    const int nb = blocked_panel_width(dtype);
    switch (dtype) {
        case DT_F32: return bp_layout<float>(n, nb).total;
        case DT_F64: return bp_layout<double>(n, nb).total;
        case DT_C64: return bp_layout<c64>(n, nb).total;
        case DT_C128: return bp_layout<c128>(n, nb).total;
    }
    return 0;
}

template <typename T>
int blocked_slog(const T* A, int64_t n, int nb, bool mixed, void* ws, double* log_out, T* phase_out,
                 hipStream_t s) {
    (void)mixed;
    if (nb <= 0 || nb % 8 != 0) return (int)hipErrorInvalidValue;   // bp_row reads P in batches of 8 pairs
    const BPLayout<T> L = bp_layout<T>(n, nb);
    char* base = (char*)ws;
    T* W[2] = {(T*)(base + L.w0), (T*)(base + L.w1)};
    BPArgs<T> a{};
    a.P = (T*)(base + L.p);
    a.R = (T*)(base + L.r);
    a.perm = (int*)(base + L.perm);
    a.part_v = (double*)(base + L.pv);
    a.part_i = (int*)(base + L.pi);
    a.st = (BPState<T>*)(base + L.st);
    a.n = n;
    a.nb = nb;
    hipError_t e = hipMemcpyAsync(W[0], A, sizeof(T) * n * n, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return (int)e;
    a.W = W[0];
    hipLaunchKernelGGL(bp_init<T>, dim3(64), dim3(256), 0, s, a);
    int cur = 0;
    auto blocks = [&](int64_t c_lo) { return (unsigned)((n - c_lo + BP_THREADS - 1) / BP_THREADS); };
    for (int64_t k0 = 0; k0 + 1 < n; k0 += 2 * nb) {
        const int np = (int)((n - k0) / 2 < nb ? (n - k0) / 2 : nb);
        a.W = W[cur];
        // first pivot row of the panel
        hipLaunchKernelGGL((bp_row<T, false, true>), dim3(blocks(k0 + 1)), dim3(BP_THREADS), 0, s, a, (int)k0, 0,
                           (int)(k0 + 1));
        for (int i = 0; i < np; ++i) {
            const int64_t k = k0 + 2 * i;
            if (k + 2 >= n) break;   // last pair of the matrix: nothing left to update
            if (i + 1 < np) {
                hipLaunchKernelGGL((bp_row<T, true, true>), dim3(blocks(k + 2)), dim3(BP_THREADS), 0, s, a,
                                   (int)k, i, (int)(k + 2));
            } else {
                hipLaunchKernelGGL((bp_row<T, true, false>), dim3(blocks(k + 2)), dim3(BP_THREADS), 0, s, a,
                                   (int)k, i, (int)(k + 2));
            }
        }
        const int64_t lo = k0 + 2 * np;
        if (lo + 1 < n) {
            const int ntile = (int)((n - lo + BP_TILE - 1) / BP_TILE);
            const int tiles = ntile * (ntile + 1) / 2;
            const size_t smem = sizeof(T) * 4 * BP_TILE * (np + 1);
            if (smem > 65536) {
                e = hipFuncSetAttribute((const void*)bp_trailing<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)smem);
                if (e != hipSuccess) return (int)e;
            }
            hipLaunchKernelGGL(bp_trailing<T>, dim3(tiles), dim3(256), smem, s, W[cur], W[cur ^ 1], a.P, a.perm,
                               n, (int)lo, nb, np, ntile);
            hipLaunchKernelGGL(bp_reset_perm<T>, dim3(64), dim3(256), 0, s, a.perm, lo, n);
            cur ^= 1;
        }
    }
    hipLaunchKernelGGL(bp_finish<T>, dim3(1), dim3(64), 0, s, a.st, log_out, phase_out);
    return (int)hipGetLastError();
}

#define PFH_INST(T) \
    template int blocked_slog<T>(const T*, int64_t, int, bool, void*, double*, T*, hipStream_t);
PFH_INST(float)
PFH_INST(double)
PFH_INST(c64)
PFH_INST(c128)
#undef PFH_INST

}  // namespace pfh
