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
#include <cstdlib>
#include <type_traits>

#include "../../include/pfaffian_hip.h"
#include "pf_batch.h"
#include "pf_blocked.h"
#include "pf_types.h"
#include "pf_wave.h"

namespace pfh {

// Digits per value of the split-precision trailing update, set by the
// |dlog| <= 1e-9 bound (measured against the FP64 oracle, |dlog| at S = 7 /
// 8 / 9: N = 8192 precision matrix 1.9e-7 / 3.9e-10 / 1.8e-11, random
// N = 4096 2.8e-8 / 3.5e-10 / 4.5e-12): S = 8 meets it by a factor 2.5 only,
// S = 9 by 50.
#ifndef PFH_MIXED_SLICES_DEFAULT
#define PFH_MIXED_SLICES_DEFAULT 9
#endif
constexpr int BP_THREADS = 128;    // columns per block in the row kernels
constexpr int BP_TILE = 64;        // trailing-update tile edge
constexpr int BC_T = 256;          // cooperative panel: threads per block = slots per block

// Upper 64 x 64 tile (ti <= tj) of block b, in an order that keeps each tile
// column on one XCD: workgroups go to the 8 XCDs round-robin, and XCD
// x = b % 8 takes the columns tj = x, x + 8, ..., each column's tiles
// ti = 0..tj in turn, so a column's operand (the same for all its tiles)
// is read into that XCD's L2 once.  False for the blocks past the last tile
// of their XCD (the grid is 8 times the largest XCD share).
__device__ __forceinline__ bool xcd_tile(int b, int ntile, int& ti, int& tj) {
    int rem = b >> 3;
    tj = b & 7;
    while (tj < ntile && rem > tj) { rem -= tj + 1; tj += 8; }
    ti = rem;
    return tj < ntile;
}
static int xcd_tile_blocks(int ntile) {
    int most = 0;
    for (int x = 0; x < 8; ++x) {
        int c = 0;
        for (int tj = x; tj < ntile; tj += 8) c += tj + 1;
        most = c > most ? c : most;
    }
    return 8 * most;
}
constexpr int BC_MAXG = 64;        // cooperative panel: blocks at most (one wave polls and reduces their records)
constexpr int BC_WAVES = BC_T / 64;
constexpr int WP_MAXG = 256;       // per-wave panel: participants at most (a poll lane reads up to four records)
// candidate records (described with bc_poll): the top bit carries the
// epoch's parity; "no candidate" is NaN bits in A and index 2^31 - 1, slot -1
// in B
constexpr unsigned long long BC_PAR = 1ull << 63;
constexpr unsigned long long BC_NONE_A = 0x7ff8000000000001ull;
constexpr unsigned long long BC_NONE_B = (0x7fffffffull << 32) | 0xffffffffull;

__host__ __device__ constexpr unsigned long long bc_par(int e) { return (e & 1) ? BC_PAR : 0ull; }

// Set by bp_finish / gj_out when a panel's poll gave up (the results of that
// call are NaN); read and cleared by blocked_take_status.  One per device.
__device__ int g_panel_timeout;

template <typename T>
struct BPState {
    int p, dead, counter, err;
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
    int* claim;      // per panel launch: the XCD-local panel's rank counter (n / (2 nb) + 2 of them)
    int64_t n;
    int nb;
    int spin;        // the panel's poll: reads of the records before its error flag
};

// Element (r, c) of the stale matrix seen through perm, from the upper triangle.
// Loads are unconditional (the diagonal when pr == pc) and the result is
// selected afterwards, so that a run of such reads is issued back to back
// instead of one branch-and-wait each.
template <typename T>
__device__ __forceinline__ T w_at(const T* W, int64_t n, int pr, int pc) {
    const int a = pr < pc ? pr : pc, b = pr < pc ? pc : pr;
    const T v = W[(int64_t)a * n + b];
    return pr < pc ? v : (pr > pc ? -v : zero<T>());
}

// c ? a : b, with complex values selected part by part (a select of whole
// structs can go through scratch memory)
template <typename T>
__device__ __forceinline__ T psel(bool c, T a, T b) {
    if constexpr (is_cplx<T>::value) return T{c ? a.re : b.re, c ? a.im : b.im};
    else return c ? a : b;
}

template <typename T>
__device__ __forceinline__ T wp(const T* W, const int* perm, int64_t n, int r, int c) {
    return w_at(W, n, perm[r], perm[c]);
}

// log|pivot| (as mantissa lm and exponent le) and the sign / phase, with a
// sign flip for an interchange
template <typename T>
__device__ __forceinline__ void accumulate(double& lm, int& le, double& ph_re, double& ph_im, T piv, bool sw) {
    double m = lm * absd(piv);
    int ex;
    m = frexp(m, &ex);
    lm = m;
    le += ex;
    if constexpr (is_cplx<T>::value) {
        const double a = hypot((double)piv.re, (double)piv.im);
        const double pr = (double)piv.re / a, pi = (double)piv.im / a;
        double nr = ph_re * pr - ph_im * pi, ni = ph_re * pi + ph_im * pr;
        const double nm = hypot(nr, ni);
        nr /= nm; ni /= nm;
        if (sw) { nr = -nr; ni = -ni; }
        ph_re = nr; ph_im = ni;
    } else {
        double s = ph_re;
        if (piv < T(0)) s = -s;
        if (sw) s = -s;
        ph_re = s;
    }
}

template <typename T>
__device__ __forceinline__ void accumulate(BPState<T>* st, T piv, bool sw) {
    accumulate(st->lm, st->le, st->ph_re, st->ph_im, piv, sw);
}

template <typename T>
__global__ void bp_init(BPArgs<T> a, unsigned long long* rec, unsigned long long* wrec) {
    const int64_t n = a.n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a.perm[i] = (int)i;
    const int64_t nclaim = n / (2 * a.nb) + 2;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nclaim; i += (int64_t)gridDim.x * blockDim.x)
        a.claim[i] = 0;
    // candidate records: "no candidate", with the parity opposite to that of
    // the first epoch each buffer serves (epochs count from 1: buffer 1
    // serves 1, 4, 7, ...; buffer 2 serves 2, 5, ...; buffer 0 serves 3, 6, ...)
    if (blockIdx.x == 0) {
        for (int q = threadIdx.x; q < 3 * BC_MAXG; q += blockDim.x) {
            const int buf = q / BC_MAXG;
            const unsigned long long pb = bc_par((buf == 0 ? 3 : buf) + 1);
            rec[2 * q] = BC_NONE_A | pb;
            rec[2 * q + 1] = BC_NONE_B | pb;
        }
        for (int q = threadIdx.x; q < 3 * WP_MAXG; q += blockDim.x) {   // the same for the per-wave panel's
            const int buf = q / WP_MAXG;
            const unsigned long long pb = bc_par((buf == 0 ? 3 : buf) + 1);
            wrec[2 * q] = BC_NONE_A | pb;
            wrec[2 * q + 1] = BC_NONE_B | pb;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        BPState<T>* st = a.st;
        st->p = 0; st->dead = 0; st->counter = 0; st->err = 0;
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
        const double m = (double)mag2(r);
        if (m == m) { best = m; bi = c; }   // NaN is never a candidate
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

// ---------------------------------------------------------------------------
// Cooperative panel: one launch factors a whole panel (np pair steps) with
// one grid barrier per pair step, in place of a launch per pair step.
//
// Columns are addressed by "slot", their index in the panel's starting W.
// Block b owns the slots k0 + 256 b + t (thread t) and keeps their panel
// vectors U_j, V_j in LDS.  An interchange only changes two maps, kept in
// LDS by every block and updated identically from the common pivot
// sequence: logical -> slot for the panel's own positions (Qpan) and
// slot -> logical for the block's slots (Qi).  Row k of pair i is stored
// as V_i directly, in slot coordinates.  At the end of the panel the
// logical -> slot map goes to `perm` for the trailing update, whose P reads
// are then by slot.
//
// After computing the next pivot row, every block publishes its best
// candidate (|value|, logical index, slot) as a record that is also its
// barrier flag: every block polls all records, reduces them itself, and then
// issues one round of loads for everything the pair step needs (the pivot's
// panel entries and value, the entries at the pre-interchange slots of rows
// k+1 and k+2, and those two rows of W at its own slots).
// ---------------------------------------------------------------------------

// acc += a b for complex values as four FMAs
template <typename R>
__device__ __forceinline__ void cfma(cplx<R>& acc, const cplx<R>& a, const cplx<R>& b) {
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(a.im, b.re, acc.im);
}
template <typename R>
__device__ __forceinline__ void cfma(R&, const R&, const R&) {}   // real types take the plain form

// U_j, V_j of one slot, adjacent in the panel's LDS store
template <typename T>
struct alignas(2 * sizeof(T)) BCPair {
    T u, v;
};

// Candidate records: the grid barrier and the exchange of the partials in
// one.  A block's record for epoch E (one per pivot row) is two 64-bit words
// in buffer E % 3, each carrying E's parity in its top bit:
//   A = the candidate's magnitude, a double >= 0 (NaN bits: no candidate);
//   B = its logical index (bits 32..62) and slot (bits 0..31).
// The buffer a block overwrites at epoch E last held its record E - 3, which
// has the other parity, and every block has finished reading it: a block
// publishing E has seen every record E - 1, each published after its
// author's reads of E - 2 and earlier.  So a reader polling for E sees either
// E - 3 (keeps polling) or E, word by word, and once both words show E, the
// block's other exchange stores have landed: they precede the record's
// stores.
//
// The readers need the winner's entries (U_j, V_j for j < i, and its value
// V_i).  Real types: the winning block copies them into its candidate row
// (buffer E % 3, the same reuse argument), one contiguous read for the
// others.  Complex types: the readers gather them from the panel vectors at
// the winner's slot, as they gather those of rows k+1 and k+2.  Measured
// (ms, gather / row): f64 N = 8192 30.3 / 25.3, N = 4096 12.4 / 9.8;
// c128 N = 4096 15.4 / 16.8.  A slot-major copy of the panel vectors, which
// makes every one of these reads contiguous but every thread's two new
// entries a store to its own line, was slower for f64 too (30.9 ms at
// N = 8192; c128 15.9 ms).

template <typename T>
struct BPCoop {
    unsigned long long* rec;   // [3][BC_MAXG][2]: records A, B by epoch % 3
    T* cand;                   // [3][BC_MAXG][2 nb]: a block's candidate entries (real types), by epoch % 3
    unsigned long long* wrec;  // [3][WP_MAXG][2]: the per-wave panel's records, by epoch % 3
};

// Accesses to the data the panel's blocks exchange within a launch (panel
// vectors, candidate rows), by exchange mode XM:
//   XM_AGENT (real types, blocks anywhere): agent-scope relaxed atomics,
//     coherent across the XCDs' L2s on their own, so that a record needs no
//     L2 write-back / invalidate, only a wait for the block's stores before
//     it (N = 4096 f64 13.5 -> 12.7 ms, N = 8192 33.6 -> 30.9 ms, results
//     bit-identical);
//   XM_FENCE (complex types, blocks anywhere): plain accesses and
//     agent-scope fences (as two atomics per value they were slower, c128
//     N = 4096 19.4 -> 20.5 ms);
//   XM_XCD (every type, all blocks on one XCD): stores at workgroup scope
//     (written through to the XCD's L2) and loads at system scope (they
//     miss the CU's L1 and are served by that L2), so an exchange never
//     leaves the XCD.  A flag round trip between two workgroups of one XCD
//     measured 420-470 ns this way against 810-1240 ns with agent-scope
//     accesses (which go past the L2); across two XCDs this form is not
//     coherent, which is why the panel checks its blocks' XCD.
constexpr int XM_AGENT = 0, XM_FENCE = 1, XM_XCD = 2;

template <int XM, typename V>
__device__ __forceinline__ void xb_st(V* p, V v) {
    if constexpr (XM == XM_XCD) {
        if constexpr (sizeof(V) == 16) {
            unsigned long long w[2];
            __builtin_memcpy(w, &v, 16);
            unsigned long long* q = reinterpret_cast<unsigned long long*>(p);
            __hip_atomic_store(q, w[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_store(q + 1, w[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else if constexpr (sizeof(V) == 8) {
            unsigned long long w;
            __builtin_memcpy(&w, &v, 8);
            __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), w, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
            static_assert(sizeof(V) == 4, "exchanged values are 4, 8 or 16 bytes");
            unsigned w;
            __builtin_memcpy(&w, &v, 4);
            __hip_atomic_store(reinterpret_cast<unsigned*>(p), w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    } else if constexpr (XM == XM_AGENT) {
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        *p = v;
    }
}
template <int XM, typename V>
__device__ __forceinline__ V xb_ld(const V* p) {
    if constexpr (XM == XM_XCD) {
        V v;
        if constexpr (sizeof(V) == 16) {
            const unsigned long long* q = reinterpret_cast<const unsigned long long*>(p);
            unsigned long long w[2];
            w[0] = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            w[1] = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __builtin_memcpy(&v, w, 16);
        } else if constexpr (sizeof(V) == 8) {
            const unsigned long long w = __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p),
                                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __builtin_memcpy(&v, &w, 8);
        } else {
            const unsigned w = __hip_atomic_load(reinterpret_cast<const unsigned*>(p), __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_SYSTEM);
            __builtin_memcpy(&v, &w, 4);
        }
        return v;
    } else if constexpr (XM == XM_AGENT) {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        return *p;
    }
}
// record words: the same scopes as the data for the XCD-local form, agent
// scope otherwise
template <int XM>
__device__ __forceinline__ void rec_st(unsigned long long* p, unsigned long long v) {
    if constexpr (XM == XM_XCD) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <int XM>
__device__ __forceinline__ unsigned long long rec_ld(const unsigned long long* p) {
    if constexpr (XM == XM_XCD) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    else return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T, bool LOCAL>
constexpr int panel_xm() { return LOCAL ? XM_XCD : (is_cplx<T>::value ? XM_FENCE : XM_AGENT); }
// The winner's entries reach the other blocks through a candidate row that
// the winning block writes before its record (real types, device-wide
// exchange), else they are gathered from the panel vectors at its slot.  In
// the XCD-local form the gather is the faster (f64 N = 4096 9.26 -> 7.93 ms,
// N = 8192 24.5 -> 21.9 ms, gradient N = 2048 5.83 -> 5.14 ms): the export
// of the winner's registers, the row's stores and the wait for them took
// ~1.3k cycles of each pair step's publication, and reads served by one
// XCD's L2 gather quickly.
template <typename T, int XM>
constexpr bool panel_cand() { return !is_cplx<T>::value && XM != XM_XCD; }

// Wait for every block's record of epoch E and reduce them (one wave, one
// block per lane): the winning block, or -1 when no block has a finite
// candidate, with its logical index and slot.  A record that does not show
// up within the spin limit sets err (the output becomes NaN); every record
// word ever written decodes to indices in range, so nothing faults then.
template <int XM>
__device__ __forceinline__ void bc_poll(const unsigned long long* rec_all, int e, int G, int limit, int* err,
                                        int* s_w, int* s_p, int* s_sp, int* s_fail) {
    if (threadIdx.x < 64) {
        const int g = threadIdx.x;
        const unsigned long long pb = bc_par(e);
        const unsigned long long* rec = rec_all + ((int64_t)(e % 3) * BC_MAXG + (g < G ? g : 0)) * 2;
        unsigned long long wa = BC_NONE_A, wb = BC_NONE_B;
        int fail = 0;
        for (int reads = 0;; ++reads) {
            if (reads >= limit) {   // seconds by default: the grid is co-resident, so this is a failure
                if (g == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                fail = 1;
                break;
            }
            if (g < G) {
                wa = rec_ld<XM>(&rec[0]);
                wb = rec_ld<XM>(&rec[1]);
            }
            // another block that gave up has set err: its record may never
            // come (checked every 16th read: at agent scope this load takes
            // longer than a record read through the XCD's L2)
            const int failed = (g == 0 && (reads & 15) == 15)
                                   ? __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
            const bool ok = g >= G || ((((wa ^ pb) | (wb ^ pb)) & BC_PAR) == 0);
            if (__shfl(failed, 0) != 0) { fail = 1; break; }
            if (__all(ok)) break;
            __builtin_amdgcn_s_sleep(1);
        }
        if (g == 0) *s_fail = fail;
        if (g >= G) { wa = BC_NONE_A; wb = BC_NONE_B; }
        const double m = __longlong_as_double((long long)(wa & ~BC_PAR));
        const double best = (m == m) ? m : -1.0;
        const int idx = (int)((wb >> 32) & 0x7fffffffull);
        const int slot = (int)(unsigned)(wb & 0xffffffffull);
        double b;
        int bi;
        wave_argmax(best, idx, b, bi);                                          // ties: lowest index
        const unsigned long long hit = __ballot(b >= 0.0 && best == b && idx == bi);
        const int w = hit ? (int)__builtin_ctzll(hit) : -1;
        if (g == 0) *s_w = w;
        if (w >= 0 && g == w) { *s_p = idx; *s_sp = slot; }
    }
    __syncthreads();
    if constexpr (XM == XM_FENCE) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Entry (r, c) of the stale matrix in slot coordinates (upper triangle stored).
template <typename T>
__device__ __forceinline__ T wsl(const T* W, int64_t n, int r, int c) {
    return w_at(W, n, r, c);
}

__device__ __forceinline__ bool bc_better(double b, int i, double bb, int bi) {
    return b > bb || (b == bb && i < bi);
}

// Own-slot panel vectors in registers.  A vector of more than 16 values with
// a dynamic index goes to scratch memory, so the NB values sit in two halves
// of at most 16; a dynamic index is only ever taken within one half, named at
// compile time (the pair steps run in stages of 16), and compiles to
// register indexing.  Every dynamic index must lie inside its half even on
// paths that do not use the result: the compiler may execute an indexed
// write speculatively (both arms of a branch, then a select), and an index
// past the vector would overwrite whatever register follows it.  The callers
// clamp such indices (masking them inside the accessors instead sent the
// vectors to scratch memory).
template <typename R, int N>
struct BPRegHalf {
    static constexpr int L0 = N > 16 ? 16 : N, L1 = N > 16 ? N - 16 : 8;
    typedef R V0 __attribute__((ext_vector_type(L0)));
    typedef R V1 __attribute__((ext_vector_type(L1)));
    V0 v0;
    V1 v1;
    template <int J> __device__ __forceinline__ R at() const {
        if constexpr (J < 16) return v0[J];
        else return v1[J - 16];
    }
    template <int J> __device__ __forceinline__ void put(R x) {
        if constexpr (J < 16) v0[J] = x;
        else v1[J - 16] = x;
    }
    template <int H> static constexpr int len() { return H == 0 ? L0 : L1; }
    template <int H> __device__ __forceinline__ R get(int i) const {
        if constexpr (H == 0) return v0[i];
        else return v1[i];
    }
    template <int H> __device__ __forceinline__ void set(int i, R x) {
        if constexpr (H == 0) v0[i] = x;
        else v1[i] = x;
    }
};
template <typename T, int N>
struct BPRegs {
    BPRegHalf<T, N> v;
    template <int J> __device__ __forceinline__ T at() const { return v.template at<J>(); }
    template <int J> __device__ __forceinline__ void put(T x) { v.template put<J>(x); }
    template <int H> __device__ __forceinline__ T get(int i) const { return v.template get<H>(i); }
    template <int H> __device__ __forceinline__ void set(int i, T x) { v.template set<H>(i, x); }
};
template <typename R, int N>
struct BPRegs<cplx<R>, N> {
    BPRegHalf<R, N> re, im;
    template <int J> __device__ __forceinline__ cplx<R> at() const { return {re.template at<J>(), im.template at<J>()}; }
    template <int J> __device__ __forceinline__ void put(cplx<R> x) { re.template put<J>(x.re); im.template put<J>(x.im); }
    template <int H> __device__ __forceinline__ cplx<R> get(int i) const {
        return {re.template get<H>(i), im.template get<H>(i)};
    }
    template <int H> __device__ __forceinline__ void set(int i, cplx<R> x) {
        re.template set<H>(i, x.re);
        im.template set<H>(i, x.im);
    }
};

// f(std::integral_constant<int, I>) for I = I0 .. N-1
template <int I, int N, typename F>
__device__ __forceinline__ void bp_sfor(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        bp_sfor<I + 1, N>(f);
    }
}

// Block argmax of (best, logical index) and publication of the winner for
// epoch e: with candidate rows, its entries U_j, V_j (j < cnt) and its value
// V_cnt first (the winning thread passes them from its registers through
// LDS, Wx, to 2 cnt + 2 storing threads); then, once this block's exchange
// stores have landed, the record.
template <typename T, int NB, int XM, typename Regs>
__device__ __forceinline__ void bc_publish(double best, int bi, int lg, int x, bool own, int cnt, int e, int rank,
                                           const BPCoop<T>& co, const Regs& Ur, const Regs& Vr, T* Wx,
                                           double* r_best, int* r_idx, int* r_slot) {
    constexpr bool CAND = panel_cand<T, XM>();
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    double b;
    int ix;
    wave_argmax(best, bi, b, ix);                                       // ties: lowest index
    if (lane == 0) { r_best[wv] = b; r_idx[wv] = ix; }
    __syncthreads();
    double gb = r_best[0];
    int gi = r_idx[0];
    for (int u = 1; u < BC_WAVES; ++u)
        if (bc_better(r_best[u], r_idx[u], gb, gi)) { gb = r_best[u]; gi = r_idx[u]; }
    const bool found = gb >= 0.0;
    if (found && own && best >= 0.0 && lg == gi) {
        *r_slot = x;
        if constexpr (CAND) {
            // pairs 0 .. cnt in chunks of 8 (the entries past V_cnt are zero)
            bp_sfor<0, NB / 8>([&](auto cc) {
                constexpr int C = decltype(cc)::value;
                if (8 * C > cnt) return;
                bp_sfor<0, 8>([&](auto qc) {
                    constexpr int J = 8 * C + decltype(qc)::value;
                    reinterpret_cast<BCPair<T>*>(Wx)[J] = BCPair<T>{Ur.template at<J>(), Vr.template at<J>()};
                });
            });
        }
    }
    if constexpr (CAND) {
        __syncthreads();
        T* cand = co.cand + ((int64_t)(e % 3) * BC_MAXG + rank) * 2 * NB;
        if (found && t < 2 * cnt + 2) xb_st<XM>(&cand[t], Wx[t]);
    }
    if constexpr (XM != XM_FENCE) __builtin_amdgcn_s_waitcnt(0);   // this thread's stores have landed
    else __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");        // this thread's stores are visible device-wide
    __syncthreads();
    if (t == 0) {
        const unsigned long long pb = bc_par(e);
        unsigned long long wa = BC_NONE_A | pb, wb = BC_NONE_B | pb;
        if (found) {
            wa = (unsigned long long)__double_as_longlong(gb) | pb;   // gb >= +0: sign bit clear
            wb = pb | ((unsigned long long)(unsigned)gi << 32) | (unsigned long long)(unsigned)*r_slot;
        }
        unsigned long long* rec = co.rec + ((int64_t)(e % 3) * BC_MAXG + rank) * 2;
        rec_st<XM>(&rec[0], wa);
        rec_st<XM>(&rec[1], wb);
    }
}

// LOCAL: the launch has 8 G blocks; those that find themselves on XCD 0
// (its id read from the hardware) claim the panel ranks 0 .. G-1 from a
// counter of this launch, the rest return at once.  Workgroups go to the 8
// XCDs round-robin (from wherever the dispatch starts), so XCD 0 receives
// exactly G of them, and the G share its L2, through which they exchange
// (XM_XCD).  Were fewer to land there, a rank would stay unclaimed and the
// poll would time out (NaN, reported by pfh_stream_status), never a wrong
// value.
//
// The own slot's panel entries U_j, V_j live in registers (the corrections
// read them with static indices, in chunks; only U_i, V_i and V_{i+1} of the
// current pair step take a dynamic index, within one 16-value half).  In LDS
// it took ~1.9k of ~10.5k cycles per pair step to read them (N = 8192 f64).
template <typename T, int NB, bool LOCAL>
__global__ void __launch_bounds__(BC_T) bp_panel(BPArgs<T> a, BPCoop<T> co, int k0, int np, int epoch0) {
    constexpr int XM = panel_xm<T, LOCAL>();
    constexpr bool CAND = panel_cand<T, XM>();
    constexpr int NH = (NB + 15) / 16;                  // stages of 16 pair steps
    constexpr int CH = is_cplx<T>::value ? 4 : 8;       // pairs per chunk of the corrections
    static_assert(NB % 8 == 0 && NB <= 32, "panel widths 8, 16, 24, 32");
    __shared__ int Qpan[2 * NB + 2];           // slot of logical k0 + q
    // entries at one slot, interleaved (U_j at 2j, V_j at 2j + 1): the
    // winner's (the new row k+1, its value V_i at 2i + 1), and those at the
    // pre-interchange slots of rows k+1 and k+2 (one array, so that choosing
    // one of them is an index, not a select between two loads): Bx[0]
    // winner, Bx[1] pre-interchange k+1, Bx[2] k+2
    __shared__ __align__(16) T Bx[3][2 * NB];
    __shared__ __align__(16) T Wx[2 * NB];     // the block winner's entries (candidate rows)
    __shared__ double r_best[BC_WAVES];
    __shared__ int r_idx[BC_WAVES], r_slot;
    __shared__ int s_w, s_p, s_sp, s_fail;

    __shared__ int s_rank;
    int rank = (int)blockIdx.x;
    const int G = LOCAL ? (int)(gridDim.x >> 3) : (int)gridDim.x;
    if constexpr (LOCAL) {
        if (__builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11)) != 0) return;   // HW_REG_XCC_ID[3:0]
        if (threadIdx.x == 0) s_rank = atomicAdd(&a.claim[k0 / (2 * NB)], 1);
        __syncthreads();
        rank = s_rank;
        if (rank >= G) return;
    }
    BPState<T>* st = a.st;
    if (st->dead || st->err) return;            // an earlier panel gave up: the output is NaN already
    const int64_t n = a.n;
    const int t = threadIdx.x;
    const int x = k0 + rank * BC_T + t;         // own slot
    const bool own = x < n;
    const int xs = own ? x : k0;                // a valid slot for masked reads
    T* PU = a.P;                                // U_j[x] at PU[j*n + x]
    T* PV = a.P + (int64_t)NB * n;              // V_j[x] at PV[j*n + x]

    double lm = 0.0, ph_re = 0.0, ph_im = 0.0;   // pivot log / phase, kept by block 0 thread 0
    int le = 0;
    const bool keeper = rank == 0 && t == 0;
    if (keeper) { lm = st->lm; le = st->le; ph_re = st->ph_re; ph_im = st->ph_im; }

    for (int q = t; q < 2 * NB + 2; q += BC_T) Qpan[q] = k0 + q;
    int lg = x;                                 // logical index of the own slot
    BPRegs<T, NB> Ur, Vr;                       // U_j, V_j at the own slot
    bp_sfor<0, NB>([&](auto jc) {
        constexpr int J = decltype(jc)::value;
        Ur.template put<J>(zero<T>());
        Vr.template put<J>(zero<T>());
    });

    // first pivot row of the panel: row k0 of W (nothing pending); epoch
    // epoch0 + 1 + i carries the candidates of pair step i
    {
        double best = -1.0;
        T r = zero<T>();
        if (own && x > k0) {
            r = a.W[(int64_t)k0 * n + x];
            xb_st<XM>(&PV[x], r);
            const double m = (double)mag2(r);
            if (m == m) best = m;
        }
        Vr.template put<0>(r);
        bc_publish<T, NB, XM>(best, lg, lg, x, own, 0, epoch0 + 1, rank, co, Ur, Vr, Wx, r_best, r_idx, &r_slot);
    }

    int quit = 0;   // 1: the panel is done (or the matrix); 2: leave at once (zero pivot, failed poll)
    bp_sfor<0, NH>([&](auto hc) {
        constexpr int H = decltype(hc)::value;
        if (quit) return;
        const int iend = min(np, 16 * H + 16);
        for (int i = 16 * H; i < iend; ++i) {
            const int il = i - 16 * H;
            const int k = k0 + 2 * i;
            const int e = epoch0 + 1 + i;
            const bool has_next = i + 1 < np;       // row k+2 is this panel's next pivot row
            const int sa = Qpan[k + 1 - k0];
            const int sb = has_next ? Qpan[k + 2 - k0] : -1;
            const int sbs = sb >= 0 ? sb : sa;

            // ---- loads that do not depend on the pivot, issued before the
            //      poll: the entries at the pre-interchange slots of rows k+1
            //      and k+2 that this epoch's records do not cover (U_j for
            //      j < i-1, V_j for j < i: stored before an earlier record;
            //      U_{i-1} and V_i come after the poll), by wave 1, so that
            //      the polling wave does not wait for them; and row k+2 of W
            //      at the own slots (its slot changes only when the pivot is
            //      k+2)
            if (t >= 64 && t < 64 + 2 * NB) {
                const int l = t - 64, j = l >> 1;
                const bool old = (l & 1) ? j < i : j < i - 1;
                const T* src = ((l & 1) ? PV : PU) + (int64_t)j * n;
                const T v1 = xb_ld<XM>(&src[sa]), v2 = xb_ld<XM>(&src[sbs]);
                Bx[1][l] = psel(old, v1, zero<T>());
                Bx[2][l] = psel(old && sb >= 0, v2, zero<T>());
            }
            T w2 = zero<T>();
            if (has_next) w2 = wsl(a.W, n, sb, xs);

            bc_poll<XM>(co.rec, e, G, a.spin, &st->err, &s_w, &s_p, &s_sp, &s_fail);
            if (s_fail) { quit = 2; return; }                  // err is set: the result will be NaN
            const int wb = s_w;
            int p, sp;
            if (wb >= 0) { p = s_p; sp = s_sp; }
            else { p = k + 1; sp = sa; }                       // no finite candidate (NaN row)
            const bool sw = p != k + 1;
            // the interchange k+1 <-> p as seen from the own slot
            if (x == sp) lg = k + 1;
            else if (sw && x == sa) lg = p;
            const int y1 = sp;
            const int y2 = (p == k + 2) ? sa : sb;
            const int y2s = has_next ? y2 : y1;

            // ---- the loads that do (each unconditional from a valid address,
            //      masked after): the winner's entries and value, U_{i-1}
            //      and V_i at the pre-interchange slots of rows k+1 and k+2,
            //      and rows k+1 (and k+2 when the pivot is k+2) of W at the
            //      own slots
            if (t < 2 * NB) {
                const int j = t >> 1;
                const bool isv = t & 1;
                const bool ok = j < i || (isv && j == i);
                const T* src = (isv ? PV : PU) + (int64_t)j * n;
                const T* w0 = CAND ? co.cand + ((int64_t)(e % 3) * BC_MAXG + (wb >= 0 ? wb : 0)) * 2 * NB + t
                                   : &src[sp];
                const T v0 = xb_ld<XM>(w0);
                Bx[0][t] = psel(ok, v0, zero<T>());
                if (isv ? j == i : j == i - 1) {
                    const T v1 = xb_ld<XM>(&src[sa]), v2 = xb_ld<XM>(&src[sbs]);
                    Bx[1][t] = v1;
                    Bx[2][t] = psel(sb >= 0, v2, zero<T>());
                }
            }
            const bool act1 = own && lg >= k + 2;
            T w1 = wsl(a.W, n, y1, xs);
            if (!act1) w1 = zero<T>();
            T w12 = wsl(a.W, n, y1, y2s);
            if (!has_next) w12 = zero<T>();
            if (has_next && p == k + 2) w2 = wsl(a.W, n, sa, xs);
            if (!(has_next && own && lg >= k + 3)) w2 = zero<T>();
            __syncthreads();
            const T* b1 = Bx[(wb >= 0) ? 0 : 1];               // the pivot's entries (row k+1 at the start)
            const T piv = b1[2 * i + 1];                       // V_i there
            if (is_zero(piv)) {                                // same decision in every block
                if (keeper) st->dead = 1;
                quit = 2;
                return;
            }
            if (keeper) accumulate(lm, le, ph_re, ph_im, piv, sw);
            // the interchange in the shared map (every read of Qpan in this
            // step came before the block barrier in bc_poll)
            if (sw && t == 0) {
                Qpan[k + 1 - k0] = sp;
                if ((unsigned)(p - k0) < (unsigned)(2 * NB + 2)) Qpan[p - k0] = sa;
            }
            const T* b2 = Bx[(p == k + 2) ? 1 : 2];
            if (k + 2 >= n) { quit = 1; return; }              // last pair of the matrix

            const T inv = recip(piv);
            // corrections of the pairs j < i, in chunks (entries past i are
            // zero in b1, b2 and the registers -- V_i aside, whose
            // coefficients b1u_i and b2u_i are zero -- so a partial chunk adds
            // nothing); the chunk's b1, b2 pairs are read first
            // (complex: each product accumulates as four FMAs into separate
            // sums of the added and the subtracted terms, which halves the
            // dependent chains; a complex product written out costs six
            // operations)
            T r1 = zero<T>(), rl = zero<T>(), rk = zero<T>();
            T r1m = zero<T>(), rlm = zero<T>(), rkm = zero<T>();
            const BCPair<T>* P1 = reinterpret_cast<const BCPair<T>*>(b1);
            const BCPair<T>* P2 = reinterpret_cast<const BCPair<T>*>(b2);
            bp_sfor<0, NB / CH>([&](auto cc) {
                constexpr int C = decltype(cc)::value;
                if (C * CH >= i) return;
                BCPair<T> c1[CH], c2[CH];
                bp_sfor<0, CH>([&](auto qc) {
                    constexpr int Q = decltype(qc)::value;
                    c1[Q] = P1[C * CH + Q];
                    c2[Q] = P2[C * CH + Q];
                });
                bp_sfor<0, CH>([&](auto qc) {
                    constexpr int Q = decltype(qc)::value;
                    constexpr int J = C * CH + Q;
                    const T uj = Ur.template at<J>(), vj = Vr.template at<J>();
                    const T b1u = c1[Q].u, b1v = c1[Q].v, b2u = c2[Q].u, b2v = c2[Q].v;
                    if constexpr (is_cplx<T>::value) {
                        cfma(r1, b1u, vj); cfma(r1m, b1v, uj);
                        cfma(rl, b2u, vj); cfma(rlm, b2v, uj);
                        cfma(rk, b1u, b2v); cfma(rkm, b1v, b2u);
                    } else {
                        r1 = r1 + b1u * vj - b1v * uj;
                        rl = rl + b2u * vj - b2v * uj;
                        rk = rk + b1u * b2v - b1v * b2u;
                    }
                });
            });
            if constexpr (is_cplx<T>::value) {
                r1 = r1 - r1m;
                rl = rl - rlm;
                rk = rk - rkm;
            }
            T u = zero<T>();
            if (act1) {
                u = (w1 + r1) * inv;
                xb_st<XM>(&PU[(int64_t)i * n + x], u);
            }
            Ur.template set<H>(il, u);
            if (!has_next) { quit = 1; return; }

            // ---- row k+2 (the next pivot row) at the own slots -> V_{i+1}
            const T u_kl = (w12 + rk) * inv;   // U_i at row k+2
            const T v_kl = b2[2 * i + 1];      // V_i at row k+2
            const T vi = Vr.template get<H>(il);
            double best = -1.0;
            T r = zero<T>();
            if (own && lg >= k + 3) {
                r = w2 + rl + u_kl * vi - v_kl * u;
                xb_st<XM>(&PV[(int64_t)(i + 1) * n + x], r);
                const double m = (double)mag2(r);
                if (m == m) best = m;
            }
            // V_{i+1}: the next half's first entry after the 16th step of a
            // stage; the clamps only keep a speculated write inside the half
            // (il + 1 is in range wherever the write is used)
            if constexpr (H + 1 < NH) {
                if (il == 15) Vr.template put<16 * (H + 1)>(r);
                else Vr.template set<H>(min(il + 1, 15), r);
            } else {
                constexpr int LH = BPRegHalf<double, NB>::template len<H>();
                Vr.template set<H>(min(il + 1, LH - 1), r);    // i + 1 < np <= NB
            }
            bc_publish<T, NB, XM>(best, lg, lg, x, own, i + 1, e + 1, rank, co, Ur, Vr, Wx, r_best, r_idx,
                                  &r_slot);
        }
    });
    if (quit == 2) return;

    if (own) a.perm[lg] = x;   // logical -> slot, for the trailing update
    if (keeper) { st->lm = lm; st->le = le; st->ph_re = ph_re; st->ph_im = ph_im; }
}

// ---------------------------------------------------------------------------
// The panel with one participant per wave (the XCD-local exchange only).
// Each wave owns 64 consecutive slots and does by itself what a block of
// bp_panel does with its four: it polls every participant's record, reduces
// them, gathers the pivot rows' panel entries into its own LDS, updates its
// slots and publishes its own record (its 64 slots' argmax).  A workgroup is
// one wave, so a pair step holds no workgroup barrier, and every load that
// depends on the pivot is issued in one run right after the poll (the
// gathered entries first, the stale rows of W after them), with nothing read
// before the poll that the poll's own waits would hold up.  The arithmetic,
// and its order, is bp_panel's: the two forms agree bit for bit.
// ---------------------------------------------------------------------------
// Flat index of entry (pr, pc) in the stored upper triangle, and the entry
// from the value read there: the load and the sign / zero fix-up are split
// so that the fix-up can sit after other work.
__device__ __forceinline__ int64_t wraw_ix(int64_t n, int pr, int pc) {
    const int a = pr < pc ? pr : pc, b = pr < pc ? pc : pr;
    return (int64_t)a * n + b;
}
template <typename T>
__device__ __forceinline__ T wfix(T v, int pr, int pc) {
    return pr < pc ? v : (pr > pc ? -v : zero<T>());
}

// Qpan[q] of the panel's first 2 NB + 2 logical rows, kept in two registers
// (lane q of q0 for q < 64, lane q - 64 of q1 above); q is wave-uniform.
__device__ __forceinline__ int wp_q(int q0, int q1, int q) {
    return q < 64 ? __builtin_amdgcn_readlane(q0, q) : __builtin_amdgcn_readlane(q1, q - 64);
}
__device__ __forceinline__ void wp_qset(int& q0, int& q1, int q, int v) {
    const int lane = (int)threadIdx.x;
    if (q < 64) q0 = lane == q ? v : q0;
    else q1 = lane == q - 64 ? v : q1;
}

// Every participant's record of epoch e, polled and reduced by the whole
// wave (lane l reads records l, l + 64, ..., NMR of them, every read issued
// before any is tested): false when a record did not show up within the spin
// limit or another participant gave up (err is set, the output becomes NaN);
// else found = some participant has a finite candidate, with its logical
// index p and slot sp (ties: lowest index).
template <int NMR>
__device__ __forceinline__ bool wp_poll_n(const unsigned long long* rec, unsigned long long pb, int G, int limit,
                                          int* err, int& p, int& sp, bool& found) {
    const int lane = (int)threadIdx.x;
    unsigned long long wa[NMR], wb[NMR];
    for (int reads = 0;; ++reads) {
        if (reads >= limit) {
            if (lane == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
#pragma unroll
        for (int m = 0; m < NMR; ++m) {
            const int g = min(lane + 64 * m, G - 1);
            wa[m] = rec_ld<XM_XCD>(&rec[2 * g]);
            wb[m] = rec_ld<XM_XCD>(&rec[2 * g + 1]);
        }
        unsigned long long bad = 0;
#pragma unroll
        for (int m = 0; m < NMR; ++m) bad |= (wa[m] ^ pb) | (wb[m] ^ pb);
        // another participant that gave up has set err: its record may never come
        if ((reads & 15) == 15 &&
            __builtin_amdgcn_readfirstlane(__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0)
            return false;
        if (__all((bad & BC_PAR) == 0)) break;
        __builtin_amdgcn_s_sleep(1);
    }
    double best = -1.0;
    int bi = 0x7fffffff, bs = 0;
#pragma unroll
    for (int m = 0; m < NMR; ++m) {
        const double mg = __longlong_as_double((long long)(wa[m] & ~BC_PAR));
        const int idx = (int)((wb[m] >> 32) & 0x7fffffffull);
        if (lane + 64 * m < G && mg == mg && bc_better(mg, idx, best, bi)) {
            best = mg;
            bi = idx;
            bs = (int)(unsigned)(wb[m] & 0xffffffffull);
        }
    }
    double b;
    int ix;
    wave_argmax(best, bi, b, ix);
    found = b >= 0.0;
    if (found) {
        const unsigned long long hit = __ballot(best == b && bi == ix);
        p = ix;
        sp = __builtin_amdgcn_readlane(bs, (int)__builtin_ctzll(hit));
    }
    return true;
}
__device__ __forceinline__ bool wp_poll(const unsigned long long* rec_all, int e, int G, int limit, int* err,
                                        int& p, int& sp, bool& found) {
    const unsigned long long* rec = rec_all + (int64_t)(e % 3) * WP_MAXG * 2;
    const unsigned long long pb = bc_par(e);
    if (G <= 64) return wp_poll_n<1>(rec, pb, G, limit, err, p, sp, found);
    if (G <= 128) return wp_poll_n<2>(rec, pb, G, limit, err, p, sp, found);
    return wp_poll_n<4>(rec, pb, G, limit, err, p, sp, found);
}

// The wave's argmax of (best, logical index) over its slots, and its record
// for epoch e once the wave's stores (its new panel entries) have landed.
__device__ __forceinline__ void wp_publish(double best, int lg, int x, int e, int rank, unsigned long long* rec_all) {
    const int lane = (int)threadIdx.x;
    double b;
    int ix;
    wave_argmax(best, lg, b, ix);                                        // ties: lowest index
    const bool found = b >= 0.0;
    int slot = 0;
    if (found) slot = __builtin_amdgcn_readlane(x, (int)__builtin_ctzll(__ballot(best == b && lg == ix)));
    __builtin_amdgcn_s_waitcnt(0);                                       // this wave's stores have landed
    if (lane == 0) {
        const unsigned long long pb = bc_par(e);
        unsigned long long wa = BC_NONE_A | pb, wb = BC_NONE_B | pb;
        if (found) {
            wa = (unsigned long long)__double_as_longlong(b) | pb;       // b >= +0: sign bit clear
            wb = pb | ((unsigned long long)(unsigned)ix << 32) | (unsigned long long)(unsigned)slot;
        }
        unsigned long long* rec = rec_all + ((int64_t)(e % 3) * WP_MAXG + rank) * 2;
        rec_st<XM_XCD>(&rec[0], wa);
        rec_st<XM_XCD>(&rec[1], wb);
    }
}

// Launched with 8 G workgroups of one wave; those on XCD 0 claim the
// participant ranks 0 .. G-1 (as bp_panel's blocks do) and the rest return.
template <typename T, int NB>
__global__ void __launch_bounds__(64) bp_wpanel(BPArgs<T> a, BPCoop<T> co, int k0, int np, int epoch0) {
    constexpr int XM = XM_XCD;
    constexpr int NH = (NB + 15) / 16;                  // stages of 16 pair steps
    constexpr int CH = is_cplx<T>::value ? 4 : 8;       // pairs per chunk of the corrections
    static_assert(NB % 8 == 0 && NB <= 32, "panel widths 8, 16, 24, 32");
    // entries at one slot, interleaved (U_j at 2j, V_j at 2j + 1): Bx[0] the
    // winner's (the new row k+1), Bx[1] the pre-interchange row k+1's,
    // Bx[2] row k+2's
    __shared__ __align__(16) T Bx[3][2 * NB];
    const int lane = (int)threadIdx.x;
    const int G = (int)(gridDim.x >> 3);
    if (__builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11)) != 0) return;   // HW_REG_XCC_ID[3:0]
    int r0 = 0;
    if (lane == 0) r0 = atomicAdd(&a.claim[k0 / (2 * NB)], 1);
    const int rank = __builtin_amdgcn_readfirstlane(r0);
    if (rank >= G) return;
    BPState<T>* st = a.st;
    if (st->dead || st->err) return;            // an earlier panel gave up: the output is NaN already
    const int64_t n = a.n;
    const int x = k0 + rank * 64 + lane;        // own slot
    const bool own = x < n;
    const int xs = own ? x : k0;                // a valid slot for masked reads
    T* PU = a.P;                                // U_j[x] at PU[j*n + x]
    T* PV = a.P + (int64_t)NB * n;              // V_j[x] at PV[j*n + x]

    double lm = 0.0, ph_re = 0.0, ph_im = 0.0;  // pivot log / phase, kept by participant 0 lane 0
    int le = 0;
    const bool keeper = rank == 0 && lane == 0;
    if (keeper) { lm = st->lm; le = st->le; ph_re = st->ph_re; ph_im = st->ph_im; }

    int q0 = k0 + lane, q1 = k0 + 64 + lane;    // Qpan: slot of logical k0 + q
    int lg = x;                                 // logical index of the own slot
    BPRegs<T, NB> Ur, Vr;                       // U_j, V_j at the own slot
    bp_sfor<0, NB>([&](auto jc) {
        constexpr int J = decltype(jc)::value;
        Ur.template put<J>(zero<T>());
        Vr.template put<J>(zero<T>());
    });

    // first pivot row of the panel: row k0 of W (nothing pending); epoch
    // epoch0 + 1 + i carries the candidates of pair step i
    {
        double best = -1.0;
        T r = zero<T>();
        if (own && x > k0) {
            r = a.W[(int64_t)k0 * n + x];
            xb_st<XM>(&PV[x], r);
            const double m = (double)mag2(r);
            if (m == m) best = m;
        }
        Vr.template put<0>(r);
        wp_publish(best, lg, x, epoch0 + 1, rank, co.wrec);
    }

    int quit = 0;   // 1: the panel is done (or the matrix); 2: leave at once (zero pivot, failed poll)
    bp_sfor<0, NH>([&](auto hc) {
        constexpr int H = decltype(hc)::value;
        if (quit) return;
        const int iend = min(np, 16 * H + 16);
        for (int i = 16 * H; i < iend; ++i) {
            const int il = i - 16 * H;
            const int k = k0 + 2 * i;
            const int e = epoch0 + 1 + i;
            const bool has_next = i + 1 < np;       // row k+2 is this panel's next pivot row
            const int sa = wp_q(q0, q1, k + 1 - k0);
            const int sb = has_next ? wp_q(q0, q1, k + 2 - k0) : -1;
            const int sbs = sb >= 0 ? sb : sa;

            int p = k + 1, sp = sa;
            bool wf = false;
            if (!wp_poll(co.wrec, e, G, a.spin, &st->err, p, sp, wf)) { quit = 2; return; }
            if (!wf) { p = k + 1; sp = sa; }                   // no finite candidate (NaN row)
            const bool sw = p != k + 1;
            // the interchange k+1 <-> p as seen from the own slot
            if (x == sp) lg = k + 1;
            else if (sw && x == sa) lg = p;
            const int y1 = sp;
            const int y2 = (p == k + 2) ? sa : sb;
            const int y2s = has_next ? y2 : y1;

            // ---- every load that depends on the pivot, in one run: the
            //      entries U_j, V_j at the winner's slot and at the
            //      pre-interchange slots of rows k+1 and k+2 (lane 2j + v),
            //      then rows k+1 and k+2 of W at the own slots and entry
            //      (k+1, k+2); each from a valid address, masked after
            const int l = lane < 2 * NB ? lane : 2 * NB - 1;
            const int jl = l >> 1;
            const bool isv = l & 1;
            const T* src = (isv ? PV : PU) + (int64_t)jl * n;
            const T g0 = xb_ld<XM>(&src[sp]);
            const T g1 = xb_ld<XM>(&src[sa]);
            const T g2 = xb_ld<XM>(&src[sbs]);
            // (the column of entry (k+1, k+2) passes through a vector register:
            // from a uniform address the compiler would move the value to a
            // scalar register at once, and the wait for the entries above
            // would then hold for this HBM read too)
            int y2v;
            asm volatile("v_mov_b32 %0, %1" : "=v"(y2v) : "v"(y2s));
            const T w1r = a.W[wraw_ix(n, y1, xs)];
            const T w2r = a.W[wraw_ix(n, y2s, xs)];
            const T w12r = a.W[wraw_ix(n, y1, y2v)];
            const bool ok = isv ? jl <= i : jl < i;             // U_j for j < i, V_j for j <= i
            __builtin_amdgcn_wave_barrier();
            if (lane < 2 * NB) {
                Bx[0][lane] = psel(ok, g0, zero<T>());
                Bx[1][lane] = psel(ok, g1, zero<T>());
                Bx[2][lane] = psel(ok && sb >= 0, g2, zero<T>());
            }
            __builtin_amdgcn_wave_barrier();
            const T* b1 = Bx[0];                               // the pivot's entries (row k+1 when no winner: sp = sa)
            const T piv = b1[2 * i + 1];                       // V_i there
            if (is_zero(piv)) {                                // same decision in every participant
                if (keeper) st->dead = 1;
                quit = 2;
                return;
            }
            if (keeper) accumulate(lm, le, ph_re, ph_im, piv, sw);
            if (sw) {
                wp_qset(q0, q1, k + 1 - k0, sp);
                if ((unsigned)(p - k0) < (unsigned)(2 * NB + 2)) wp_qset(q0, q1, p - k0, sa);
            }
            const T* b2 = Bx[(p == k + 2) ? 1 : 2];
            if (k + 2 >= n) { quit = 1; return; }              // last pair of the matrix

            const T inv = recip(piv);
            // corrections of the pairs j < i, in chunks, as in bp_panel
            T r1 = zero<T>(), rl = zero<T>(), rk = zero<T>();
            T r1m = zero<T>(), rlm = zero<T>(), rkm = zero<T>();
            const BCPair<T>* P1 = reinterpret_cast<const BCPair<T>*>(b1);
            const BCPair<T>* P2 = reinterpret_cast<const BCPair<T>*>(b2);
            bp_sfor<0, NB / CH>([&](auto cc) {
                constexpr int C = decltype(cc)::value;
                if (C * CH >= i) return;
                BCPair<T> c1[CH], c2[CH];
                bp_sfor<0, CH>([&](auto qc) {
                    constexpr int Q = decltype(qc)::value;
                    c1[Q] = P1[C * CH + Q];
                    c2[Q] = P2[C * CH + Q];
                });
                bp_sfor<0, CH>([&](auto qc) {
                    constexpr int Q = decltype(qc)::value;
                    constexpr int J = C * CH + Q;
                    const T uj = Ur.template at<J>(), vj = Vr.template at<J>();
                    const T b1u = c1[Q].u, b1v = c1[Q].v, b2u = c2[Q].u, b2v = c2[Q].v;
                    if constexpr (is_cplx<T>::value) {
                        cfma(r1, b1u, vj); cfma(r1m, b1v, uj);
                        cfma(rl, b2u, vj); cfma(rlm, b2v, uj);
                        cfma(rk, b1u, b2v); cfma(rkm, b1v, b2u);
                    } else {
                        r1 = r1 + b1u * vj - b1v * uj;
                        rl = rl + b2u * vj - b2v * uj;
                        rk = rk + b1u * b2v - b1v * b2u;
                    }
                });
            });
            if constexpr (is_cplx<T>::value) {
                r1 = r1 - r1m;
                rl = rl - rlm;
                rk = rk - rkm;
            }
            // every lane stores U_i and V_{i+1} (zero at the slots not
            // active, which are never read back; lanes past n at slot k0):
            // a store under a branch would make the compiler wait at the
            // branch's end for it and for the reads it consumes
            const bool act1 = own && lg >= k + 2;
            T u = zero<T>();
            if (act1) u = (wfix(w1r, y1, xs) + r1) * inv;
            xb_st<XM>(&PU[(int64_t)i * n + xs], u);
            Ur.template set<H>(il, u);
            if (!has_next) { quit = 1; return; }

            // ---- row k+2 (the next pivot row) at the own slots -> V_{i+1}
            const T u_kl = (wfix(w12r, y1, y2s) + rk) * inv;   // U_i at row k+2
            const T v_kl = b2[2 * i + 1];                      // V_i at row k+2
            const T vi = Vr.template get<H>(il);
            const bool act2 = own && lg >= k + 3;
            T r = zero<T>();
            if (act2) r = wfix(w2r, y2, xs) + rl + u_kl * vi - v_kl * u;
            xb_st<XM>(&PV[(int64_t)(i + 1) * n + xs], r);
            const double m = (double)mag2(r);
            const double best = (act2 && m == m) ? m : -1.0;
            if constexpr (H + 1 < NH) {
                if (il == 15) Vr.template put<16 * (H + 1)>(r);
                else Vr.template set<H>(min(il + 1, 15), r);
            } else {
                constexpr int LH = BPRegHalf<double, NB>::template len<H>();
                Vr.template set<H>(min(il + 1, LH - 1), r);    // i + 1 < np <= NB
            }
            wp_publish(best, lg, x, e + 1, rank, co.wrec);
        }
    });
    if (quit == 2) return;

    if (own) a.perm[lg] = x;   // logical -> slot, for the trailing update
    if (keeper) { st->lm = lm; st->le = le; st->ph_re = ph_re; st->ph_im = ph_im; }
}

// Trailing update over upper tiles of the block [lo, n)^2:
// Wn[r][c] = Wo[perm r][perm c] + sum_{j<np} (U_j[s r] V_j[s c] - V_j[s r] U_j[s c]),  r < c,
// with s = perm when the panel vectors are stored by slot (pslot), else identity.
// 256 threads; thread (tx, ty) owns rows r0+ty+16a, cols c0+tx+16b (a, b < 4).
template <typename T>
__global__ void __launch_bounds__(256) bp_trailing(const T* __restrict__ Wo, T* __restrict__ Wn,
                                                    const T* __restrict__ P, const int* __restrict__ perm,
                                                    int64_t n, int lo, int nb, int np, int ntile, int pslot) {
    extern __shared__ unsigned char smem_raw[];
    T* sm = reinterpret_cast<T*>(smem_raw);
    const int ld = np + 1;                       // padded row length
    T* Ur = sm;
    T* Vr = Ur + BP_TILE * ld;
    T* Uc = Vr + BP_TILE * ld;
    T* Vc = Uc + BP_TILE * ld;
    int ti, tj;
    if (!xcd_tile(blockIdx.x, ntile, ti, tj)) return;
    const int r0 = lo + ti * BP_TILE, c0 = lo + tj * BP_TILE;
    for (int e = threadIdx.x; e < BP_TILE * np; e += 256) {
        const int q = e % BP_TILE, j = e / BP_TILE;   // consecutive threads: consecutive rows (coalesced)
        const int r = r0 + q, c = c0 + q;
        const int pr = (r < n) ? (pslot ? perm[r] : r) : 0;
        const int pc = (c < n) ? (pslot ? perm[c] : c) : 0;
        Ur[q * ld + j] = (r < n) ? P[(int64_t)j * n + pr] : zero<T>();
        Vr[q * ld + j] = (r < n) ? P[(int64_t)(nb + j) * n + pr] : zero<T>();
        Uc[q * ld + j] = (c < n) ? P[(int64_t)j * n + pc] : zero<T>();
        Vc[q * ld + j] = (c < n) ? P[(int64_t)(nb + j) * n + pc] : zero<T>();
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
    // stale entries: all sixteen gathers issued before any store
    const int n1 = (int)n - 1;
    int pr[4], pc[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) pr[x] = perm[min(r0 + ty + 16 * x, n1)];
#pragma unroll
    for (int y = 0; y < 4; ++y) pc[y] = perm[min(c0 + tx + 16 * y, n1)];
    T wv[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) wv[x][y] = w_at(Wo, n, pr[x], pc[y]);
#pragma unroll
    for (int x = 0; x < 4; ++x) {
        const int r = r0 + ty + 16 * x;
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const int c = c0 + tx + 16 * y;
            if (r < n && c < n && c > r) Wn[(int64_t)r * n + c] = wv[x][y] + acc[x][y];
        }
    }
}

// FP64 trailing update on the matrix cores: the same sum as bp_trailing, as
// the product  C += A B  (K = 2 nb) per 64 x 64 upper tile, with the two
// terms of a pair adjacent in K (they nearly cancel, and summed apart the
// partial sums grow and lose digits):
//   A[r][2j] = U_j[s r],  A[r][2j+1] = V_j[s r],
//   B[2j][c] = V_j[s c],  B[2j+1][c] = -U_j[s c].  Wave w owns
// tile rows 16 w .. 16 w + 15 and four 16 x 16 accumulators across the
// columns, fed by v_mfma_f64_16x16x4f64 (A operand: lane l holds
// A[l % 16][l / 16], B operand B[l / 16][l % 16], accumulator register e holds
// (l / 16 + 4 e, l % 16)).  Operands come straight from P (small, L2
// resident) with no LDS, so many waves per CU hide the latency of the
// stale-matrix gather; the accumulators start from the gathered entries,
// so W_old is read once and W_new written once.
typedef double bt_d4 __attribute__((ext_vector_type(4)));

template <int NB>
__global__ void __launch_bounds__(256) bp_trailing_mfma(const double* __restrict__ Wo, double* __restrict__ Wn,
                                                        const double* __restrict__ P, const int* __restrict__ perm,
                                                        int64_t n, int lo, int np, int ntile, int pslot) {
    constexpr int K = 2 * NB;
    int ti, tj;
    if (!xcd_tile(blockIdx.x, ntile, ti, tj)) return;
    const int r0 = lo + ti * BP_TILE + 16 * (threadIdx.x >> 6), c0 = lo + tj * BP_TILE;
    const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
    const int n1 = (int)n - 1;
    // order of the loads: perm, then the L2-resident operands, then the stale
    // entries from HBM.  Loads complete in issue order, so the products can
    // start as soon as the operands are in while the gathers are in flight.
    int pr[4], pc[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) pr[e] = perm[min(r0 + lq + 4 * e, n1)];
#pragma unroll
    for (int u = 0; u < 4; ++u) pc[u] = perm[min(c0 + 16 * u + lr, n1)];
    const int pa = perm[min(r0 + lr, n1)];
    const int sa = pslot ? pa : min(r0 + lr, n1);   // slot of this lane's A row
    int sb[4];                                       // slots of its B columns
#pragma unroll
    for (int u = 0; u < 4; ++u) sb[u] = pslot ? pc[u] : min(c0 + 16 * u + lr, n1);
    double av[K / 4], bv[K / 4][4];
#pragma unroll
    for (int g = 0; g < K / 4; ++g) {
        const int q = 4 * g + lq;                  // this lane's K index: pair q / 2
        const int j = q >> 1;
        const bool on = j < np;
        const int qa = (q & 1) ? NB + j : j;       // A row: U_j, V_j
        const int qb = (q & 1) ? j : NB + j;       // B row: V_j, -U_j
        const double sg = (q & 1) ? -1.0 : 1.0;
        const double va = P[(int64_t)qa * n + sa];
        av[g] = on ? va : 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double vb = P[(int64_t)qb * n + sb[u]];
            bv[g][u] = on ? sg * vb : 0.0;
        }
    }
    double wv[4][4];   // stale entries (r0 + lq + 4e, c0 + 16u + lr)
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int e = 0; e < 4; ++e) wv[u][e] = w_at(Wo, n, pr[e], pc[u]);   // clamped indices are valid
    bt_d4 acc[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[u] = bt_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int g = 0; g < K / 4; ++g) {
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[g], bv[g][u], acc[u], 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int r = r0 + lq + 4 * e, c = c0 + 16 * u + lr;
            if (c < n && c > r) Wn[(int64_t)r * n + c] = wv[u][e] + acc[u][e];
        }
    }
}

// complex128 form of bp_trailing_mfma: the same tiles, lane layout and K
// order, with each complex product as four real ones,
//   Re C += Re A Re B - Im A Im B,   Im C += Re A Im B + Im A Re B.
template <int NB>
__global__ void __launch_bounds__(256) bp_trailing_mfma_c(const c128* __restrict__ Wo, c128* __restrict__ Wn,
                                                          const c128* __restrict__ P, const int* __restrict__ perm,
                                                          int64_t n, int lo, int np, int ntile, int pslot) {
    constexpr int K = 2 * NB;
    int ti, tj;
    if (!xcd_tile(blockIdx.x, ntile, ti, tj)) return;
    const int r0 = lo + ti * BP_TILE + 16 * (threadIdx.x >> 6), c0 = lo + tj * BP_TILE;
    const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
    const int n1 = (int)n - 1;
    int pr[4], pc[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) pr[e] = perm[min(r0 + lq + 4 * e, n1)];
#pragma unroll
    for (int u = 0; u < 4; ++u) pc[u] = perm[min(c0 + 16 * u + lr, n1)];
    const int pa = perm[min(r0 + lr, n1)];
    const int sa = pslot ? pa : min(r0 + lr, n1);
    int sb[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) sb[u] = pslot ? pc[u] : min(c0 + 16 * u + lr, n1);
    bt_d4 are[4], aim[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) { are[u] = bt_d4{0.0, 0.0, 0.0, 0.0}; aim[u] = bt_d4{0.0, 0.0, 0.0, 0.0}; }
#pragma unroll
    for (int g = 0; g < K / 4; ++g) {
        const int q = 4 * g + lq;                  // pair q / 2
        const int j = q >> 1;
        const bool on = j < np;
        const int qa = (q & 1) ? NB + j : j;
        const int qb = (q & 1) ? j : NB + j;
        const double sg = (q & 1) ? -1.0 : 1.0;
        const c128 va = P[(int64_t)qa * n + sa];
        const double ar = on ? va.re : 0.0, ai = on ? va.im : 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const c128 vb = P[(int64_t)qb * n + sb[u]];
            const double br = on ? sg * vb.re : 0.0, bi = on ? sg * vb.im : 0.0;
            are[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, are[u], 0, 0, 0);
            are[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, bi, are[u], 0, 0, 0);
            aim[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, aim[u], 0, 0, 0);
            aim[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, aim[u], 0, 0, 0);
        }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int r = r0 + lq + 4 * e, c = c0 + 16 * u + lr;
            const c128 w = w_at(Wo, n, pr[e], pc[u]);   // clamped indices are valid
            if (c < n && c > r) Wn[(int64_t)r * n + c] = c128{w.re + are[u][e], w.im + aim[u][e]};
        }
    }
}

// ---------------------------------------------------------------------------
// Trailing updates with the B operand staged in LDS.  The four waves of a
// 64 x 64 tile share its K x 64 B operand: here it is read once per tile
// from L2 into LDS (16 values per thread) instead of once per wave, and each
// wave keeps only its own A rows in registers.  The stale-entry gathers are
// issued after the staging stores and before the barrier, so they overlap
// the barrier and the products.  Same products,
// same K order and same lane layout as the register forms above.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void tr_mma(double a, double b, bt_d4& re, bt_d4& im) {
    (void)im;
    re = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, re, 0, 0, 0);
}
__device__ __forceinline__ void tr_mma(c128 a, c128 b, bt_d4& re, bt_d4& im) {
    re = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.re, re, 0, 0, 0);
    re = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.im, b.im, re, 0, 0, 0);
    im = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.im, im, 0, 0, 0);
    im = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, b.re, im, 0, 0, 0);
}
__device__ __forceinline__ double tr_out(double w, const bt_d4& re, const bt_d4& im, int e) {
    (void)im;
    return w + re[e];
}
__device__ __forceinline__ c128 tr_out(c128 w, const bt_d4& re, const bt_d4& im, int e) {
    return c128{w.re + re[e], w.im + im[e]};
}

// Factorization: Wn[r][c] = Wo[q r][q c] + sum_j (U_j V_j' - V_j U_j') over
// the upper tiles of [lo, n)^2 (as bp_trailing_mfma / bp_trailing_mfma_c).
template <typename T, int NB>
__global__ void __launch_bounds__(256) bp_trailing_staged(const T* __restrict__ Wo, T* __restrict__ Wn,
                                                          const T* __restrict__ P, const int* __restrict__ perm,
                                                          int64_t n, int lo, int np, int ntile, int pslot) {
    constexpr int K = 2 * NB;
    __shared__ T Bs[K][BP_TILE];
    int ti, tj;
    if (!xcd_tile(blockIdx.x, ntile, ti, tj)) return;
    const int r0 = lo + ti * BP_TILE + 16 * (threadIdx.x >> 6), c0 = lo + tj * BP_TILE;
    const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
    const int col = threadIdx.x & 63, kq = threadIdx.x >> 6;   // staging: column, first K row
    const int n1 = (int)n - 1;
    int pr[4], pc[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) pr[e] = perm[min(r0 + lq + 4 * e, n1)];
#pragma unroll
    for (int u = 0; u < 4; ++u) pc[u] = perm[min(c0 + 16 * u + lr, n1)];
    const int pa = perm[min(r0 + lr, n1)];
    const int pcs = perm[min(c0 + col, n1)];
    const int sa = pslot ? pa : min(r0 + lr, n1);
    const int sbs = pslot ? pcs : min(c0 + col, n1);
    T sv[K / 4];   // B[kq + 4 i][col]: V_j, then -U_j
#pragma unroll
    for (int i = 0; i < K / 4; ++i) {
        const int q = kq + 4 * i, j = q >> 1;
        const T v = P[(int64_t)((q & 1) ? j : NB + j) * n + sbs];
        sv[i] = (j < np) ? ((q & 1) ? -v : v) : zero<T>();
    }
    T av[K / 4];   // A[r][4 g + lq]: U_j, V_j
#pragma unroll
    for (int g = 0; g < K / 4; ++g) {
        const int q = 4 * g + lq, j = q >> 1;
        const T v = P[(int64_t)((q & 1) ? NB + j : j) * n + sa];
        av[g] = (j < np) ? v : zero<T>();
    }
#pragma unroll
    for (int i = 0; i < K / 4; ++i) Bs[kq + 4 * i][col] = sv[i];
    T wv[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int e = 0; e < 4; ++e) wv[u][e] = w_at(Wo, n, pr[e], pc[u]);
    __syncthreads();
    bt_d4 re[4], im[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) { re[u] = bt_d4{0.0, 0.0, 0.0, 0.0}; im[u] = bt_d4{0.0, 0.0, 0.0, 0.0}; }
#pragma unroll
    for (int g = 0; g < K / 4; ++g) {
#pragma unroll
        for (int u = 0; u < 4; ++u) tr_mma(av[g], Bs[4 * g + lq][16 * u + lr], re[u], im[u]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int r = r0 + lq + 4 * e, c = c0 + 16 * u + lr;
            if (c < n && c > r) Wn[(int64_t)r * n + c] = tr_out(wv[u][e], re[u], im[u], e);
        }
    }
}

// ---------------------------------------------------------------------------
// Split-precision trailing update (PFH_METHOD_MIXED, float64): the product
// sum_j (U_j V_j' - V_j U_j') on the bf16 matrix cores with exact products,
// recombined in FP64.  The panel stays FP64.
//
// bp_slice writes, for every slot x of the panel vectors and each half h of
// the K = 64 values (pairs 0-15, 16-31: one MFMA's K range each), a scale
// E[x][h] (the exponent of the largest of those values, so every one is
// below 2^E in magnitude) and S signed 7-bit digits of each value:
//   v = sum_t d_t 2^(E - 7 (t + 1)) + rho,  |d_t| <= 127,  |rho| < 2^(E - 7 S).
// The digits are small integers, exact in bf16, stored D[t][x][q] (q = 2j:
// U_j, q = 2j + 1: V_j; 64 per slot).  (One scale per slot over all 64
// values was measured first: on the N = 8192 precision matrix, whose panel
// vectors span six decades, its |dlog| was 1.9e-9 at S = 8.)
//
// bp_trailing_mixed forms, per 16 x 16 block of a 64 x 64 upper tile, per
// half h and per digit order d = ta + tb < S,
//   acc_{h,d} = sum_{ta + tb = d} Dta(A)_h Dtb(B)_h
// with v_mfma_f32_16x16x32_bf16.  Every product of two digits is exact in
// fp32 and every partial sum is an integer of magnitude at most
// S * 32 * 127^2 < 2^24 (S <= 16), so acc_{h,d} is exact whatever the order
// of the additions.  Then
//   C = sum_h 2^(E[r][h] + E[c][h]) sum_d 2^(-7 (d + 2)) acc_{h,d}
// in FP64.  The digit products left out (ta + tb >= S) and the remainders
// bound the error of an entry by about 32 (S + 1) 2^(-7 S) 2^(E[r][h] +
// E[c][h]) per half: relative to the largest operand entries of that half,
// 2^-44 at S = 7, 2^-51 at S = 8, 2^-58 at S = 9 (FP64 products: 2^-53).
// The B operand is the A-order digits with each pair (U_j, V_j) turned into
// (V_j, -U_j): a swap of the two bf16 halves of a dword and a sign flip.
// ---------------------------------------------------------------------------
typedef __bf16 mx_bf8 __attribute__((ext_vector_type(8)));
typedef float mx_f4 __attribute__((ext_vector_type(4)));
constexpr int MX_K = 64;   // 2 nb at nb = 32

// one wave per slot x in [x0, x1): lane q holds value q (A order)
template <int S>
__global__ void __launch_bounds__(256) bp_slice(const double* __restrict__ P, int64_t n, int nb, int np, int x0,
                                                int x1, __bf16* __restrict__ D, int* __restrict__ E) {
    const int x = x0 + (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (x >= x1) return;
    const int q = threadIdx.x & 63, j = q >> 1;
    const double a = (j < np) ? P[(int64_t)((q & 1) ? nb + j : j) * n + x] : 0.0;
    double m = fabs(a);   // the largest of this lane's half (32 lanes)
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    int e = 0;
    if (m > 0.0 && m <= 1.7976931348623157e308) (void)frexp(m, &e);   // |a| < 2^e; NaN / inf keep 0 and propagate
    double v = ldexp(a, -e);                                          // |v| < 1
#pragma unroll
    for (int t = 0; t < S; ++t) {
        v *= 128.0;
        const double dgt = trunc(v);                                  // |dgt| <= 127, exact
        v -= dgt;                                                     // exact
        D[((int64_t)t * n + x) * MX_K + q] = (__bf16)(float)dgt;
    }
    if ((q & 31) == 0) E[2 * x + (q >> 5)] = e;
}

// (U, V) pairs of a chunk of 8 A-order digits -> (V, -U): B-order digits
__device__ __forceinline__ mx_bf8 mx_to_b(mx_bf8 a) {
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    u4 w;
    __builtin_memcpy(&w, &a, 16);
#pragma unroll
    for (int i = 0; i < 4; ++i) w[i] = ((w[i] >> 16) | (w[i] << 16)) ^ 0x80000000u;   // low half V, high half -U
    mx_bf8 b;
    __builtin_memcpy(&b, &w, 16);
    return b;
}

template <int S>
__global__ void __launch_bounds__(256) bp_trailing_mixed(const double* __restrict__ Wo, double* __restrict__ Wn,
                                                         const __bf16* __restrict__ D, const int* __restrict__ E,
                                                         const int* __restrict__ perm, int64_t n, int lo,
                                                         int ntile, int pslot) {
    int ti, tj;
    if (!xcd_tile(blockIdx.x, ntile, ti, tj)) return;
    const int r0 = lo + ti * BP_TILE + 16 * (threadIdx.x >> 6), c0 = lo + tj * BP_TILE;
    const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
    const int n1 = (int)n - 1;
    // output rows of this lane: r0 + 4 lq + e; column c0 + 16 u + lr
    int pr[4], pc[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) pr[e] = perm[min(r0 + 4 * lq + e, n1)];
#pragma unroll
    for (int u = 0; u < 4; ++u) pc[u] = perm[min(c0 + 16 * u + lr, n1)];
    const int sa = pslot ? perm[min(r0 + lr, n1)] : min(r0 + lr, n1);   // slot of this lane's A row
    int sb[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) sb[u] = pslot ? pc[u] : min(c0 + 16 * u + lr, n1);
    int er[4][2], ec[4][2];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int x = pslot ? pr[e] : min(r0 + 4 * lq + e, n1);
        er[e][0] = E[2 * x];
        er[e][1] = E[2 * x + 1];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) { ec[u][0] = E[2 * sb[u]]; ec[u][1] = E[2 * sb[u] + 1]; }
    mx_bf8 ad[S][2];
#pragma unroll
    for (int t = 0; t < S; ++t)
#pragma unroll
        for (int h = 0; h < 2; ++h)
            ad[t][h] = *reinterpret_cast<const mx_bf8*>(D + ((int64_t)t * n + sa) * MX_K + 32 * h + 8 * lq);
    double wv[4][4];   // stale entries (r0 + 4 lq + e, c0 + 16 u + lr)
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int e = 0; e < 4; ++e) wv[u][e] = w_at(Wo, n, pr[e], pc[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        mx_bf8 bd[S][2];
#pragma unroll
        for (int t = 0; t < S; ++t)
#pragma unroll
            for (int h = 0; h < 2; ++h)
                bd[t][h] = mx_to_b(*reinterpret_cast<const mx_bf8*>(D + ((int64_t)t * n + sb[u]) * MX_K + 32 * h + 8 * lq));
        double res[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
#pragma unroll
        for (int d = 0; d < S; ++d) {
            mx_f4 acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ta = 0; ta <= d; ++ta) {
                acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ad[ta][0], bd[d - ta][0], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ad[ta][1], bd[d - ta][1], acc1, 0, 0, 0);
            }
            const double sc = ldexp(1.0, -7 * (d + 2));
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                res[0][e] = fma((double)acc0[e], sc, res[0][e]);
                res[1][e] = fma((double)acc1[e], sc, res[1][e]);
            }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int r = r0 + 4 * lq + e, c = c0 + 16 * u + lr;
            const double cv = ldexp(res[0][e], er[e][0] + ec[u][0]) + ldexp(res[1][e], er[e][1] + ec[u][1]);
            if (c < n && c > r) Wn[(int64_t)r * n + c] = wv[u][e] + cv;
        }
    }
}

// Digit count of the split-precision update: PFH_TUNE_MIXED_SLICES (5 .. 9)
// or PFH_MIXED_SLICES, else the built-in choice.
static int mixed_slices() {
    const int tv = tune_value(PFH_TUNE_MIXED_SLICES);
    if (tv >= 5 && tv <= 9) return tv;
    const char* e = std::getenv("PFH_MIXED_SLICES");
    if (e && e[0] >= '5' && e[0] <= '9' && e[1] == 0) return e[0] - '0';
    return PFH_MIXED_SLICES_DEFAULT;
}

template <int S>
static void launch_mixed_trailing(const double* Wo, double* Wn, const double* P, __bf16* D, int* E,
                                  const int* perm, int64_t n, int k0, int lo, int nb, int np, int ntile, int tiles,
                                  int pslot, hipStream_t s) {
    const int nslot = (int)(n - k0);
    hipLaunchKernelGGL(bp_slice<S>, dim3((unsigned)((nslot + 3) / 4)), dim3(256), 0, s, P, n, nb, np, k0, (int)n, D, E);
    hipLaunchKernelGGL(bp_trailing_mixed<S>, dim3(tiles), dim3(256), 0, s, Wo, Wn, D, E, perm, n, lo, ntile, pslot);
}

// Which form runs, by measurement (ms, staged / register): the f64
// factorization keeps the register form (N = 4096: 14.2 / 13.9, N = 8192:
// 35.6 / 34.8); its c128 form (N = 4096: 21.7 / 22.3) and the block sweep
// (f64 N = 2048: 8.0 / 8.6, N = 4096: 18.5 / 22.5; c128 N = 4096: 30.5 /
// 32.1) use the staged form.  PFH_TRAIL_STAGED=1 or 0 forces one form
// everywhere (read per call).
static bool trail_staged(int key, bool dflt) {
    const int tv = tune_value(key);
    if (tv >= 0) return tv == 1;
    const char* e = std::getenv("PFH_TRAIL_STAGED");
    if (e && e[0] == '0') return false;
    if (e && e[0] == '1') return true;
    return dflt;
}

template <typename T>
__global__ void bp_finish(const BPState<T>* st, double* log_out, T* phase_out) {
    if (threadIdx.x != 0) return;
    if (st->err) {
        atomicOr(&g_panel_timeout, 1);
        *log_out = __builtin_nan("");
        *phase_out = nan_of<T>();
        return;
    }
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
    size_t w0, w1, p, r, perm, pv, pi, st, claim, rec, wrec, cand, mxd, mxe, total;
};

template <typename T>
static BPLayout<T> bp_layout(int64_t n, int nb, bool mixed = false) {
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
    L.claim = o; o = al(o + sizeof(int) * (size_t)(n / (2 * nb) + 2));
    L.rec = o; o = al(o + sizeof(unsigned long long) * 3 * BC_MAXG * 2);
    L.wrec = o; o = al(o + sizeof(unsigned long long) * 3 * WP_MAXG * 2);
    L.cand = o; o = al(o + sizeof(T) * 3 * BC_MAXG * 2 * nb);
    L.mxd = L.mxe = o;
    if (mixed) {   // split-precision digits (up to 8 per value) and slot scales
        L.mxd = o; o = al(o + sizeof(__bf16) * 9 * (size_t)n * MX_K);
        L.mxe = o; o = al(o + sizeof(int) * 2 * (size_t)n);
    }
    L.total = o;
    return L;
}

// Pair steps per panel of the blocked factorization: 32 (float64, and the
// real / complex64 types), 16 for complex128, whose panel vectors would not
// fit the panel kernel's LDS at 32; PFH_TUNE_BLOCKED_NB (float64: 16, 24 or
// 32) and PFH_TUNE_BLOCKED_NB_C128 (8 or 16) choose another instantiated
// width.  The split-precision method keeps 32 (its digit products take
// K = 64), and the gradient's block sweep keeps 32 / 16 (grad_panel_width).
int blocked_panel_width(int dtype) {
    if (dtype == DT_C128) {
        const int tv = tune_value(PFH_TUNE_BLOCKED_NB_C128);
        return (tv == 8 || tv == 16) ? tv : 16;
    }
    if (dtype == DT_F64) {
        const int tv = tune_value(PFH_TUNE_BLOCKED_NB);
        return (tv == 16 || tv == 24 || tv == 32) ? tv : 32;
    }
    return 32;
}
// Pair steps per group of the blocked block sweep: complex128 16; float64
// 16 (PFH_TUNE_GRAD_NB may set 32: measured at n = 2048, 16 pairs 3.46 ms
// against 3.55 at 32 -- shorter pair steps and a 32 x 32 inversion per group
// outweigh the twice as many trailing passes); other types 32.
static int grad_panel_width(int dtype) {
    if (dtype == DT_C128) return 16;
    if (dtype == DT_F64) return tune_value(PFH_TUNE_GRAD_NB) == 32 ? 32 : 16;
    return 32;
}

// The panel kernel runs on a device with at least BC_MAXG CUs (so that a
// panel's blocks, one per CU, are resident together), unless
// PFH_BLOCKED_COOP=0, and while the panel's columns fit BC_MAXG blocks of
// BC_T slots.
static bool coop_enabled() {
    static DevCache cache;   // 1 enabled, 2 not
    return cache.get([] {
        const char* e = std::getenv("PFH_BLOCKED_COOP");
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            cus = 0;
        return (cus >= BC_MAXG && !(e && e[0] == '0')) ? 1 : 2;
    }) == 1;
}

// Largest panel grid that the XCD-local form takes: one XCD's CUs (the
// blocks take a CU each), 0 when the form is off.  On by default for every
// type.  Measured with the own-slot panel vectors in registers (ms,
// XCD-local / device-wide): f64 N = 4096 9.20 / 11.28, N = 8192 24.5 /
// 28.6, gradient N = 2048 5.82 / 6.59; c128 N = 4096 12.4 / 14.4.  (With
// the panel vectors in LDS, real types had lost: f64 N = 4096 12.1 / 9.9,
// the strided reads of W's columns, one cache line per entry, all going
// through one XCD's L2.)  The panel form: 2 one participant per wave
// (bp_wpanel, XCD-local; the default), 1 one per block (bp_panel,
// XCD-local), 0 one per block device-wide; PFH_PANEL_XCD=0/1/2 or
// PFH_TUNE_PANEL_XCD set it for every type, read per call.  A form whose
// participants do not fit leaves the panel to the next one down.
static int panel_form() {
    const int tv = tune_value(PFH_TUNE_PANEL_XCD);
    int f = 2;
    if (tv >= 0) {
        f = tv;
    } else {
        const char* e = std::getenv("PFH_PANEL_XCD");
        if (e && e[0] >= '0' && e[0] <= '2') f = e[0] - '0';
    }
    return device_cu_count() % 8 == 0 ? f : 0;
}
static int panel_xcd_max_grid() { return panel_form() >= 1 ? device_cu_count() / 8 : 0; }

// Waves of bp_wpanel<T, NB> that XCD 0 holds at once (its CUs times the
// kernel's workgroups per CU), 0 when the query fails.
template <typename T, int NB>
static int wpanel_capacity() {
    static DevCache cache;   // capacity + 1
    return cache.get([] {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bp_wpanel<T, NB>, 64, 0) != hipSuccess)
            per_cu = 0;
        return per_cu * (device_cu_count() / 8) + 1;
    }) - 1;
}

template <typename T, int NB, bool LOCAL>
static hipError_t launch_panel_form(const BPArgs<T>& a, const BPCoop<T>& co, int k0, int np, int epoch0, int G,
                                    hipStream_t s) {
    // A plain launch: the G working blocks (at most 64; their registers
    // allow two per CU) are resident together once the stream's earlier work
    // has drained.  A block that cannot become resident (the CUs held by
    // other work for seconds) ends in the barrier's bounded spin, which sets
    // the error flag: the result is NaN and pfh_stream_status reports
    // PFH_ETIMEOUT; nothing hangs.  (A cooperative launch, which asks the
    // runtime for co-residency, was measured slower, 13.5-20.5 ms against
    // 12.6 ms at N = 4096, and was removed.)
    const int grid = LOCAL ? 8 * G : G;
    hipLaunchKernelGGL((bp_panel<T, NB, LOCAL>), dim3(grid), dim3(BC_T), 0, s, a, co, k0, np, epoch0);
    return hipGetLastError();
}

// One panel launch over the columns [k0, n): the per-wave form while its
// participants fit XCD 0 together, else the per-block forms.  Each form
// keeps its own epoch count (its records' parities follow it), advanced by
// the panel's np.
template <typename T, int NB>
static hipError_t launch_panel(const BPArgs<T>& a, const BPCoop<T>& co, int k0, int np, int& epoch, int& wepoch,
                               hipStream_t s) {
    const int64_t m = a.n - k0;
    const int Gw = (int)((m + 63) / 64);
    if (panel_form() == 2 && Gw <= WP_MAXG && Gw <= wpanel_capacity<T, NB>()) {
        hipLaunchKernelGGL((bp_wpanel<T, NB>), dim3(8 * Gw), dim3(64), 0, s, a, co, k0, np, wepoch);
        wepoch += np;
        return hipGetLastError();
    }
    const int G = (int)((m + BC_T - 1) / BC_T);
    const hipError_t e = G <= panel_xcd_max_grid() ? launch_panel_form<T, NB, true>(a, co, k0, np, epoch, G, s)
                                                   : launch_panel_form<T, NB, false>(a, co, k0, np, epoch, G, s);
    epoch += np;
    return e;
}

int blocked_take_status(hipStream_t s, int* timed_out) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return (int)e;
    int v = 0;
    e = hipMemcpyFromSymbol(&v, HIP_SYMBOL(g_panel_timeout), sizeof(int), 0, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return (int)e;
    if (v) {
        const int zero = 0;
        e = hipMemcpyToSymbol(HIP_SYMBOL(g_panel_timeout), &zero, sizeof(int), 0, hipMemcpyHostToDevice);
        if (e != hipSuccess) return (int)e;
    }
    *timed_out = v;
    return 0;
}

size_t blocked_workspace_bytes(int dtype, int64_t n, bool mixed) {
    const int nb = (mixed && dtype == DT_F64) ? 32 : blocked_panel_width(dtype);
    switch (dtype) {
        case DT_F32: return bp_layout<float>(n, nb).total;
        case DT_F64: return bp_layout<double>(n, nb, mixed && nb == 32).total;
        case DT_C64: return bp_layout<c64>(n, nb).total;
        case DT_C128: return bp_layout<c128>(n, nb).total;
    }
    return 0;
}

template <typename T>
int blocked_slog(const T* A, int64_t n, int nb, bool mixed, void* ws, double* log_out, T* phase_out,
                 hipStream_t s) {
    if (nb <= 0 || nb % 8 != 0) return (int)hipErrorInvalidValue;   // bp_row reads P in batches of 8 pairs
    // the split-precision trailing update serves float64 at nb = 32; other
    // types run the FP64 (or their own precision's) update under MIXED
    const bool mx = mixed && std::is_same<T, double>::value && nb == 32;
    const BPLayout<T> L = bp_layout<T>(n, nb, mx);
    char* base = (char*)ws;
    T* W[2] = {(T*)(base + L.w0), (T*)(base + L.w1)};
    BPArgs<T> a{};
    a.P = (T*)(base + L.p);
    a.R = (T*)(base + L.r);
    a.perm = (int*)(base + L.perm);
    a.part_v = (double*)(base + L.pv);
    a.part_i = (int*)(base + L.pi);
    a.st = (BPState<T>*)(base + L.st);
    a.claim = (int*)(base + L.claim);
    a.n = n;
    a.nb = nb;
    a.spin = tune_value(PFH_TUNE_PANEL_SPIN) >= 0 ? tune_value(PFH_TUNE_PANEL_SPIN) : (1 << 22);
    BPCoop<T> co{};
    co.rec = (unsigned long long*)(base + L.rec);
    co.wrec = (unsigned long long*)(base + L.wrec);
    co.cand = (T*)(base + L.cand);
    const bool coop = coop_enabled() &&
                      (std::is_same<T, c128>::value ? (nb == 8 || nb == 16)
                       : std::is_same<T, double>::value ? (nb == 16 || nb == 24 || nb == 32) : nb == 32);
    hipError_t e = hipMemcpyAsync(W[0], A, sizeof(T) * n * n, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return (int)e;
    a.W = W[0];
    hipLaunchKernelGGL(bp_init<T>, dim3(64), dim3(256), 0, s, a, co.rec, co.wrec);
    int cur = 0;
    int epoch = 0, wepoch = 0;   // epochs passed so far: per-block and per-wave panels
    auto blocks = [&](int64_t c_lo) { return (unsigned)((n - c_lo + BP_THREADS - 1) / BP_THREADS); };
    for (int64_t k0 = 0; k0 + 1 < n; k0 += 2 * nb) {
        const int np = (int)((n - k0) / 2 < nb ? (n - k0) / 2 : nb);
        a.W = W[cur];
        const int G = (int)((n - k0 + BC_T - 1) / BC_T);
        const bool panel_coop = coop && G <= BC_MAXG;
        if (panel_coop) {
            if constexpr (std::is_same<T, c128>::value) {
                e = (nb == 8) ? launch_panel<T, 8>(a, co, (int)k0, np, epoch, wepoch, s)
                              : launch_panel<T, 16>(a, co, (int)k0, np, epoch, wepoch, s);
            } else if constexpr (std::is_same<T, double>::value) {
                e = (nb == 16) ? launch_panel<T, 16>(a, co, (int)k0, np, epoch, wepoch, s)
                  : (nb == 24) ? launch_panel<T, 24>(a, co, (int)k0, np, epoch, wepoch, s)
                               : launch_panel<T, 32>(a, co, (int)k0, np, epoch, wepoch, s);
            } else {
                e = launch_panel<T, 32>(a, co, (int)k0, np, epoch, wepoch, s);
            }
            if (e != hipSuccess) return (int)e;
        } else {
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
        }
        const int64_t lo = k0 + 2 * np;
        if (lo + 1 < n) {
            const int ntile = (int)((n - lo + BP_TILE - 1) / BP_TILE);
            const int tiles = xcd_tile_blocks(ntile);
            const size_t smem = sizeof(T) * 4 * BP_TILE * (np + 1);
            if (smem > 65536) {
                e = hipFuncSetAttribute((const void*)bp_trailing<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)smem);
                if (e != hipSuccess) return (int)e;
            }
            if constexpr (std::is_same<T, double>::value) {
                if (mx) {
                    __bf16* D = (__bf16*)(base + L.mxd);
                    int* E = (int*)(base + L.mxe);
                    const int pslot = panel_coop ? 1 : 0;
                    switch (mixed_slices()) {
                        case 5: launch_mixed_trailing<5>(W[cur], W[cur ^ 1], a.P, D, E, a.perm, n, (int)k0, (int)lo, nb, np, ntile, tiles, pslot, s); break;
                        case 6: launch_mixed_trailing<6>(W[cur], W[cur ^ 1], a.P, D, E, a.perm, n, (int)k0, (int)lo, nb, np, ntile, tiles, pslot, s); break;
                        case 7: launch_mixed_trailing<7>(W[cur], W[cur ^ 1], a.P, D, E, a.perm, n, (int)k0, (int)lo, nb, np, ntile, tiles, pslot, s); break;
                        case 8: launch_mixed_trailing<8>(W[cur], W[cur ^ 1], a.P, D, E, a.perm, n, (int)k0, (int)lo, nb, np, ntile, tiles, pslot, s); break;
                        default: launch_mixed_trailing<9>(W[cur], W[cur ^ 1], a.P, D, E, a.perm, n, (int)k0, (int)lo, nb, np, ntile, tiles, pslot, s); break;
                    }
                } else if (nb == 32 && trail_staged(PFH_TUNE_TRAIL_F64, false)) {
                    hipLaunchKernelGGL((bp_trailing_staged<double, 32>), dim3(tiles), dim3(256), 0, s, W[cur],
                                       W[cur ^ 1], a.P, a.perm, n, (int)lo, np, ntile, panel_coop ? 1 : 0);
                } else if (nb == 32) {
                    hipLaunchKernelGGL(bp_trailing_mfma<32>, dim3(tiles), dim3(256), 0, s, W[cur], W[cur ^ 1], a.P,
                                       a.perm, n, (int)lo, np, ntile, panel_coop ? 1 : 0);
                } else if (nb == 24) {
                    hipLaunchKernelGGL(bp_trailing_mfma<24>, dim3(tiles), dim3(256), 0, s, W[cur], W[cur ^ 1], a.P,
                                       a.perm, n, (int)lo, np, ntile, panel_coop ? 1 : 0);
                } else if (nb == 16) {
                    hipLaunchKernelGGL(bp_trailing_mfma<16>, dim3(tiles), dim3(256), 0, s, W[cur], W[cur ^ 1], a.P,
                                       a.perm, n, (int)lo, np, ntile, panel_coop ? 1 : 0);
                } else {
                    hipLaunchKernelGGL(bp_trailing<T>, dim3(tiles), dim3(256), smem, s, W[cur], W[cur ^ 1], a.P,
                                       a.perm, n, (int)lo, nb, np, ntile, panel_coop ? 1 : 0);
                }
            } else if constexpr (std::is_same<T, c128>::value) {
                if (nb == 16 && trail_staged(PFH_TUNE_TRAIL_C128, true)) {
                    hipLaunchKernelGGL((bp_trailing_staged<c128, 16>), dim3(tiles), dim3(256), 0, s, W[cur],
                                       W[cur ^ 1], a.P, a.perm, n, (int)lo, np, ntile, panel_coop ? 1 : 0);
                } else if (nb == 8) {
                    hipLaunchKernelGGL((bp_trailing_staged<c128, 8>), dim3(tiles), dim3(256), 0, s, W[cur],
                                       W[cur ^ 1], a.P, a.perm, n, (int)lo, np, ntile, panel_coop ? 1 : 0);
                } else if (nb == 16) {
                    hipLaunchKernelGGL(bp_trailing_mfma_c<16>, dim3(tiles), dim3(256), 0, s, W[cur], W[cur ^ 1],
                                       a.P, a.perm, n, (int)lo, np, ntile, panel_coop ? 1 : 0);
                } else {
                    hipLaunchKernelGGL(bp_trailing<T>, dim3(tiles), dim3(256), smem, s, W[cur], W[cur ^ 1], a.P,
                                       a.perm, n, (int)lo, nb, np, ntile, panel_coop ? 1 : 0);
                }
            } else {
                hipLaunchKernelGGL(bp_trailing<T>, dim3(tiles), dim3(256), smem, s, W[cur], W[cur ^ 1], a.P, a.perm,
                                   n, (int)lo, nb, np, ntile, panel_coop ? 1 : 0);
            }
            // the per-step row kernels start a panel from the identity map; the
            // cooperative panel builds its own (the next panel is cooperative
            // whenever this one was: the panel only shrinks)
            if (!panel_coop) hipLaunchKernelGGL(bp_reset_perm<T>, dim3(64), dim3(256), 0, s, a.perm, lo, n);
            cur ^= 1;
        }
    }
    hipLaunchKernelGGL(bp_finish<T>, dim3(1), dim3(64), 0, s, a.st, log_out, phase_out);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Gradient of log|Pf| for one large matrix: a blocked block sweep
// (Gauss-Jordan with the 2x2 Parlett-Reid pivots) that leaves A^-1 in pivot
// order, from which grad = 1/2 A^-T = -1/2 A^-1 (A^-1 is skew).
//
// Per group of np pair steps, with G = [k0, k0 + 2np) the group's positions
// after its interchanges and T every other position:
//   bp_panel     chooses the group's pivots on the unswept block [k0, n),
//                which is exactly the factorization's Schur complement, so
//                the factorization's panel serves unchanged (perm: logical
//                -> slot on [k0, n); positions below k0 keep their place)
//   gj_rows      Ai = A_GG^-1 by the unblocked sweep of the 2np x 2np block
//                (its pivots are already in place: no search), in every
//                block; then S = A_GT and Z = Ai A_GT at the block's columns
//                of T, in logical order; at the columns of G, S = -I and
//                Z = -Ai
//   gj_trailing  Wn[r][c] = base + sum_h S[h][r] Z[h][c] over every upper
//                tile, base = W[perm r][perm c] for r, c in T, else 0
// This is the principal pivot transform of the matrix on G,
//   new_GG = A_GG^-1,  new_GT = -A_GG^-1 A_GT,  new_TT = A_TT - A_TG A_GG^-1 A_GT,
// i.e. the group's np sweeps of the unblocked form at once: one read and
// one write of the upper triangle per group instead of per pair step.
// `gp` composes the groups' interchanges (logical -> caller's index).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool gj_in(int x, int k0, int m) { return x >= k0 && x < k0 + m; }

// position x of the current matrix in the previous one
__device__ __forceinline__ int gj_src(const int* perm, int x, int k0) {
    const int p = perm[x];
    return x < k0 ? x : p;
}

// A_GG^-1 by the unblocked sweep of the group's 2np x 2np block (its pivots
// are already in place: no search), over the NT threads of a block: thread
// t holds the entries (t / TPR, t % TPR + TPR i) in registers, returned in
// v.  Per pair step the two pivot rows go through LDS (double-buffered by
// step parity: one barrier per step), with the pivot's reciprocal, taken
// once by the thread that holds the pivot.  The pair steps are unrolled, so
// the pivot row and columns are known at compile time: an entry that cannot
// lie in a pivot column takes the plain update without the selects of the
// special cases.
template <typename T, int NB, int NT>
__device__ __forceinline__ void gj_sweep(const T* __restrict__ W, const int* __restrict__ perm, int64_t n, int k0,
                                         int np, T (&v)[2 * NB / (NT / (2 * NB))]) {
    constexpr int M = 2 * NB, TPR = NT / M, EPT = M / TPR;
    static_assert(NT % M == 0 && M % TPR == 0, "the threads tile the block");
    __shared__ T xr[2][M], yr[2][M], rinv[2];
    __shared__ int pg[M];
    const int m = 2 * np, t = threadIdx.x;
    const int r = t / TPR, cq = t % TPR;
    if (t < M) pg[t] = perm[k0 + min(t, m - 1)];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
        const int c = cq + TPR * i;
        const T w = w_at(W, n, pg[min(r, m - 1)], pg[min(c, m - 1)]);
        v[i] = (r < m && c < m) ? w : zero<T>();
    }
    // a full group (m = M, every group but possibly a matrix's last) has no
    // padding entries to keep at zero: its steps skip that select
    auto sweep = [&](auto fullc) {
    constexpr bool FULL = decltype(fullc)::value;
    bp_sfor<0, NB>([&](auto jc) {
        constexpr int J = decltype(jc)::value, K = 2 * J, PAR = J & 1;
        if (J >= np) return;
        if (r == K || r == K + 1) {
#pragma unroll
            for (int i = 0; i < EPT; ++i) {
                if (r == K) xr[PAR][cq + TPR * i] = v[i];
                else yr[PAR][cq + TPR * i] = v[i];
            }
            constexpr int IP = (K + 1) / TPR, CP = (K + 1) % TPR;   // where the pivot (K, K+1) sits
            if (r == K && cq == CP) rinv[PAR] = recip(v[IP]);
        }
        __syncthreads();
        T xc[EPT], yc[EPT];
#pragma unroll
        for (int i = 0; i < EPT; ++i) { xc[i] = xr[PAR][cq + TPR * i]; yc[i] = yr[PAR][cq + TPR * i]; }
        const T inv = rinv[PAR];
        const T xrr = xr[PAR][r], yrr = yr[PAR][r];
        const bool pr = r == K || r == K + 1;
        const T ninv = -inv;
        if (!pr) {
            bp_sfor<0, EPT>([&](auto ic) {
                constexpr int I = decltype(ic)::value;
                const int c = cq + TPR * I;
                const T gen = v[I] + (yrr * xc[I] - xrr * yc[I]) * inv;
                T u = gen;
                if constexpr (K / TPR == I || (K + 1) / TPR == I) {   // the pivot columns may be here
                    const T colv = psel(c == K, -(yrr * inv), xrr * inv);
                    u = psel(c == K || c == K + 1, colv, gen);
                }
                v[I] = FULL ? u : psel(r < m && c < m, u, v[I]);
            });
        } else {
#pragma unroll
            for (int i = 0; i < EPT; ++i) {
                const int c = cq + TPR * i;
                const bool pc = c == K || c == K + 1;
                const T rowv = psel(r == K, yc[i] * inv, -(xc[i] * inv));
                const T pp = psel(r == c, zero<T>(), psel(r == K, ninv, inv));
                const T u = psel(pc, pp, rowv);
                v[i] = FULL ? u : psel(r < m && c < m, u, v[i]);
            }
        }
    });
    };
    if (m == M) sweep(std::true_type{});
    else sweep(std::false_type{});
}

#ifndef GJ_CB64
#define GJ_CB64 8
#endif
// gj_rows: columns per block for a group of M rows (256 / CB row groups)
template <int M>
constexpr int gj_cb() { return M >= 64 ? GJ_CB64 : 8; }

// Each block takes A_GG^-1 itself (gj_sweep, the same for every block) and
// then forms S and Z at its CB columns; the reads of A_GT are issued before
// the sweep, so their latency passes under it.
template <typename T, int NB>
__global__ void __launch_bounds__(256) gj_rows(const T* __restrict__ W, const int* __restrict__ perm, int64_t n,
                                               int k0, int np, const BPState<T>* st, T* __restrict__ S,
                                               T* __restrict__ Z, const int* __restrict__ gp_old,
                                               int* __restrict__ gp_new) {
    constexpr int M = 2 * NB, CB = gj_cb<M>(), HG = 256 / CB, HPT = M / HG;
    constexpr int TPR = 256 / M, EPT = M / TPR, SPT = (M * CB + 255) / 256;
    static_assert(256 % CB == 0 && M % HG == 0, "row groups tile the block");
    __shared__ T As[M][M + 1];
    __shared__ T Sc[M][CB + 1];
    if (st->dead) return;
    const int m = 2 * np, t = threadIdx.x;
    const int x0 = blockIdx.x * CB;
    if (t < CB && x0 + t < n) gp_new[x0 + t] = gp_old[gj_src(perm, x0 + t, k0)];
    T scv[SPT];
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int e = t + 256 * q;
        const int h = e / CB, cx = e % CB, x = x0 + cx;
        T v = zero<T>();
        if (e < M * CB && h < m && x < n) {
            if (gj_in(x, k0, m)) v = (x - k0 == h) ? -one<T>() : zero<T>();
            else v = w_at(W, n, perm[k0 + h], gj_src(perm, x, k0));
        }
        scv[q] = v;
    }
    {
        T v[EPT];
        gj_sweep<T, NB, 256>(W, perm, n, k0, np, v);
        const int r = t / TPR, cq = t % TPR;
#pragma unroll
        for (int i = 0; i < EPT; ++i) As[r][cq + TPR * i] = v[i];
    }
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int e = t + 256 * q;
        if (e < M * CB) Sc[e / CB][e % CB] = scv[q];
    }
    __syncthreads();
    const int cx = t % CB, hg = t / CB, x = x0 + cx;
    T acc[HPT];
#pragma unroll
    for (int i = 0; i < HPT; ++i) acc[i] = zero<T>();
    if (gj_in(x, k0, m)) {
#pragma unroll
        for (int i = 0; i < HPT; ++i) acc[i] = -As[hg + HG * i][x - k0];
    } else {
        for (int hp = 0; hp < m; ++hp) {
            const T sv = Sc[hp][cx];
#pragma unroll
            for (int i = 0; i < HPT; ++i) acc[i] = acc[i] + As[hg + HG * i][hp] * sv;
        }
    }
    if (x < n) {
#pragma unroll
        for (int i = 0; i < HPT; ++i) {
            const int h = hg + HG * i;
            if (h < m) { S[(int64_t)h * n + x] = Sc[h][cx]; Z[(int64_t)h * n + x] = acc[i]; }
        }
    }
}

// upper 64 x 64 tile (ti, tj), tj >= ti, of this block (xcd_tile's order)
__device__ __forceinline__ bool gj_tile(int ntile, int& ti, int& tj) {
    return xcd_tile(blockIdx.x, ntile, ti, tj);
}

// any element type: thread (tx, ty) owns rows r0+ty+16a, cols c0+tx+16b
template <typename T>
__global__ void __launch_bounds__(256) gj_trailing(const T* __restrict__ Wo, T* __restrict__ Wn,
                                                   const T* __restrict__ S, const T* __restrict__ Z,
                                                   const int* __restrict__ perm, int64_t n, int k0, int m, int ntile,
                                                   const BPState<T>* st) {
    if (st->dead) return;
    int ti, tj;
    if (!gj_tile(ntile, ti, tj)) return;
    const int r0 = ti * BP_TILE, c0 = tj * BP_TILE;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4, n1 = (int)n - 1;
    int rr[4], cc[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) { rr[x] = min(r0 + ty + 16 * x, n1); cc[x] = min(c0 + tx + 16 * x, n1); }
    T acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = zero<T>();
    for (int h = 0; h < m; ++h) {
        T sr[4], zc[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) { sr[x] = S[(int64_t)h * n + rr[x]]; zc[x] = Z[(int64_t)h * n + cc[x]]; }
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) acc[x][y] = acc[x][y] + sr[x] * zc[y];
    }
#pragma unroll
    for (int x = 0; x < 4; ++x) {
        const int r = r0 + ty + 16 * x;
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const int c = c0 + tx + 16 * y;
            if (c < n && c > r) {
                T b = zero<T>();
                if (!gj_in(r, k0, m) && !gj_in(c, k0, m)) b = w_at(Wo, n, gj_src(perm, r, k0), gj_src(perm, c, k0));
                Wn[(int64_t)r * n + c] = b + acc[x][y];
            }
        }
    }
}

// FP64 form on the matrix cores: C = S^T Z per 64 x 64 upper tile with
// v_mfma_f64_16x16x4f64 (lane layout as in bp_trailing_mfma: A operand
// A[l % 16][l / 16] = S[k][r], B operand B[l / 16][l % 16] = Z[k][c],
// accumulator register e holds (l / 16 + 4 e, l % 16)).  S and Z are in
// logical order, so their reads are coalesced and need no map.
template <int M>
__global__ void __launch_bounds__(256) gj_trailing_mfma(const double* __restrict__ Wo, double* __restrict__ Wn,
                                                        const double* __restrict__ S, const double* __restrict__ Z,
                                                        const int* __restrict__ perm, int64_t n, int k0, int m,
                                                        int ntile, const BPState<double>* st) {
    if (st->dead) return;
    int ti, tj;
    if (!gj_tile(ntile, ti, tj)) return;
    const int r0 = ti * BP_TILE + 16 * (threadIdx.x >> 6), c0 = tj * BP_TILE;
    const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
    const int n1 = (int)n - 1;
    int pr[4], pc[4];
    bool tr[4], tc[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int r = min(r0 + lq + 4 * e, n1);
        pr[e] = gj_src(perm, r, k0);
        tr[e] = !gj_in(r, k0, m);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int c = min(c0 + 16 * u + lr, n1);
        pc[u] = gj_src(perm, c, k0);
        tc[u] = !gj_in(c, k0, m);
    }
    const int ra = min(r0 + lr, n1);
    double av[M / 4], bv[M / 4][4];
#pragma unroll
    for (int g = 0; g < M / 4; ++g) {
        const int q = 4 * g + lq;
        const bool on = q < m;
        const double va = S[(int64_t)q * n + ra];
        av[g] = on ? va : 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double vb = Z[(int64_t)q * n + min(c0 + 16 * u + lr, n1)];
            bv[g][u] = on ? vb : 0.0;
        }
    }
    double wv[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const double w = w_at(Wo, n, pr[e], pc[u]);
            wv[u][e] = (tr[e] && tc[u]) ? w : 0.0;
        }
    bt_d4 acc[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[u] = bt_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int g = 0; g < M / 4; ++g) {
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[g], bv[g][u], acc[u], 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int r = r0 + lq + 4 * e, c = c0 + 16 * u + lr;
            if (c < n && c > r) Wn[(int64_t)r * n + c] = wv[u][e] + acc[u][e];
        }
    }
}

// complex128 form: each complex product as four real ones
template <int M>
__global__ void __launch_bounds__(256) gj_trailing_mfma_c(const c128* __restrict__ Wo, c128* __restrict__ Wn,
                                                          const c128* __restrict__ S, const c128* __restrict__ Z,
                                                          const int* __restrict__ perm, int64_t n, int k0, int m,
                                                          int ntile, const BPState<c128>* st) {
    if (st->dead) return;
    int ti, tj;
    if (!gj_tile(ntile, ti, tj)) return;
    const int r0 = ti * BP_TILE + 16 * (threadIdx.x >> 6), c0 = tj * BP_TILE;
    const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
    const int n1 = (int)n - 1;
    int pr[4], pc[4];
    bool tr[4], tc[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int r = min(r0 + lq + 4 * e, n1);
        pr[e] = gj_src(perm, r, k0);
        tr[e] = !gj_in(r, k0, m);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int c = min(c0 + 16 * u + lr, n1);
        pc[u] = gj_src(perm, c, k0);
        tc[u] = !gj_in(c, k0, m);
    }
    const int ra = min(r0 + lr, n1);
    bt_d4 are[4], aim[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) { are[u] = bt_d4{0.0, 0.0, 0.0, 0.0}; aim[u] = bt_d4{0.0, 0.0, 0.0, 0.0}; }
#pragma unroll
    for (int g = 0; g < M / 4; ++g) {
        const int q = 4 * g + lq;
        const bool on = q < m;
        const c128 va = S[(int64_t)q * n + ra];
        const double ar = on ? va.re : 0.0, ai = on ? va.im : 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const c128 vb = Z[(int64_t)q * n + min(c0 + 16 * u + lr, n1)];
            const double br = on ? vb.re : 0.0, bi = on ? vb.im : 0.0;
            are[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, are[u], 0, 0, 0);
            are[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, bi, are[u], 0, 0, 0);
            aim[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, aim[u], 0, 0, 0);
            aim[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, aim[u], 0, 0, 0);
        }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int r = r0 + lq + 4 * e, c = c0 + 16 * u + lr;
            const c128 w = w_at(Wo, n, pr[e], pc[u]);
            const bool b = tr[e] && tc[u];
            if (c < n && c > r) Wn[(int64_t)r * n + c] = c128{(b ? w.re : 0.0) + are[u][e], (b ? w.im : 0.0) + aim[u][e]};
        }
    }
}

// Block sweep: Wn[r][c] = base + sum_h S[h][r] Z[h][c] over every upper
// tile (as gj_trailing_mfma / gj_trailing_mfma_c).
template <typename T, int M>
__global__ void __launch_bounds__(256) gj_trailing_staged(const T* __restrict__ Wo, T* __restrict__ Wn,
                                                          const T* __restrict__ S, const T* __restrict__ Z,
                                                          const int* __restrict__ perm, int64_t n, int k0, int m,
                                                          int ntile, const BPState<T>* st) {
    __shared__ T Bs[M][BP_TILE];
    if (st->dead) return;
    int ti, tj;
    if (!gj_tile(ntile, ti, tj)) return;
    const int r0 = ti * BP_TILE + 16 * (threadIdx.x >> 6), c0 = tj * BP_TILE;
    const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
    const int col = threadIdx.x & 63, kq = threadIdx.x >> 6;
    const int n1 = (int)n - 1;
    int pr[4], pc[4];
    bool tr[4], tc[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int r = min(r0 + lq + 4 * e, n1);
        pr[e] = gj_src(perm, r, k0);
        tr[e] = !gj_in(r, k0, m);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int c = min(c0 + 16 * u + lr, n1);
        pc[u] = gj_src(perm, c, k0);
        tc[u] = !gj_in(c, k0, m);
    }
    const int ra = min(r0 + lr, n1), cs = min(c0 + col, n1);
    T sv[M / 4];
#pragma unroll
    for (int i = 0; i < M / 4; ++i) {
        const int q = kq + 4 * i;
        const T v = Z[(int64_t)q * n + cs];
        sv[i] = q < m ? v : zero<T>();
    }
    T av[M / 4];
#pragma unroll
    for (int g = 0; g < M / 4; ++g) {
        const int q = 4 * g + lq;
        const T v = S[(int64_t)q * n + ra];
        av[g] = q < m ? v : zero<T>();
    }
#pragma unroll
    for (int i = 0; i < M / 4; ++i) Bs[kq + 4 * i][col] = sv[i];
    T wv[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const T w = w_at(Wo, n, pr[e], pc[u]);
            wv[u][e] = (tr[e] && tc[u]) ? w : zero<T>();
        }
    __syncthreads();
    bt_d4 re[4], im[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) { re[u] = bt_d4{0.0, 0.0, 0.0, 0.0}; im[u] = bt_d4{0.0, 0.0, 0.0, 0.0}; }
#pragma unroll
    for (int g = 0; g < M / 4; ++g) {
#pragma unroll
        for (int u = 0; u < 4; ++u) tr_mma(av[g], Bs[4 * g + lq][16 * u + lr], re[u], im[u]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int r = r0 + lq + 4 * e, c = c0 + 16 * u + lr;
            if (c < n && c > r) Wn[(int64_t)r * n + c] = tr_out(wv[u][e], re[u], im[u], e);
        }
    }
}

// (skipped once the matrix is dead or a barrier failed: the maps may then
// be unset, and gj_out reads neither)
template <typename T>
__global__ void gj_invperm(const int* __restrict__ gp, int* __restrict__ ig, int64_t n, const BPState<T>* st) {
    if (st->dead || st->err) return;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x)
        ig[gp[x]] = (int)x;
}

// G[i][j] = -1/2 A^-1[i][j]: the swept matrix read through the inverse of
// the composed interchanges (the writes are coalesced)
template <typename T>
__global__ void __launch_bounds__(256) gj_out(const T* __restrict__ W, const int* __restrict__ ig, int64_t n,
                                              const BPState<T>* st, T* __restrict__ G) {
    using R = typename real_of<T>::type;
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= n * n) return;
    if (st->dead || st->err) { G[e] = nan_of<T>(); return; }
    const int64_t i = e / n, j = e % n;
    G[e] = w_at(W, n, ig[i], ig[j]) * (R)(-0.5);
}

template <typename T>
struct GJLayout {
    BPLayout<T> bp;
    size_t s, z, gp0, gp1, ig, total;
};

template <typename T>
static GJLayout<T> gj_layout(int64_t n, int nb) {
    GJLayout<T> L{};
    L.bp = bp_layout<T>(n, nb);
    size_t o = L.bp.total;
    L.s = o; o = al(o + sizeof(T) * (size_t)2 * nb * n);
    L.z = o; o = al(o + sizeof(T) * (size_t)2 * nb * n);
    L.gp0 = o; o = al(o + sizeof(int) * (size_t)n);
    L.gp1 = o; o = al(o + sizeof(int) * (size_t)n);
    L.ig = o; o = al(o + sizeof(int) * (size_t)n);
    L.total = o;
    return L;
}

// The blocked sweep needs the cooperative panel over the whole first
// block; PFH_GRAD_BLOCKED=0 selects the unblocked sweep (read per call).
bool blocked_grad_enabled(int64_t n) {
    const int tv = tune_value(PFH_TUNE_GRAD_BLOCKED);
    const char* e = std::getenv("PFH_GRAD_BLOCKED");
    if (tv == 0 || (tv < 0 && e && e[0] == '0')) return false;
    return coop_enabled() && (n + BC_T - 1) / BC_T <= BC_MAXG && n % 2 == 0;
}

size_t blocked_grad_workspace_bytes(int dtype, int64_t n) {
    const int nb = grad_panel_width(dtype);
    switch (dtype) {
        case DT_F32: return gj_layout<float>(n, nb).total;
        case DT_F64: return gj_layout<double>(n, nb).total;
        case DT_C64: return gj_layout<c64>(n, nb).total;
        case DT_C128: return gj_layout<c128>(n, nb).total;
    }
    return 0;
}

template <typename T, int NB>
static int blocked_grad_nb(const T* A, int64_t n, void* ws, double* log_out, T* phase_out, T* G, hipStream_t s) {
    if (n % 2 != 0 || (n + BC_T - 1) / BC_T > BC_MAXG) return (int)hipErrorInvalidValue;
    const GJLayout<T> L = gj_layout<T>(n, NB);
    char* base = (char*)ws;
    T* W[2] = {(T*)(base + L.bp.w0), (T*)(base + L.bp.w1)};
    BPArgs<T> a{};
    a.P = (T*)(base + L.bp.p);
    a.R = (T*)(base + L.bp.r);
    a.perm = (int*)(base + L.bp.perm);
    a.part_v = (double*)(base + L.bp.pv);
    a.part_i = (int*)(base + L.bp.pi);
    a.st = (BPState<T>*)(base + L.bp.st);
    a.claim = (int*)(base + L.bp.claim);
    a.n = n;
    a.nb = NB;
    a.spin = tune_value(PFH_TUNE_PANEL_SPIN) >= 0 ? tune_value(PFH_TUNE_PANEL_SPIN) : (1 << 22);
    BPCoop<T> co{};
    co.rec = (unsigned long long*)(base + L.bp.rec);
    co.wrec = (unsigned long long*)(base + L.bp.wrec);
    co.cand = (T*)(base + L.bp.cand);
    T* S = (T*)(base + L.s);
    T* Z = (T*)(base + L.z);
    int* gp[2] = {(int*)(base + L.gp0), (int*)(base + L.gp1)};
    int* ig = (int*)(base + L.ig);
    hipError_t e = hipMemcpyAsync(W[0], A, sizeof(T) * n * n, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return (int)e;
    a.W = W[0];
    hipLaunchKernelGGL(bp_init<T>, dim3(64), dim3(256), 0, s, a, co.rec, co.wrec);
    hipLaunchKernelGGL(bp_reset_perm<T>, dim3(64), dim3(256), 0, s, gp[0], (int64_t)0, n);
    const int ntile = (int)((n + BP_TILE - 1) / BP_TILE);
    const int tiles = xcd_tile_blocks(ntile);
    int cur = 0, gc = 0, epoch = 0, wepoch = 0;
    for (int64_t k0 = 0; k0 + 1 < n; k0 += 2 * NB) {
        const int np = (int)((n - k0) / 2 < NB ? (n - k0) / 2 : NB);
        const int m = 2 * np;
        a.W = W[cur];
        e = launch_panel<T, NB>(a, co, (int)k0, np, epoch, wepoch, s);
        if (e != hipSuccess) return (int)e;
        hipLaunchKernelGGL((gj_rows<T, NB>), dim3((unsigned)((n + gj_cb<2 * NB>() - 1) / gj_cb<2 * NB>())), dim3(256), 0, s, W[cur], a.perm, n,
                           (int)k0, np, a.st, S, Z, gp[gc], gp[gc ^ 1]);
        if constexpr (std::is_same<T, double>::value || std::is_same<T, c128>::value) {
            if (trail_staged(PFH_TUNE_TRAIL_SWEEP, true)) {
                hipLaunchKernelGGL((gj_trailing_staged<T, 2 * NB>), dim3(tiles), dim3(256), 0, s, W[cur], W[cur ^ 1],
                                   S, Z, a.perm, n, (int)k0, m, ntile, a.st);
                cur ^= 1;
                gc ^= 1;
                continue;
            }
        }
        if constexpr (std::is_same<T, double>::value) {
            hipLaunchKernelGGL(gj_trailing_mfma<2 * NB>, dim3(tiles), dim3(256), 0, s, W[cur], W[cur ^ 1], S, Z,
                               a.perm, n, (int)k0, m, ntile, a.st);
        } else if constexpr (std::is_same<T, c128>::value) {
            hipLaunchKernelGGL(gj_trailing_mfma_c<2 * NB>, dim3(tiles), dim3(256), 0, s, W[cur], W[cur ^ 1], S, Z,
                               a.perm, n, (int)k0, m, ntile, a.st);
        } else {
            hipLaunchKernelGGL(gj_trailing<T>, dim3(tiles), dim3(256), 0, s, W[cur], W[cur ^ 1], S, Z, a.perm, n,
                               (int)k0, m, ntile, a.st);
        }
        cur ^= 1;
        gc ^= 1;
    }
    hipLaunchKernelGGL(bp_finish<T>, dim3(1), dim3(64), 0, s, a.st, log_out, phase_out);
    hipLaunchKernelGGL(gj_invperm<T>, dim3(64), dim3(256), 0, s, gp[gc], ig, n, a.st);
    hipLaunchKernelGGL(gj_out<T>, dim3((unsigned)((n * n + 255) / 256)), dim3(256), 0, s, W[cur], ig, n, a.st, G);
    return (int)hipGetLastError();
}

template <typename T>
int blocked_grad(const T* A, int64_t n, void* ws, double* log_out, T* phase_out, T* G, hipStream_t s) {
    if constexpr (std::is_same<T, c128>::value) {
        return blocked_grad_nb<T, 16>(A, n, ws, log_out, phase_out, G, s);
    } else if constexpr (std::is_same<T, double>::value) {
        if (grad_panel_width(DT_F64) == 32) return blocked_grad_nb<T, 32>(A, n, ws, log_out, phase_out, G, s);
        return blocked_grad_nb<T, 16>(A, n, ws, log_out, phase_out, G, s);
    } else {
        return blocked_grad_nb<T, 32>(A, n, ws, log_out, phase_out, G, s);
    }
}

#define PFH_INST(T) \
    template int blocked_grad<T>(const T*, int64_t, void*, double*, T*, T*, hipStream_t);
PFH_INST(float)
PFH_INST(double)
PFH_INST(c64)
PFH_INST(c128)
#undef PFH_INST

#define PFH_INST(T) \
    template int blocked_slog<T>(const T*, int64_t, int, bool, void*, double*, T*, hipStream_t);
PFH_INST(float)
PFH_INST(double)
PFH_INST(c64)
PFH_INST(c128)
#undef PFH_INST

}  // namespace pfh
