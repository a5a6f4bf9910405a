// FP64 workgroup-per-matrix Parlett-Reid on MFMA tiles.
//
// The upper triangle of the matrix (tiles (I, J), I <= J, of 16 x 16) lives
// in the accumulator registers of v_mfma_f64_16x16x4f64: lane l of a wave
// holds D[l/16 + 4*r][l%16], r = 0..3, of each tile it owns (rows
// interleaved by 4, one row group per register; operand A[i][k] and B[k][j]
// sit in lane i + 16k and j + 16k).  Pivoting is
// virtual (nothing moves): each wave keeps the logical label of every
// physical column in registers and finds positions with a ballot; the pivot
// search is a DPP reduction, so a pair step touches LDS only to exchange the
// two pivot rows.  Two pair steps
// form a panel:
//   pair A: row f = pos[k] is gathered from the tiles (row f%16 of tiles
//           (f/16, J) and, negated, column f%16 of tiles (I, f/16), I < f/16),
//           pivot q = argmax over live columns, row q gathered the same way;
//           V_A = x, U_A = y / a.
//   pair B: rows f' and q' gathered with pair A's correction
//           + U_A[f'] V_A[c] - V_A[f'] U_A[c]  (left-looking inside the panel).
//   update: every owned tile takes the rank-4 skew update
//           A[r][c] += U_A[r] V_A[c] - V_A[r] U_A[c] + U_B[r] V_B[c] - V_B[r] U_B[c]
//           as ONE mfma: A-operand rows [U_A, V_A, U_B, V_B], B-operand
//           columns [V_A, -U_A, V_B, -U_B].
// Each lane feeds the matrix core one distinct double per operand, so the
// update reads 1 KB of LDS per 1024 FMAs, where the column-per-lane kernel
// broadcasts 16 B per row to every lane.  The arithmetic is the skew Schur
// complement S = A22 + (y x^T - x y^T)/a of pf_sm_calc (src/pfaffian_sm.cu:
// 71-89), taken two pair steps at a time.
#include "pf_batch.h"
#include "pf_types.h"

namespace pfh {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double d4_get(const d4& v, int r) {
    // r is wave-uniform
    if (r < 2) return r == 0 ? v[0] : v[1];
    return r == 2 ? v[2] : v[3];
}

// Wave-wide argmax of a candidate (non-negative magnitude m, key), larger m
// first, then the smaller key.  Magnitudes compare as their IEEE bit
// patterns (valid for m >= 0).  The reduction runs in DPP within rows of 16
// lanes and with two lane swaps across rows; every lane ends with the result.
struct Cand {
    unsigned hi, lo, key;
};
__device__ __forceinline__ bool cand_better(const Cand& a, const Cand& b) {
    if (a.hi != b.hi) return a.hi > b.hi;
    if (a.lo != b.lo) return a.lo > b.lo;
    return a.key < b.key;
}
template <int CTRL>
__device__ __forceinline__ Cand cand_dpp(const Cand& c) {
    Cand o;
    o.hi = (unsigned)__builtin_amdgcn_update_dpp((int)c.hi, (int)c.hi, CTRL, 0xF, 0xF, false);
    o.lo = (unsigned)__builtin_amdgcn_update_dpp((int)c.lo, (int)c.lo, CTRL, 0xF, 0xF, false);
    o.key = (unsigned)__builtin_amdgcn_update_dpp((int)c.key, (int)c.key, CTRL, 0xF, 0xF, false);
    return o;
}
__device__ __forceinline__ Cand cand_xor(const Cand& c, int m) {
    Cand o;
    o.hi = (unsigned)__shfl_xor((int)c.hi, m);
    o.lo = (unsigned)__shfl_xor((int)c.lo, m);
    o.key = (unsigned)__shfl_xor((int)c.key, m);
    return o;
}
__device__ __forceinline__ Cand wave_argmax(Cand c) {
    Cand o;
    o = cand_dpp<0xB1>(c);  if (cand_better(o, c)) c = o;   // quad_perm [1,0,3,2]
    o = cand_dpp<0x4E>(c);  if (cand_better(o, c)) c = o;   // quad_perm [2,3,0,1]
    o = cand_dpp<0x141>(c); if (cand_better(o, c)) c = o;   // row_half_mirror
    o = cand_dpp<0x140>(c); if (cand_better(o, c)) c = o;   // row_mirror
    o = cand_xor(c, 16);    if (cand_better(o, c)) c = o;
    o = cand_xor(c, 32);    if (cand_better(o, c)) c = o;
    return c;
}

// tile slot -> (I, J) over the upper triangle, row-major
__device__ __forceinline__ void tile_ij(int t, int NT, int& I, int& J) {
    int i = 0;
    while (t >= NT - i) { t -= NT - i; ++i; }
    I = i;
    J = i + t;
}

template <int NP, int W, int MODE>
__global__ void __launch_bounds__(W * 64)
pf_tile_kernel(BatchDesc d) {
    constexpr int NT = NP / 16;
    constexpr int NTILES = NT * (NT + 1) / 2;
    constexpr int TPW = (NTILES + W - 1) / W;
    __shared__ double Xs[2][2][NP];   // [panel parity][pair] pivot rows x
    __shared__ double Us[2][2][NP];   // [panel parity][pair] y / a
    constexpr int NM = (NP + 63) / 64;   // columns per lane in the label copy

    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int lr = lane & 15;        // column within a tile
    const int lq = lane >> 4;        // tile rows lq, lq+4, lq+8, lq+12 (registers 0..3)

    int TI[TPW], TJ[TPW];
#pragma unroll
    for (int s = 0; s < TPW; ++s) {
        const int ts = w + W * s;
        int I = 0, J = 0;
        if (ts < NTILES) tile_ij(ts, NT, I, J);
        else { I = -1; J = -1; }
        TI[s] = __builtin_amdgcn_readfirstlane(I);
        TJ[s] = __builtin_amdgcn_readfirstlane(J);
    }

    int64_t count = d.count;
    if (d.count_dev) {
        int64_t cd = *d.count_dev;
        count = cd < count ? cd : count;
    }
    for (int64_t j = blockIdx.x; j < count; j += gridDim.x) {
        const int idx = d.order ? d.order[(d.begin_dev ? *d.begin_dev : 0) + j] : (int)j;
        const int n = __builtin_amdgcn_readfirstlane(d.sizes ? d.sizes[idx] : d.dense_n);
        const int64_t off = d.offsets ? d.offsets[idx] : (int64_t)idx * n * n;
        const double* base = (const double*)d.data + off;
        __syncthreads();

        if (n < 0 || n > NP) {
            if (t == 0) {
                if (MODE == MODE_PROD) ((double*)d.prod_out)[idx] = __builtin_nan("");
                else { d.log_out[idx] = __builtin_nan(""); ((double*)d.phase_out)[idx] = __builtin_nan(""); }
            }
            continue;
        }
        if (n == 0 || (n & 1)) {
            if (t == 0) {
                if (MODE == MODE_PROD) ((double*)d.prod_out)[idx] = 0.0;
                else { d.log_out[idx] = -INFINITY; ((double*)d.phase_out)[idx] = 0.0; }
            }
            continue;
        }

        d4 acc[TPW];
#pragma unroll
        for (int s = 0; s < TPW; ++s) {
            const int I = TI[s], J = TJ[s];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * I + lq + 4 * r, col = 16 * J + lr;
                acc[s][r] = (I >= 0 && row < n && col < n) ? base[(int64_t)row * n + col] : 0.0;
            }
        }
        // every wave keeps its own copy of the labels (logical index of each
        // physical column lane + 64 m); all waves apply the same interchanges
        int labr[NM];
#pragma unroll
        for (int m = 0; m < NM; ++m) labr[m] = lane + 64 * m;
        __syncthreads();

        double prod = 1.0;
        LogAcc la;
        Phase<double> ph;
        bool dead = false;
        int f = 0;
        int par = 0;

        // gather physical row `row` of the stored matrix into dst[c] = fin(c, A[row][c])
        auto gather = [&](int row, double* dst, auto fin) {
            const int F = row >> 4, rr = row & 15;
#pragma unroll
            for (int s = 0; s < TPW; ++s) {
                const int I = TI[s], J = TJ[s];
                if (I == F) {
                    if (lq == (rr & 3)) {
                        const int c = 16 * J + lr;
                        dst[c] = fin(c, d4_get(acc[s], rr >> 2));
                    }
                } else if (J == F) {
                    if (lr == rr) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int c = 16 * I + lq + 4 * r;
                            dst[c] = fin(c, -acc[s][r]);
                        }
                    }
                }
            }
        };
        // pivot search over logical columns (k, n): returns (logical, physical)
        auto argmax = [&](const double* X, int k, int& lq_out, int& q_out) {
            Cand best{0u, 0u, 0xffffffffu};
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                const int c = lane + 64 * m;
                const int l = labr[m];
                if (c < NP && l > k && l < n) {
                    const double v = fabs(X[c]);
                    if (v == v) {   // NaN never wins, as in the reference's strict '>'
                        const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
                        Cand cd{(unsigned)(b >> 32), (unsigned)b, ((unsigned)l << 16) | (unsigned)c};
                        if (cand_better(cd, best)) best = cd;
                    }
                }
            }
            best = wave_argmax(best);
            const unsigned key = __builtin_amdgcn_readfirstlane(best.key);
            if (key == 0xffffffffu) {   // no finite candidate: keep logical k+1
                lq_out = k + 1;
                q_out = -1;
            } else {
                lq_out = (int)(key >> 16);
                q_out = (int)(key & 0xffffu);
            }
        };
        // physical index of logical position `l` (from the label copy)
        auto pos_of = [&](int l) {
            int found = 0;
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                const unsigned long long mask = __builtin_amdgcn_ballot_w64(labr[m] == l);
                if (mask) found = (int)__builtin_ctzll(mask) + 64 * m;
            }
            return __builtin_amdgcn_readfirstlane(found);
        };
        auto relabel = [&](int a_phys, int la, int b_phys, int lb) {
            // physical a_phys takes logical la, b_phys takes lb
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                const int c = lane + 64 * m;
                if (c == a_phys) labr[m] = la;
                else if (c == b_phys) labr[m] = lb;
            }
        };

        for (int k = 0; k < n; k += 4) {
            double* XA = Xs[par][0];
            double* UA = Us[par][0];
            double* XB = Xs[par][1];
            double* UB = Us[par][1];
            const bool has_b = (k + 2 < n);

            // ---- pair A ----
            gather(f, XA, [](int, double v) { return v; });
            __syncthreads();   // B1
            int lqa, qa;
            argmax(XA, k, lqa, qa);
            const int p1a = pos_of(k + 1);
            if (qa < 0) qa = p1a;
            const double pa = XA[qa];
            if (pa == 0.0) { dead = true; break; }
            const bool swa = (lqa != k + 1);
            const int fb = has_b ? ((lqa == k + 2) ? p1a : pos_of(k + 2)) : 0;
            if (swa) relabel(qa, k + 1, p1a, lqa);
            const double inva = 1.0 / pa;
            gather(qa, UA, [inva](int, double v) { return v * inva; });
            __syncthreads();   // B2
            if constexpr (MODE == MODE_PROD) { prod *= pa; if (swa) prod = -prod; }
            else { la.mul(fabs(pa)); ph.mul(pa); if (swa) ph.flip(); }

            int fnext = 0;
            if (has_b) {
                // ---- pair B (left-looking on pair A) ----
                const double uaf = UA[fb], xaf = XA[fb];
                gather(fb, XB, [&](int c, double v) { return v + (uaf * XA[c] - xaf * UA[c]); });
                __syncthreads();   // B3
                const int kb = k + 2;
                int lqb, qb;
                argmax(XB, kb, lqb, qb);
                const int p1b = pos_of(kb + 1);
                if (qb < 0) qb = p1b;
                const double pb = XB[qb];
                if (pb == 0.0) { dead = true; break; }
                const bool swb = (lqb != kb + 1);
                fnext = (kb + 2 < n) ? ((lqb == kb + 2) ? p1b : pos_of(kb + 2)) : 0;
                if (swb) relabel(qb, kb + 1, p1b, lqb);
                const double invb = 1.0 / pb;
                const double uaq = UA[qb], xaq = XA[qb];
                gather(qb, UB, [&](int c, double v) { return (v + (uaq * XA[c] - xaq * UA[c])) * invb; });
                __syncthreads();   // B4
                if constexpr (MODE == MODE_PROD) { prod *= pb; if (swb) prod = -prod; }
                else { la.mul(fabs(pb)); ph.mul(pb); if (swb) ph.flip(); }
            } else {
                for (int c = t; c < NP; c += W * 64) { XB[c] = 0.0; UB[c] = 0.0; }
                __syncthreads();
            }

            // ---- rank-4 skew update of every owned tile ----
            // A-operand: lane (lq, lr) supplies P_lq[16 I + lr]; B-operand: Q_lq[16 J + lr]
            // P = [U_A, X_A, U_B, X_B], Q = [X_A, -U_A, X_B, -U_B]
            const double* Pv = (lq == 0) ? UA : (lq == 1) ? XA : (lq == 2) ? UB : XB;
            const double* Qv = (lq == 0) ? XA : (lq == 1) ? UA : (lq == 2) ? XB : UB;
            const double qs = (lq & 1) ? -1.0 : 1.0;
            // operands for every slot first (slots past the last tile work on a
            // dummy accumulator that nothing reads), then the MFMAs back to back
            double av[TPW], bv[TPW];
#pragma unroll
            for (int s = 0; s < TPW; ++s) {
                const int I = TI[s] < 0 ? 0 : TI[s], J = TJ[s] < 0 ? 0 : TJ[s];
                av[s] = Pv[16 * I + lr];
                bv[s] = qs * Qv[16 * J + lr];
            }
#pragma unroll
            for (int s = 0; s < TPW; ++s)
                acc[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s], bv[s], acc[s], 0, 0, 0);
            f = fnext;
            par ^= 1;
        }

        if (t == 0) {
            if constexpr (MODE == MODE_PROD) {
                ((double*)d.prod_out)[idx] = dead ? 0.0 : prod;
            } else {
                d.log_out[idx] = dead ? -INFINITY : la.value();
                ((double*)d.phase_out)[idx] = dead ? 0.0 : ph.s;
            }
        }
    }
}

static int tile_num_cu() {
    static int v = 0;
    if (v == 0) {
        int dev = 0;
        hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
    }
    return v;
}

template <int NP, int W, int MODE>
static int tile_launch(const BatchDesc& d, hipStream_t s) {
    if (d.count <= 0) return 0;
    auto kern = pf_tile_kernel<NP, W, MODE>;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, W * 64, 0) != hipSuccess || occ <= 0) occ = 1;
    int64_t grid = (int64_t)occ * tile_num_cu();
    if (grid > d.count) grid = d.count;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(W * 64), 0, s, d);
    return (int)hipGetLastError();
}

int tile_max_n() { return 256; }

int launch_tile_f64(int mode, const BatchDesc& d, int maxn, hipStream_t s) {
    if (mode == MODE_SLOG) {
        if (maxn <= 32) return tile_launch<32, 1, MODE_SLOG>(d, s);
        if (maxn <= 64) return tile_launch<64, 2, MODE_SLOG>(d, s);
        if (maxn <= 128) return tile_launch<128, 8, MODE_SLOG>(d, s);
        if (maxn <= 192) return tile_launch<192, 12, MODE_SLOG>(d, s);
        if (maxn <= 256) return tile_launch<256, 16, MODE_SLOG>(d, s);
    } else if (mode == MODE_PROD) {
        if (maxn <= 32) return tile_launch<32, 1, MODE_PROD>(d, s);
        if (maxn <= 64) return tile_launch<64, 2, MODE_PROD>(d, s);
    }
    return (int)hipErrorInvalidValue;
}

}  // namespace pfh
