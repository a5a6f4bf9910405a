// FP64 workgroup-per-matrix Parlett-Reid with the matrix resident in LDS.
//
// The strict upper triangle of the matrix, padded to N16 = 16*ceil(n/16)
// rows, lives in LDS packed by rows: A[i][j] (i < j) at off(i) + j with
// off(i) = i*(N16-2) - i*(i-1)/2 - 1.  A row r is then one LDS read per
// position c: off(r) + c for c > r, and -(off(c) + r) for c < r.  At
// n <= 192 the triangle (146688 bytes) and the four update vectors fit one
// CU's 160 KB of LDS.  Above 192 the first S = N16 - 192 rows (a prefix of
// the packed order, and the first to be eliminated) live in a global slab
// owned by the workgroup and the remaining 192 rows in LDS: after S/2 pair
// steps the matrix is LDS-only.
//
// Pivoting is physical, as in the reference (pf_sm_calc swaps rows and
// columns k+1 and kp, src/pfaffian_sm.cu:48-66), so the trailing matrix
// shrinks.  Two pair steps form a panel.  The pivot chain of a panel runs in
// ONE wave with no workgroup barrier inside it: lane l holds positions
// l + 64 m of the rows it handles, a pivot search is a DPP max-reduction,
// and the vectors stay in the panel's starting coordinates while every LDS
// access maps its position through the interchanges made so far.  Then all
// waves apply the panel's two rank-2 updates together as one rank-4
// v_mfma_f64_16x16x4f64 per live 16 x 16 tile:
//   pair A (k, k+1): x = row k, pivot q = argmax |x[c]|, c > k (lowest
//     index on ties, NaN never wins); y = row q, z = row k+1; row/column q
//     := z (the stale row k+1 is dead from here on).
//   pair B (k+2, k+3): rows taken with pair A's pending correction
//     + U_A[r] X_A[c] - X_A[r] U_A[c] (left-looking inside the panel); z is
//     the raw row.
//   update: A[r][c] += U_A[r] X_A[c] - X_A[r] U_A[c] + U_B[r] X_B[c] - X_B[r] U_B[c]
//     (U = y / pivot): A-operand rows [U_A, X_A, U_B, X_B], B-operand
//     columns [X_A, -U_A, X_B, -U_B].
// This is the skew Schur complement S = A22 + (y x^T - x y^T)/a of
// pf_sm_calc (src/pfaffian_sm.cu:71-89), two pair steps at a time.
#include <cstdlib>

#include "pf_batch.h"
#include "pf_types.h"
#include "pf_wave.h"

namespace pfh {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) double lds_f64;
typedef __attribute__((address_space(3))) int lds_i32;

__device__ __forceinline__ int row_off(int i, int n16) { return i * (n16 - 2) - (i * (i - 1)) / 2 - 1; }

__device__ __forceinline__ int swap_pos(int c, int p, int q) { return c == p ? q : (c == q ? p : c); }

// value held for position p (lane p % 64, register p / 64); p wave-uniform.
// Every register is read out and the choice is made on scalars, so the
// array never needs a run-time register index.
template <int NM>
__device__ __forceinline__ double pos_value(const double (&v)[NM], int p) {
    const int m = p >> 6, l = p & 63;
    double x = readlane_f64(v[0], l);
#pragma unroll
    for (int i = 1; i < NM; ++i) {
        const double xi = readlane_f64(v[i], l);
        x = (m == i) ? xi : x;
    }
    return x;
}

// pivot search: among candidates (key in (lo_key, n), magnitude not NaN) the
// largest magnitude, ties to the smallest key.  Returns the key, or
// `none` when there is no candidate.
template <int NM>
__device__ __forceinline__ int pivot_search(const double (&v)[NM], const int (&key)[NM], int lo_key, int n, int none) {
    double a[NM];
    bool ok[NM];
    double mx = -1.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        a[m] = fabs(v[m]);
        ok[m] = (key[m] > lo_key) & (key[m] < n) & (a[m] == a[m]);   // NaN never wins, as the reference's strict '>'
        mx = fmax(mx, ok[m] ? a[m] : -1.0);
    }
    mx = wave_max(mx);
    if (mx < 0.0) return none;
    int kk = 0x7fffffff;
#pragma unroll
    for (int m = 0; m < NM; ++m) kk = min(kk, (ok[m] && a[m] == mx) ? key[m] : 0x7fffffff);
    return wave_min(kk);
}

}  // namespace

template <int NP, int W, int MODE>
__global__ void __launch_bounds__(W * 64)
pf_lds_kernel(BatchDesc d) {
    constexpr int NM = (NP + 63) / 64;          // positions per lane in the pivot wave
    constexpr bool SPILL = NP > 192;            // first rows in a global slab
    constexpr int TRI = SPILL ? 192 * 191 / 2 : NP * (NP - 1) / 2;   // LDS triangle capacity
    constexpr int SINK = TRI + 4 * NP;          // 256 doubles: sink for masked tile writes
    extern __shared__ double lds_raw[];
    lds_f64* P = (lds_f64*)lds_raw;             // packed triangle (rows >= S)
    lds_f64* OPV = P + TRI;                     // [U_A, X_A, U_B, X_B], NP each
    lds_i32* flag = (lds_i32*)(P + SINK + 256);
    double* G = SPILL ? d.slab + (int64_t)blockIdx.x * d.slab_elems : nullptr;   // rows < S

    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int lr = lane & 15;
    const int lq = lane >> 4;

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
        __syncthreads();   // LDS of the previous matrix is free

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
        const int NTn = (n + 15) >> 4;
        const int N16 = 16 * NTn;
        // rows < S live in G at their packed index, the rest in P at index - G0
        const int S = (SPILL && N16 > 192) ? N16 - 192 : 0;
        const int G0 = S * (N16 - 1) - (S * (S - 1)) / 2;
        auto ld = [&](int ix) -> double {
            if (SPILL && ix < G0) return G[ix];
            return P[ix - G0];
        };
        auto st = [&](int ix, double v) {
            if (SPILL && ix < G0) G[ix] = v;
            else P[ix - G0] = v;
        };

        // ---- load the strict upper triangle, one row per wave at a time ----
        for (int i = w; i < N16; i += W) {
            const int oi = row_off(i, N16);
            for (int jj = i + 1 + lane; jj < N16; jj += 64)
                st(oi + jj, (i < n && jj < n) ? base[(int64_t)i * n + jj] : 0.0);
        }

        // per-lane constants
        int cpos[NM], coff[NM];   // pivot wave: position lane + 64 m and its row offset
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            cpos[m] = lane + 64 * m;
            coff[m] = row_off(cpos[m], N16);
        }
        int tx[4], toff[4], toffl[4];   // tile lanes: row lq + 4 r, its offset, and off(lr) + lq + 4 r
        bool tlow[4];                   // (lq + 4 r) < lr: entry above a diagonal tile's diagonal
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            tx[r] = lq + 4 * r;
            toff[r] = row_off(tx[r], N16) + lr;
            toffl[r] = row_off(lr, N16) + tx[r];
            tlow[r] = tx[r] < lr;
        }
        __syncthreads();

        // row r at mapped positions p[m] (offsets po[m] = off(p[m])) of this lane
        // (positions below lo are eliminated: not read, 0)
        auto read_row = [&](int r, int lo, const int (&p)[NM], const int (&po)[NM], double (&v)[NM]) {
            const int sr = row_off(r, N16);
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                const int pm = p[m];
                const int ix = pm > r ? sr + pm : po[m] + r;
                double x = 0.0;
                if (cpos[m] < N16 && pm >= lo && pm != r) x = ld(ix);
                v[m] = pm > r ? x : -x;
            }
        };
        // row and column r := v at mapped positions
        auto write_row = [&](int r, int lo, const int (&p)[NM], const int (&po)[NM], const double (&v)[NM]) {
            const int sr = row_off(r, N16);
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                const int pm = p[m];
                const int ix = pm > r ? sr + pm : po[m] + r;
                if (cpos[m] < N16 && pm >= lo && pm != r) st(ix, pm > r ? v[m] : -v[m]);
            }
        };

        double prod = 1.0;
        LogAcc la;
        Phase<double> ph;
        bool dead = false;

        for (int k = 0; k < n; k += 4) {
            const bool has_b = (k + 2 < n);
            if (w == 0) {
                // Vectors are kept in the panel's starting coordinates; an
                // interchange only changes where an LDS access lands.
                double x[NM], z[NM], r2[NM], r3[NM];
                const int lo = (k >> 4) << 4;
                read_row(k, lo, cpos, coff, x);
                read_row(k + 1, lo, cpos, coff, z);
                if (has_b) { read_row(k + 2, lo, cpos, coff, r2); read_row(k + 3, lo, cpos, coff, r3); }
                int st = 0;
                // ---- pair A ----
                const int pA = k + 1;
                const int qA = pivot_search(x, cpos, k, n, pA);
                const double a = pos_value(x, qA);
                if (a == 0.0) st = 1;
                if (!st) {
                    const bool swA = (qA != pA);
                    double ua[NM];
                    read_row(qA, lo, cpos, coff, ua);
                    const double inva = 1.0 / a;
#pragma unroll
                    for (int m = 0; m < NM; ++m) ua[m] *= inva;
                    // new position of each held (old) position under interchange A
                    int pa[NM], pao[NM];
                    const int offpA = row_off(pA, N16), offqA = row_off(qA, N16);
#pragma unroll
                    for (int m = 0; m < NM; ++m) {
                        const int c = cpos[m];
                        pa[m] = swap_pos(c, pA, qA);
                        pao[m] = c == pA ? offqA : (c == qA ? offpA : coff[m]);
                    }
                    if (swA) write_row(qA, lo, pa, pao, z);
                    if constexpr (MODE == MODE_PROD) { prod *= a; if (swA) prod = -prod; }
                    else { la.mul(fabs(a)); ph.mul(a); if (swA) ph.flip(); }

                    if (has_b) {
                        // ---- pair B ----
                        const int kb = k + 2, pB = k + 3;
                        // row kb / pB after interchange A, in old coordinates
                        const bool kbz = (qA == kb), pbz = (qA == pB);
                        const int okb = swap_pos(kb, pA, qA);
                        const double uaf = pos_value(ua, okb), xaf = pos_value(x, okb);
                        double xb[NM];
#pragma unroll
                        for (int m = 0; m < NM; ++m) xb[m] = (kbz ? z[m] : r2[m]) + (uaf * x[m] - xaf * ua[m]);
                        const int qB = pivot_search(xb, pa, kb, n, pB);
                        const double b = pos_value(xb, swap_pos(qB, pA, qA));
                        if (b == 0.0) st = 1;
                        if (!st) {
                            const bool swB = (qB != pB);
                            double ub[NM];
                            read_row(qB, lo, pa, pao, ub);   // row qB after interchange A (written above)
                            const int oqB = swap_pos(qB, pA, qA);
                            const double uaq = pos_value(ua, oqB), xaq = pos_value(x, oqB);
                            const double invb = 1.0 / b;
#pragma unroll
                            for (int m = 0; m < NM; ++m) ub[m] = (ub[m] + (uaq * x[m] - xaq * ua[m])) * invb;
                            // final position of each held position
                            int pab[NM], pabo[NM];
                            const int offpB = row_off(pB, N16), offqB = row_off(qB, N16);
#pragma unroll
                            for (int m = 0; m < NM; ++m) {
                                pab[m] = swap_pos(pa[m], pB, qB);
                                pabo[m] = pa[m] == pB ? offqB : (pa[m] == qB ? offpB : pao[m]);
                            }
                            if (swB) {
                                double zb[NM];
#pragma unroll
                                for (int m = 0; m < NM; ++m) zb[m] = pbz ? z[m] : r3[m];
                                write_row(qB, lo, pab, pabo, zb);
                            }
                            if constexpr (MODE == MODE_PROD) { prod *= b; if (swB) prod = -prod; }
                            else { la.mul(fabs(b)); ph.mul(b); if (swB) ph.flip(); }
                            if (k + 4 < n) {
#pragma unroll
                                for (int m = 0; m < NM; ++m) {
                                    if (cpos[m] < N16) {
                                        const int q = pab[m];
                                        OPV[0 * NP + q] = ua[m];
                                        OPV[1 * NP + q] = x[m];
                                        OPV[2 * NP + q] = ub[m];
                                        OPV[3 * NP + q] = xb[m];
                                    }
                                }
                            }
                        }
                    }
                }
                if (lane == 0) *flag = st;
            }
            __syncthreads();
            if (*flag) { dead = true; break; }
            if (k + 4 >= n) break;

            // ---- rank-4 skew update of every live tile, four tiles per batch ----
            // Per batch: every load (tile entries and MFMA operands) first,
            // then the MFMAs, then the stores.  Off-diagonal tiles take the
            // straight path; a diagonal tile (one per tile row) reads its
            // lower half as the negated upper half and stores the upper half.
            {
                const int I0 = (k + 4) >> 4;
                const int mt = NTn - I0;
                const int nt = mt * (mt + 1) / 2;
                const lds_f64* Pv = OPV + lq * NP;          // U_A, X_A, U_B, X_B
                const lds_f64* Qv = OPV + (lq ^ 1) * NP;    // X_A, U_A, X_B, U_B
                const double qs = (lq & 1) ? -1.0 : 1.0;
                int I = I0, J = I0 + w;
                while (J >= NTn && I < NTn) { ++I; J = I + (J - NTn); }
                constexpr int GB = 4;
                for (int tt = w; tt < nt; tt += W * GB) {
                    int Ig[GB], Jg[GB];
#pragma unroll
                    for (int g = 0; g < GB; ++g) {
                        Ig[g] = (tt + W * g < nt) ? I : -1;
                        Jg[g] = J;
                        J += W;
                        while (J >= NTn && I < NTn) { ++I; J = I + (J - NTn); }
                    }
                    d4 acc[GB];
                    double av[GB], bv[GB];
                    int ix[GB][4];
#pragma unroll
                    for (int g = 0; g < GB; ++g) {
                        const int I_ = Ig[g], J_ = Jg[g];
                        if (I_ < 0) continue;
                        const int sI = 16 * I_ * (N16 - 2) - 128 * I_ * I_ + 8 * I_ + 16 * J_;   // S(I) + 16 J
                        const int tI = 16 * I_;
                        const bool inG = SPILL && (16 * I_ < S);   // tile rows in the global slab
                        if (I_ != J_) {
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const int e = toff[r] - __mul24(tI, tx[r]) + sI;   // (16I + tx, 16J + lr)
                                ix[g][r] = e;
                                acc[g][r] = inG ? G[e] : P[e - G0];
                            }
                        } else {
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const int ia = toff[r] - __mul24(tI, tx[r]) + sI;   // above the diagonal
                                const int ib = toffl[r] - __mul24(tI, lr) + sI;     // below it: (j, i)
                                const bool on = tx[r] == lr;
                                const int e = tlow[r] ? ia : ib;
                                double v = 0.0;
                                if (!on) v = inG ? G[e] : P[e - G0];
                                acc[g][r] = tlow[r] ? v : -v;
                                ix[g][r] = tlow[r] ? e : -1;
                            }
                        }
                        av[g] = Pv[16 * I_ + lr];
                        bv[g] = qs * Qv[16 * J_ + lr];
                    }
#pragma unroll
                    for (int g = 0; g < GB; ++g) {
                        if (Ig[g] < 0) continue;
                        acc[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[g], bv[g], acc[g], 0, 0, 0);
                    }
#pragma unroll
                    for (int g = 0; g < GB; ++g) {
                        const int I_ = Ig[g];
                        if (I_ < 0) continue;
                        const bool inG = SPILL && (16 * I_ < S);
                        if (I_ != Jg[g]) {
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                if (inG) G[ix[g][r]] = acc[g][r];
                                else P[ix[g][r] - G0] = acc[g][r];
                            }
                        } else {
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const int e = ix[g][r];
                                if (e >= 0) {
                                    if (inG) G[e] = acc[g][r];
                                    else P[e - G0] = acc[g][r];
                                }
                            }
                        }
                    }
                }
            }
            __syncthreads();
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

namespace {

int lds_num_cu() { return device_cu_count(); }

template <int NP>
constexpr size_t lds_bytes() {
    return sizeof(double) * ((NP > 192 ? 192 * 191 / 2 : NP * (NP - 1) / 2) + 4 * NP + 256 + 1);
}

template <int NP, int W, int MODE>
int lds_launch(const BatchDesc& d, hipStream_t s) {
    if (d.count <= 0) return 0;
    auto kern = pf_lds_kernel<NP, W, MODE>;
    constexpr size_t bytes = lds_bytes<NP>();
    static DevCache resident;   // resident workgroups on the device, once its LDS limit is set
    hipError_t e = hipSuccess;
    int64_t grid = resident.get([&] {
        e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) return 0;
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, W * 64, bytes) != hipSuccess || occ <= 0) occ = 1;
        return occ * lds_num_cu();
    });
    if (grid == 0) return (int)e;
    if (NP > 192) {
        // one spill slab per workgroup
        if (!d.slab || d.slab_elems < lds_spill_elems()) return (int)hipErrorInvalidValue;
        if (grid > lds_spill_slabs()) grid = lds_spill_slabs();
    }
    if (grid > d.count) grid = d.count;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(W * 64), bytes, s, d);
    return (int)hipGetLastError();
}

int lds_waves_override() {
    static const int v = [] {
        const char* e = getenv("PFH_LDS_WAVES");
        return e ? atoi(e) : -1;
    }();
    return v;
}

}  // namespace

int lds_max_n() { return 192; }

// rows 0..63 of a 256-padded matrix: 64 * 255 - 64 * 63 / 2 doubles
int64_t lds_spill_elems() { return 64 * 255 - (64 * 63) / 2; }

int lds_spill_slabs() { return lds_num_cu(); }

int launch_lds_f64(int mode, const BatchDesc& d, int maxn, hipStream_t s) {
    const int ow = lds_waves_override();
    if (mode == MODE_SLOG) {
        if (maxn <= 32) return lds_launch<32, 1, MODE_SLOG>(d, s);
        if (maxn <= 64) {
            if (ow == 1) return lds_launch<64, 1, MODE_SLOG>(d, s);
            return lds_launch<64, 2, MODE_SLOG>(d, s);
        }
        if (maxn <= 128) {
            if (ow == 2) return lds_launch<128, 2, MODE_SLOG>(d, s);
            return lds_launch<128, 4, MODE_SLOG>(d, s);
        }
        if (maxn <= 192) {
            if (ow == 4) return lds_launch<192, 4, MODE_SLOG>(d, s);
            return lds_launch<192, 8, MODE_SLOG>(d, s);
        }
        if (maxn <= 256) {
            if (ow == 4) return lds_launch<256, 4, MODE_SLOG>(d, s);
            return lds_launch<256, 8, MODE_SLOG>(d, s);
        }
    } else if (mode == MODE_PROD) {
        if (maxn <= 32) return lds_launch<32, 1, MODE_PROD>(d, s);
        if (maxn <= 64) return lds_launch<64, 1, MODE_PROD>(d, s);
    }
    return (int)hipErrorInvalidValue;
}

}  // namespace pfh
