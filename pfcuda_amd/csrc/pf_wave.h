// Wave64 helpers shared by the kernels: DPP reductions and lane reads.
//
// The reductions run in six DPP steps (quad_perm [1,0,3,2], [2,3,0,1],
// row_half_mirror, row_mirror, row_bcast:15, row_bcast:31) and read the
// result from lane 63, so every lane gets the same value.  All lanes of the
// wave must be active.
#pragma once
#include <hip/hip_runtime.h>

namespace pfh {

__device__ __forceinline__ double readlane_f64(double x, int l) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)b, (int)(unsigned)b, CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(unsigned)(b >> 32), (int)(unsigned)(b >> 32), CTRL, ROW_MASK, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// wave maximum (no NaN inputs), in every lane
__device__ __forceinline__ double wave_max(double m) {
    m = fmax(m, dpp_f64<0xB1, 0xF>(m));
    m = fmax(m, dpp_f64<0x4E, 0xF>(m));
    m = fmax(m, dpp_f64<0x141, 0xF>(m));
    m = fmax(m, dpp_f64<0x140, 0xF>(m));
    m = fmax(m, dpp_f64<0x142, 0xA>(m));
    m = fmax(m, dpp_f64<0x143, 0xC>(m));
    return readlane_f64(m, 63);
}
// wave minimum, in every lane; the DPP moves leave INT_MAX (the minimum's
// identity) in the lanes they do not write, so each folds into the minimum
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_min_i32(int v) {
    return __builtin_amdgcn_update_dpp(0x7fffffff, v, CTRL, ROW_MASK, 0xF, false);
}
__device__ __forceinline__ int wave_min(int k) {
    k = min(k, dpp_min_i32<0xB1, 0xF>(k));
    k = min(k, dpp_min_i32<0x4E, 0xF>(k));
    k = min(k, dpp_min_i32<0x141, 0xF>(k));
    k = min(k, dpp_min_i32<0x140, 0xF>(k));
    k = min(k, dpp_min_i32<0x142, 0xA>(k));
    k = min(k, dpp_min_i32<0x143, 0xC>(k));
    return __builtin_amdgcn_readlane(k, 63);
}

// Steps whose DPP pattern gives every lane a value: no "old" operand, so
// no copy of the source before the move.
template <int CTRL>
__device__ __forceinline__ double dpp_all_f64(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// DPP move with 0 (the identity of an unsigned maximum) in the lanes the
// pattern does not write, so that the move folds into the maximum
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ unsigned dpp_u32(unsigned v) {
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xF, false);
}
// Wave argmax: the largest `best` (inputs >= -1, no NaN) and, among the
// lanes holding it, the smallest key; (-1, 0x7fffffff) when every lane has
// -1.  When one lane alone holds the maximum (the usual case) its key is read
// directly; ties fall back to a min-reduction of the keys, so the result is
// exactly that of wave_max followed by wave_min.  (A form on the 32-bit
// words of the maximum, one DPP maximum per step, takes fewer instructions
// but a longer chain of dependent ones: the panel's pair steps, which wait
// on it, got slower.)
__device__ __forceinline__ void wave_argmax(double best, int key, double& out_best, int& out_key) {
    double m = best;
    m = fmax(m, dpp_all_f64<0xB1>(m));
    m = fmax(m, dpp_all_f64<0x4E>(m));
    m = fmax(m, dpp_all_f64<0x141>(m));
    m = fmax(m, dpp_all_f64<0x140>(m));
    m = fmax(m, dpp_f64<0x142, 0xA>(m));
    m = fmax(m, dpp_f64<0x143, 0xC>(m));
    const double wb = readlane_f64(m, 63);
    out_best = wb;
    const unsigned long long hit = __ballot(best == wb && wb >= 0.0);
    if (hit == 0) { out_key = 0x7fffffff; return; }
    if ((hit & (hit - 1)) == 0) {                     // one lane holds the maximum
        out_key = __builtin_amdgcn_readlane(key, (int)__builtin_ctzll(hit));
        return;
    }
    out_key = wave_min((best == wb) ? key : 0x7fffffff);
}

}  // namespace pfh
