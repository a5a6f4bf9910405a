// Register-resident row storage for the workgroup Pfaffian kernels.
//
// A thread keeps N elements of one matrix column in VGPRs.  Elements are
// grouped into ext_vector chunks of 32 dwords so that an access with a
// wave-uniform run-time index lowers to s_set_gpr_idx moves instead of a
// trip through scratch memory; accesses with compile-time indices are plain
// register reads.
#pragma once
#include "pf_types.h"

namespace pfh {

template <typename R, int N>
struct RealRegs {
    static constexpr int CH = 128 / (int)sizeof(R);   // elements per 32-dword chunk
    static constexpr int NCH = (N + CH - 1) / CH;
    static_assert(NCH <= 4, "at most four register chunks per thread");
    typedef R vec_t __attribute__((ext_vector_type(CH)));
    // Separate members rather than an array: an array of vectors is not
    // promoted to registers by the compiler once it spans three chunks.
    vec_t v0, v1, v2, v3;

    template <int CHK> __device__ __forceinline__ vec_t& chunk() {
        if constexpr (CHK == 0) return v0;
        else if constexpr (CHK == 1) return v1;
        else if constexpr (CHK == 2) return v2;
        else return v3;
    }
    template <int CHK> __device__ __forceinline__ const vec_t& chunk() const {
        if constexpr (CHK == 0) return v0;
        else if constexpr (CHK == 1) return v1;
        else if constexpr (CHK == 2) return v2;
        else return v3;
    }

    // i is a compile-time constant after unrolling.
    __device__ __forceinline__ R get(int i) const {
        const int ch = i / CH, o = i % CH;
        if (ch == 0) return v0[o];
        if (ch == 1) return v1[o];
        if (ch == 2) return v2[o];
        return v3[o];
    }
    __device__ __forceinline__ void set(int i, R x) {
        const int ch = i / CH, o = i % CH;
        if (ch == 0) v0[o] = x;
        else if (ch == 1) v1[o] = x;
        else if (ch == 2) v2[o] = x;
        else v3[o] = x;
    }

    // i must be wave-uniform.
    __device__ __forceinline__ R dget(int i) const {
        const int ch = i / CH, o = i - (i / CH) * CH;
        R out = R(0);
        if (ch == 0) out = v0[o];
        else if (NCH > 1 && ch == 1) out = v1[o];
        else if (NCH > 2 && ch == 2) out = v2[o];
        else if (NCH > 3) out = v3[o];
        return out;
    }
    __device__ __forceinline__ void dset(int i, R x) {
        const int ch = i / CH, o = i - (i / CH) * CH;
        if (ch == 0) v0[o] = x;
        else if (NCH > 1 && ch == 1) v1[o] = x;
        else if (NCH > 2 && ch == 2) v2[o] = x;
        else if (NCH > 3) v3[o] = x;
    }
};

template <typename T, int N>
struct Regs {
    RealRegs<T, N> r;
    __device__ __forceinline__ T get(int i) const { return r.get(i); }
    __device__ __forceinline__ void set(int i, T x) { r.set(i, x); }
    __device__ __forceinline__ T dget(int i) const { return r.dget(i); }
    __device__ __forceinline__ void dset(int i, T x) { r.dset(i, x); }
};

template <typename R, int N>
struct Regs<cplx<R>, N> {
    RealRegs<R, N> re, im;
    __device__ __forceinline__ cplx<R> get(int i) const { return {re.get(i), im.get(i)}; }
    __device__ __forceinline__ void set(int i, cplx<R> x) { re.set(i, x.re); im.set(i, x.im); }
    __device__ __forceinline__ cplx<R> dget(int i) const { return {re.dget(i), im.dget(i)}; }
    __device__ __forceinline__ void dset(int i, cplx<R> x) { re.dset(i, x.re); im.dset(i, x.im); }
};

}  // namespace pfh
