// Register-resident row storage for the workgroup Pfaffian kernels.
//
// A thread keeps N elements of one matrix column in VGPRs, grouped into
// ext_vector chunks of 32 dwords.  Accesses with compile-time indices are
// plain register accesses; a read with a wave-uniform run-time index is a
// short binary search of scalar branches.  There is deliberately no
// run-time-indexed write: the compiler copies a whole chunk around each one.
#pragma once
#include "pf_types.h"

namespace pfh {

// NCH chunks held as separate members (an array of vectors is not promoted
// to registers once it spans three chunks); unused members are empty.
template <typename V, bool USED> struct Chunk { V v; };
template <typename V> struct Chunk<V, false> {};

template <typename R, int N>
struct RealRegs {
    static constexpr int CH = 128 / (int)sizeof(R);   // elements per 32-dword chunk
    static constexpr int NCH = (N + CH - 1) / CH;
    static_assert(NCH <= 4, "at most four register chunks per thread");
    typedef R vec_t __attribute__((ext_vector_type(CH)));
    Chunk<vec_t, true> c0;
    Chunk<vec_t, (NCH > 1)> c1;
    Chunk<vec_t, (NCH > 2)> c2;
    Chunk<vec_t, (NCH > 3)> c3;

    // i is a compile-time constant after unrolling.
    __device__ __forceinline__ R get(int i) const {
        const int ch = i / CH, o = i % CH;
        if constexpr (NCH > 3) if (ch == 3) return c3.v[o];
        if constexpr (NCH > 2) if (ch == 2) return c2.v[o];
        if constexpr (NCH > 1) if (ch == 1) return c1.v[o];
        return c0.v[o];
    }
    __device__ __forceinline__ void set(int i, R x) {
        const int ch = i / CH, o = i % CH;
        if constexpr (NCH > 3) if (ch == 3) { c3.v[o] = x; return; }
        if constexpr (NCH > 2) if (ch == 2) { c2.v[o] = x; return; }
        if constexpr (NCH > 1) if (ch == 1) { c1.v[o] = x; return; }
        c0.v[o] = x;
    }

    // Read with a wave-uniform run-time index (s_set_gpr_idx).
    // Read with a wave-uniform run-time index: a binary search of uniform
    // branches down to a static register read.  (Indexing the vectors with
    // the run-time value directly lets the optimiser turn the store into
    // memory once it spans three chunks.)
    template <int LO, int HI>
    __device__ __forceinline__ R dget_bs(int i) const {
        if constexpr (HI - LO == 1) {
            return get(LO);
        } else {
            constexpr int MID = (LO + HI) / 2;
            if (i < MID) return dget_bs<LO, MID>(i);
            return dget_bs<MID, HI>(i);
        }
    }
    __device__ __forceinline__ R dget(int i) const { return dget_bs<0, N>(i); }
};

template <typename T, int N>
struct Regs {
    RealRegs<T, N> r;
    __device__ __forceinline__ T get(int i) const { return r.get(i); }
    __device__ __forceinline__ void set(int i, T x) { r.set(i, x); }
    __device__ __forceinline__ T dget(int i) const { return r.dget(i); }
};

template <typename R, int N>
struct Regs<cplx<R>, N> {
    RealRegs<R, N> re, im;
    __device__ __forceinline__ cplx<R> get(int i) const { return {re.get(i), im.get(i)}; }
    __device__ __forceinline__ void set(int i, cplx<R> x) { re.set(i, x.re); im.set(i, x.im); }
    __device__ __forceinline__ cplx<R> dget(int i) const { return {re.dget(i), im.dget(i)}; }
};

}  // namespace pfh
