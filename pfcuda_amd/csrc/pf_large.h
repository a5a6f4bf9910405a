// Device-resident state and launchers of the unblocked large-matrix path.
#pragma once
#include "pf_batch.h"
#include "pf_types.h"

namespace pfh {

// Pivot state carried between the per-step launches; lives in device memory.
template <typename T>
struct LargeState {
    int p;          // pivot index chosen for the current pair step
    int dead;       // 1 once an exact zero pivot was met
    int swaps;      // number of interchanges so far
    int le;         // exponent of the running |product|
    T piv;          // current pivot value
    double lm;      // mantissa of the running |product|
    double ph_re, ph_im;   // running sign (real) or unit phase (complex)
};

// slog of one n x n row-major matrix A (device).  W: n*n scratch, st: state.
template <typename T>
int large_slog(const T* A, int64_t n, T* W, LargeState<T>* st, double* log_out, T* phase_out, hipStream_t s);

// log|Pf|, sign/phase and G = 1/2 A^-T of one n x n matrix (block sweep).
// W: n*n scratch, perm: n ints, st: state.
template <typename T>
int large_grad(const T* A, int64_t n, T* W, int* perm, LargeState<T>* st, double* log_out, T* phase_out, T* G,
               hipStream_t s);

// One workgroup per matrix over per-workgroup slabs of slab_elems elements.
template <typename T>
int launch_glb(const BatchDesc& d, T* slabs, int64_t slab_elems, int grid, hipStream_t s);

}  // namespace pfh
