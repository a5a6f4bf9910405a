// Launcher of the blocked large-matrix path (pf_blocked.hip).
#pragma once
#include "pf_batch.h"
#include "pf_types.h"

namespace pfh {

// Panel width (pair steps per panel) used for a dtype.
int blocked_panel_width(int dtype);

// Workspace bytes for one n x n matrix.
size_t blocked_workspace_bytes(int dtype, int64_t n, bool mixed);

// log|Pf| and sign/phase of one row-major n x n matrix A (device).
template <typename T>
int blocked_slog(const T* A, int64_t n, int nb, bool mixed, void* ws, double* log_out, T* phase_out,
                 hipStream_t s);

// Gradient path: true when the blocked sweep serves an n x n matrix
// (PFH_GRAD_BLOCKED=0 turns it off).
bool blocked_grad_enabled(int64_t n);

// Workspace bytes of blocked_grad for one n x n matrix.
size_t blocked_grad_workspace_bytes(int dtype, int64_t n);

// log|Pf|, sign/phase and G = 1/2 A^-T of one row-major n x n matrix A
// (device), by the blocked block sweep.
template <typename T>
int blocked_grad(const T* A, int64_t n, void* ws, double* log_out, T* phase_out, T* G, hipStream_t s);

// Synchronizes the stream, then reports (1) whether a panel's poll on the
// current device gave up since the last call, and clears that report.
int blocked_take_status(hipStream_t s, int* timed_out);

}  // namespace pfh
