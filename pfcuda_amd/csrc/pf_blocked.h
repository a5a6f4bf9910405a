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

}  // namespace pfh
