/* C ABI of the MI355X (gfx950) Pfaffian library, libpfaffian_hip.so.
 *
 * Every entry point takes device pointers, sizes and a hipStream_t, enqueues
 * its work on that stream, never synchronises the host, never throws, and
 * returns 0 on success, a negative PFH_E* code for a rejected argument, or a
 * positive hipError_t value.  Matrices are dense row-major n x n slabs packed
 * back to back (the logical matrix A[i][j] at i*n + j).  Complex data are
 * interleaved (re, im) pairs of float (c64) or double (c128).
 *
 * Zero pivots are detected exactly (pivot == 0) and give Pf = 0, log|Pf| =
 * -inf and sign/phase = 0; odd n gives the same.
 */
#ifndef PFAFFIAN_HIP_H
#define PFAFFIAN_HIP_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFH_OK 0
#define PFH_EINVAL (-1)      /* bad size, count or pointer */
#define PFH_EWORKSPACE (-2)  /* workspace too small */
#define PFH_ETOOBIG (-3)     /* n beyond what the entry point supports */
#define PFH_ETIMEOUT (-4)    /* a blocked panel's blocks were not resident together (see pfh_stream_status) */

/* dtype codes */
#define PFH_F32 0
#define PFH_F64 1
#define PFH_C64 2
#define PFH_C128 3

/* slog methods */
#define PFH_METHOD_AUTO 0
#define PFH_METHOD_UNBLOCKED 1   /* plain all-FP64 path, one rank-2 update per pair step */
#define PFH_METHOD_BLOCKED 2     /* blocked, FP64 trailing update */
#define PFH_METHOD_MIXED 3       /* blocked, FP64 panel, float64 trailing update on the bf16 matrix cores with
                                    exact digit products recombined in FP64 (other dtypes: as BLOCKED) */

/* Short text for a status code. */
const char* pfh_status_string(int status);

/* Largest n that pfh_pfaffian_* accepts (64). */
int pfh_pfaffian_max_n(void);

/* Waits for the work enqueued on `stream`, then returns PFH_ETIMEOUT if a
 * blocked panel on the current device gave up waiting for its blocks since
 * the last call (the outputs of the call it belonged to are NaN), and clears
 * that report; else 0, or a hipError_t.  The blocked panel assumes its
 * <= 64 blocks become resident together; work from other streams holding
 * the CUs for seconds can defeat that, and this is how a caller learns of
 * it.  No counterpart in the reference, whose large path synchronizes with
 * the host at every pivot (src/slog_pfaffian_lg.cu:178). */
int pfh_stream_status(hipStream_t stream);

/* Pf of `batch` matrices of size n (product form, the input dtype).
 * Replaces the XLA FFI handlers pfaffian_{f32,f64,c64,c128}
 * (reference bindings/jax_bindings.cu:82-112 and :152-182), which loop over
 * the batch launching one 32-thread block per matrix
 * (src/pfaffian.cu:42-48, src/pfaffian_sm.cu:101-105).  n <= 64. */
int pfh_pfaffian_f32(const float* A, int64_t batch, int32_t n, float* out, hipStream_t stream);
int pfh_pfaffian_f64(const double* A, int64_t batch, int32_t n, double* out, hipStream_t stream);
int pfh_pfaffian_c64(const void* A, int64_t batch, int32_t n, void* out, hipStream_t stream);
int pfh_pfaffian_c128(const void* A, int64_t batch, int32_t n, void* out, hipStream_t stream);

/* Workspace bytes needed by pfh_slog_pfaffian_* for (dtype, batch, n, method). */
int pfh_slog_pfaffian_workspace_size(int dtype, int64_t batch, int32_t n, int method, size_t* bytes);

/* log|Pf| (double) and sign or unit phase (input dtype) of `batch` matrices
 * of size n.  Replaces the XLA FFI handlers slog_pfaffian_{f32,f64,c64,c128}
 * (reference bindings/jax_bindings.cu:19-80 and :114-149) and the
 * host-driven slog_pfaffian_lg (src/slog_pfaffian_lg.cu:151-211).  The log
 * output is double for every dtype.  The input is never modified.
 * `workspace` may be NULL, in which case the library uses a cached
 * per-device buffer grown with hipMalloc (not capture-safe).  Every such
 * call on a device shares that buffer: calls that overlap (other streams,
 * other host threads) race on it, so concurrent callers pass their own
 * workspace.  Growing it synchronizes the device first. */
int pfh_slog_pfaffian_f32(const float* A, int64_t batch, int32_t n, double* log_out, float* sign_out,
                          void* workspace, size_t workspace_bytes, hipStream_t stream);
int pfh_slog_pfaffian_f64(const double* A, int64_t batch, int32_t n, double* log_out, double* sign_out,
                          void* workspace, size_t workspace_bytes, hipStream_t stream);
int pfh_slog_pfaffian_c64(const void* A, int64_t batch, int32_t n, double* log_out, void* phase_out,
                          void* workspace, size_t workspace_bytes, hipStream_t stream);
int pfh_slog_pfaffian_c128(const void* A, int64_t batch, int32_t n, double* log_out, void* phase_out,
                           void* workspace, size_t workspace_bytes, hipStream_t stream);

/* Same, with an explicit method (PFH_METHOD_*) for matrices past the
 * register-resident kernels. No counterpart in the reference: it selects
 * between the plain and blocked forms of its slog_pfaffian_lg. */
int pfh_slog_pfaffian_method(int dtype, const void* A, int64_t batch, int32_t n, int method,
                             double* log_out, void* phase_out,
                             void* workspace, size_t workspace_bytes, hipStream_t stream);

/* Workspace bytes for a ragged batch of `count` matrices no larger than max_n. */
int pfh_slog_pfaffian_ragged_workspace_size(int dtype, int64_t count, int32_t max_n, size_t* bytes);

/* log|Pf| and sign/phase of a ragged batch in one pass: matrix i is the
 * sizes[i] x sizes[i] row-major slab starting at element offsets[i] of
 * `data`.  offsets and sizes are device arrays; max_n bounds every size.
 * Matrices are bucketed by size on the device and each bucket runs as one
 * persistent launch, so the call has no host synchronisation.  No
 * counterpart in the reference, whose handlers take one dense shape and
 * loop over it serially (bindings/jax_bindings.cu:52-72). */
int pfh_slog_pfaffian_ragged(int dtype, const void* data, const int64_t* offsets, const int32_t* sizes,
                             int64_t count, int32_t max_n, double* log_out, void* phase_out,
                             void* workspace, size_t workspace_bytes, hipStream_t stream);

/* pfh_slog_pfaffian_ragged without any allocation: no counterpart in the
 * reference; for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  The
 * workspace is a fixed device array inside the library,
 * PFH_NOALLOC_SCRATCH_BYTES (256 MiB), which every device the library's code
 * is loaded on holds until the process ends.  Limits: max_n <=
 * PFH_NOALLOC_MAX_N (256) and count <= PFH_NOALLOC_MAX_COUNT (2^22), else
 * PFH_ETOOBIG.  The batched blocked kernel (f64 / c128, 64 < n <= 256) runs
 * with as many workgroups as slab pairs fit in the array (fewer than
 * pfh_slog_pfaffian_ragged may use, same results).  Every call on a device
 * shares the one array, and each call clears what must start at zero:
 * calls that overlap on one device (other streams, other host threads)
 * race on it and must not be made. */
#define PFH_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
#define PFH_NOALLOC_MAX_N 256
#define PFH_NOALLOC_MAX_COUNT ((int64_t)1 << 22)
int pfh_slog_pfaffian_ragged_noalloc(int dtype, const void* data, const int64_t* offsets, const int32_t* sizes,
                                     int64_t count, int32_t max_n, double* log_out, void* phase_out,
                                     hipStream_t stream);

/* Workspace bytes for pfh_slog_pfaffian_grad. */
int pfh_slog_pfaffian_grad_workspace_size(int dtype, int64_t batch, int32_t n, size_t* bytes);

/* log|Pf|, sign/phase and the gradient G = d log Pf / dA = 1/2 A^-T of
 * `batch` matrices of size n, from one block-sweep pass (Gauss-Jordan with
 * the forward pass's 2x2 pivots), so the factorization is recomputed rather
 * than stored.  grad_out has the input's shape and dtype.  Replaces the
 * reference's reverse-mode path, which differentiates its JVP rule
 * (pfcuda/cuda_api.py:93-107) through jnp.linalg.inv (:101). */
int pfh_slog_pfaffian_grad(int dtype, const void* A, int64_t batch, int32_t n, double* log_out,
                           void* phase_out, void* grad_out, void* workspace, size_t workspace_bytes,
                           hipStream_t stream);

/* Host-pointer forms: one row-major n x n matrix in HOST memory, copied to
 * the current device, computed, and the result copied back; the call
 * returns when the result is in *h_out (synchronous, on the null stream),
 * and frees every device buffer it made.  Replace the reference's
 * host-pointer templates pfaffian<T>(h_A, n) (src/pfaffian.cu:4-40) and
 * slog_pfaffian<T>(h_A, n, h_log_abs, h_phase) (src/slog_pfaffian.cu:4-32),
 * which read column-major input and leak their device buffers (slog).
 * pfh_pfaffian_host takes n <= 64 (PFH_ETOOBIG past it) and writes the
 * input dtype; pfh_slog_pfaffian_host takes any n and writes log|Pf|
 * (double) and the sign / phase (input dtype). */
int pfh_pfaffian_host(int dtype, const void* h_A, int32_t n, void* h_out);
int pfh_slog_pfaffian_host(int dtype, const void* h_A, int32_t n, double* h_log, void* h_phase);

/* Pfaffian of one row-major n x n float64 matrix in HOST memory, on the
 * CPU (FP64, one thread, product form), working in place on A: the
 * counterpart of the reference's public pfaffian_cpu (pfcuda/cpp_api.py:3-4,
 * bindings/pybind_bindings.cpp:8-30, src/pfaffian_cpu.cpp:7-106).  n = 0
 * gives 0, n = 2 and 4 the closed forms; odd n returns PFH_EINVAL.  A
 * separate host entry point: no GPU path calls it. */
int pfh_pfaffian_cpu_f64(double* A, int32_t n, double* out);

/* Tuning.  Each key names a choice the library makes at run time.  Its
 * built-in value was measured on MI355X.  A PFH_* environment variable can
 * override it, and a value set here overrides both for the rest of the
 * process; -1 restores the built-in choice.  The autotuner of the host package
 * (pfcuda_amd/autotune.py) times the candidates on the current device and
 * sets these.  No counterpart in the reference, whose launch shapes are
 * fixed at compile time. */
#define PFH_TUNE_F64_BUCKET_128 0   /* f64 slog, 64 < n <= 128: 0 column-per-lane, 2 LDS-resident, 3 batched blocked */
#define PFH_TUNE_F64_BUCKET_192 1   /* f64 slog, 128 < n <= 192: same values */
#define PFH_TUNE_F64_BUCKET_256 2   /* f64 slog, 192 < n <= 256: same values */
#define PFH_TUNE_C128_BUCKET_128 3  /* c128 slog, 64 < n <= 128: 0 column-per-lane, 3 batched blocked */
#define PFH_TUNE_TRAIL_F64 4        /* blocked factorization, f64 trailing update: 0 register form, 1 LDS-staged */
#define PFH_TUNE_TRAIL_C128 5       /* blocked factorization, c128 trailing update: same values */
#define PFH_TUNE_TRAIL_SWEEP 6      /* block sweep (gradient) trailing update: same values */
#define PFH_TUNE_PANEL_LAUNCH 7     /* blocked panel launch form: 0 plain launch (the only form) */
#define PFH_TUNE_GRAD_BLOCKED 8     /* gradient past the register kernel: 0 unblocked sweep, 1 blocked sweep */
#define PFH_TUNE_PANEL_SPIN 9       /* blocked panel: reads of the candidate records a poll makes before the
                                       panel sets its error flag (output NaN), 1 .. 4194304 (the default);
                                       0 sets it at once.  For tests of that path, not a tuning choice */
#define PFH_TUNE_PANEL_XCD 10      /* blocked panel exchange: 2 one participant per wave through one XCD's L2,
                                       1 one per block of 256 slots through one XCD's L2, 0 one per block
                                       device-wide (a form whose participants do not fit one XCD falls back to
                                       the next); built-in: 2 */
#define PFH_TUNE_MIXED_SLICES 11   /* PFH_METHOD_MIXED: 7-bit digits per value of the split-precision trailing
                                       update, 5 .. 9 */
#define PFH_TUNE_BAT_NB_128 12     /* f64 batched blocked kernel, 64 < n <= 128: pairs per panel, 6, 8 or 12 */
#define PFH_TUNE_BAT_NB_192 13     /* f64, 128 < n <= 192: 6, 8 or 12 */
#define PFH_TUNE_BAT_NB_256 14     /* f64, 192 < n <= 256: 8, 10 or 12 */
#define PFH_TUNE_BAT_NB_C128_128 15  /* c128 batched blocked kernel, 64 < n <= 128: 4, 6 or 8 */
#define PFH_TUNE_BAT_NB_C128_192 16  /* c128, 128 < n <= 192: 4, 6 or 8 */
#define PFH_TUNE_BAT_NB_C128_256 17  /* c128, 192 < n <= 256: 6 or 8 */
#define PFH_TUNE_BLOCKED_NB 18     /* f64 blocked factorization: pairs per panel, 16, 24 or 32 */
#define PFH_TUNE_BLOCKED_NB_C128 19  /* c128 blocked factorization: 8 or 16 */
#define PFH_TUNE_SWITCH_N 20       /* dense slog of fewer matrices than the device has CUs: the largest n the
                                      batched kernels take (64, 96, ..., 256); larger ones run blocked */
#define PFH_TUNE_SMALL_KERNEL 21   /* pfaffian / slog, n <= 32: 1 one wave per matrix updating the live rows (the
                                      default), 0 the column-per-lane kernel that updates the whole square */
#define PFH_TUNE_GRAD_NB 22         /* blocked gradient sweep, float64: pairs per group, 16 (the default) or 32 */
#define PFH_TUNE_NKEYS 23

/* Sets a tuning key (value -1: the built-in choice); PFH_EINVAL for an
 * unknown key or a value outside the key's list.  Takes effect for calls
 * enqueued afterwards; not for use while another thread enqueues work. */
int pfh_set_tuning(int key, int value);

/* The value set for a key, or -1 when none is (the environment or the
 * built-in choice decides). */
int pfh_get_tuning(int key, int* value);

#ifdef __cplusplus
}
#endif

#endif /* PFAFFIAN_HIP_H */
