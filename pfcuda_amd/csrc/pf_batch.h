// Internal launcher interfaces shared by the kernel translation units and
// the C-ABI layer.  Not part of the public C-ABI (see include/pfaffian_hip.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>

namespace pfh {

// Per-device caches of facts the launchers look up on every call (the CU
// count, grids sized from occupancy, dynamic-LDS limits already set): one
// atomic slot per device, filled on first use with that device current, so
// that host threads calling at once and processes that switch devices see
// the current device's value.  A slot holds 0 until known; get() stores
// only the nonzero results of f.
constexpr int kMaxDevices = 64;

inline int current_device() {
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess || d < 0) d = 0;
    return d < kMaxDevices ? d : kMaxDevices - 1;
}

struct DevCache {
    std::atomic<int> v[kMaxDevices];
    template <typename F>
    int get(F f) {
        std::atomic<int>& slot = v[current_device()];
        int x = slot.load(std::memory_order_acquire);
        if (x == 0) {
            x = f();
            if (x != 0) slot.store(x, std::memory_order_release);
        }
        return x;
    }
};

// Multiprocessor (CU) count of the current device (256 when the query fails).
inline int device_cu_count() {
    static DevCache cache;
    return cache.get([] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
            v = 256;
        return v;
    });
}

// One batch of matrices for the workgroup kernels.  Matrix j of the launch
// is order[j] when `order` is set (a size bucket), else j.  Ragged batches
// give per-matrix element offsets and sizes; dense batches give dense_n and
// matrices packed back to back.  Every matrix is row-major n x n.
struct BatchDesc {
    const void* data;
    const int64_t* offsets;   // ragged: element offset of each matrix (nullptr: dense)
    const int32_t* sizes;     // ragged: n of each matrix (nullptr: dense_n)
    const int32_t* order;     // optional list of matrix indices to process
    const int32_t* begin_dev; // optional device-side start of this launch's range in `order`
    const int32_t* count_dev; // optional device-side length of that range
    int64_t count;            // number of entries (upper bound when count_dev is set)
    int32_t dense_n;
    void* prod_out;           // MODE_PROD: Pf as the input type
    double* log_out;          // MODE_SLOG: log|Pf|
    void* phase_out;          // MODE_SLOG: sign / unit phase as the input type
    void* grad_out;           // MODE_GRAD: 1/2 A^-T, laid out like the input
    double* slab;             // f64 slog, 192 < n <= 256: one spill slab per resident workgroup
    int64_t slab_elems;       //   (lds_spill_elems() doubles each), nullptr when absent
    double* bslab;            // f64 / c128 slog, batched blocked kernel: slab pairs of the resident
    int64_t bslab_elems;      //   workgroups (doubles available), nullptr when absent
    int32_t* queue;           // optional work counter, zero before the launch: past its first matrix
                              //   a workgroup takes the next unclaimed one (nullptr: a fixed stride)
};

enum { MODE_PROD = 0, MODE_SLOG = 1, MODE_GRAD = 2 };

// The value set with pfh_set_tuning for a key (PFH_TUNE_*), or -1.
int tune_value(int key);
enum { DT_F32 = 0, DT_F64 = 1, DT_C64 = 2, DT_C128 = 3 };

// Largest n the register-resident workgroup kernels take for a dtype
// (MODE_GRAD: reg_grad_max_n).
int reg_max_n(int dtype);
int reg_grad_max_n(int dtype);

// FP64 LDS-resident kernels (pf_lds.hip): MODE_SLOG up to n = 256 (above
// 192 with d.slab set: the first rows of the triangle spill to a global
// slab per workgroup), MODE_PROD up to 64.
int launch_lds_f64(int mode, const BatchDesc& d, int maxn, hipStream_t stream);
int lds_max_n();                 // 192 without spill slabs
int64_t lds_spill_elems();       // doubles per spill slab
int lds_spill_slabs();           // slabs needed (one per resident workgroup)

// Batched blocked kernel (pf_bat.hip), f64 and c128: MODE_SLOG, n <= 256,
// one workgroup per matrix over a pair of global NP x NP slabs per resident
// workgroup, taken from d.bslab (d.bslab_elems doubles; the grid shrinks to
// what fits, bat_ws_elems(dtype, maxn, count) holds a full one for such a
// batch).
int launch_bat(int dtype, const BatchDesc& d, int maxn, hipStream_t stream);
int64_t bat_ws_elems(int dtype, int maxn, int64_t count);

// One wave per matrix for n <= 32 (pf_c32.hip): MODE_PROD and MODE_SLOG,
// every dtype.
int launch_c32(int dtype, int mode, const BatchDesc& d, hipStream_t stream);

// Launches the workgroup kernel whose capacity covers `maxn` for every
// matrix in the batch.  Returns 0 or a hipError_t value.
int launch_reg(int dtype, int mode, const BatchDesc& d, int maxn, hipStream_t stream);

}  // namespace pfh
