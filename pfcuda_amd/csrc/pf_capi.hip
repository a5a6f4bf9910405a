// C-ABI entry points (declared in include/pfaffian_hip.h) and the
// device-side size bucketing of ragged batches.
#include <atomic>
#include <mutex>
#include <vector>

#include "../../include/pfaffian_hip.h"
#include "pf_batch.h"
#include "pf_large.h"
#include "pf_blocked.h"

using namespace pfh;

namespace {

size_t dtype_bytes(int dt) {
    switch (dt) {
        case PFH_F32: return 4;
        case PFH_F64: return 8;
        case PFH_C64: return 8;
        case PFH_C128: return 16;
    }
    return 0;
}

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Tuning keys set through the C ABI, stored as value + 1 (0: not set).
std::atomic<int> g_tune[PFH_TUNE_NKEYS];

bool tune_value_ok(int key, int v) {
    if (v == -1) return true;
    switch (key) {
        case PFH_TUNE_F64_BUCKET_128:
        case PFH_TUNE_F64_BUCKET_192:
        case PFH_TUNE_F64_BUCKET_256: return v == 0 || v == 2 || v == 3;
        case PFH_TUNE_C128_BUCKET_128: return v == 0 || v == 3;
        case PFH_TUNE_TRAIL_F64:
        case PFH_TUNE_TRAIL_C128:
        case PFH_TUNE_TRAIL_SWEEP:
        case PFH_TUNE_GRAD_BLOCKED:
        case PFH_TUNE_SMALL_KERNEL: return v == 0 || v == 1;
        case PFH_TUNE_PANEL_XCD: return v >= 0 && v <= 2;
        case PFH_TUNE_PANEL_LAUNCH: return v == 0;   // the plain launch is the only form
        case PFH_TUNE_PANEL_SPIN: return v >= 0 && v <= (1 << 22);
        case PFH_TUNE_MIXED_SLICES: return v >= 5 && v <= 9;
        case PFH_TUNE_BAT_NB_128:
        case PFH_TUNE_BAT_NB_192: return v == 6 || v == 8 || v == 12;
        case PFH_TUNE_BAT_NB_256: return v == 8 || v == 10 || v == 12;
        case PFH_TUNE_BAT_NB_C128_128:
        case PFH_TUNE_BAT_NB_C128_192: return v == 4 || v == 6 || v == 8;
        case PFH_TUNE_BAT_NB_C128_256: return v == 6 || v == 8;
        case PFH_TUNE_BLOCKED_NB: return v == 16 || v == 24 || v == 32;
        case PFH_TUNE_BLOCKED_NB_C128: return v == 8 || v == 16;
        case PFH_TUNE_SWITCH_N: return v >= 64 && v <= 256 && v % 32 == 0;
        case PFH_TUNE_GRAD_NB: return v == 16 || v == 32;
    }
    return false;
}

// Cached per-device workspace for callers that pass none.
struct WsCache {
    std::mutex mu;
    std::vector<void*> ptr;
    std::vector<size_t> bytes;
};
WsCache g_ws;

int cached_workspace(size_t need, void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    std::lock_guard<std::mutex> lk(g_ws.mu);
    if ((int)g_ws.ptr.size() <= dev) { g_ws.ptr.resize(dev + 1, nullptr); g_ws.bytes.resize(dev + 1, 0); }
    if (g_ws.bytes[dev] < need) {
        if (g_ws.ptr[dev]) {
            e = hipDeviceSynchronize();
            if (e != hipSuccess) return (int)e;
            (void)hipFree(g_ws.ptr[dev]);
            g_ws.ptr[dev] = nullptr;
            g_ws.bytes[dev] = 0;
        }
        e = hipMalloc(&g_ws.ptr[dev], need);
        if (e != hipSuccess) { g_ws.ptr[dev] = nullptr; return (int)e; }
        g_ws.bytes[dev] = need;
    }
    *out = g_ws.ptr[dev];
    return 0;
}

int resolve_ws(void* ws, size_t ws_bytes, size_t need, void** out) {
    if (need == 0) { *out = ws; return 0; }
    if (ws) {
        if (ws_bytes < need) return PFH_EWORKSPACE;
        *out = ws;
        return 0;
    }
    return cached_workspace(need, out);
}

// ---------------------------------------------------------------------------
// ragged bucketing: counting sort of matrix indices by size, largest first
// ---------------------------------------------------------------------------
constexpr int kMaxBuckets = 8;

// zero the size histogram and the buckets' work counters (a kernel, not a
// memset, so that a captured graph clears them on every replay)
__global__ void rg_zero(int32_t* hist, int nbins, int32_t* queue, int nq) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nbins; i += gridDim.x * blockDim.x) hist[i] = 0;
    if (blockIdx.x == 0 && (int)threadIdx.x < nq) queue[threadIdx.x] = 0;
}

__global__ void rg_hist(const int32_t* sizes, int64_t count, int32_t max_n, int32_t* hist) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        int s = sizes[i];
        s = s < 0 ? 0 : (s > max_n ? max_n : s);
        atomicAdd(&hist[s], 1);
    }
}

// One thread: positions in descending-size order and bucket ranges.
// Bucket b holds sizes in (edge[b-1], edge[b]] (edge[-1] = -1).
// Exclusive scan of the size histogram in descending size order (largest
// first), then each bucket's [begin, begin + count) range of the order list.
// 256 threads: each sums a contiguous run of bins, one thread scans the 256
// partial sums in LDS, then every thread writes its run's cursors.
__global__ void __launch_bounds__(256) rg_scan(const int32_t* hist, int32_t max_n, const int32_t* edges, int nb,
                                               int32_t* cursor, int32_t* bucket_begin, int32_t* bucket_count) {
    __shared__ int32_t part[256];
    __shared__ int32_t total;
    const int t = threadIdx.x;
    const int nbins = max_n + 1;
    const int per = (nbins + 255) / 256;
    const int i0 = t * per, i1 = (i0 + per < nbins) ? i0 + per : nbins;   // descending positions i -> bin max_n - i
    int32_t sum = 0;
    for (int i = i0; i < i1; ++i) sum += hist[max_n - i];
    part[t] = sum;
    __syncthreads();
    if (t == 0) {
        int32_t run = 0;
        for (int u = 0; u < 256; ++u) {
            const int32_t v = part[u];
            part[u] = run;
            run += v;
        }
        total = run;
    }
    __syncthreads();
    int32_t pos = part[t];
    for (int i = i0; i < i1; ++i) {
        const int s = max_n - i;
        cursor[s] = pos;
        pos += hist[s];
    }
    __syncthreads();
    if (t == 0) {
        for (int b = 0; b < nb; ++b) {
            const int hi = edges[b] > max_n ? max_n : edges[b];
            const int lo = b == 0 ? -1 : edges[b - 1];
            if (hi <= lo) { bucket_begin[b] = 0; bucket_count[b] = 0; continue; }
            const int32_t begin = cursor[hi];
            const int32_t end = (lo >= 0) ? cursor[lo] : total;
            bucket_begin[b] = begin;
            bucket_count[b] = end - begin;
        }
    }
}

__global__ void rg_scatter(const int32_t* sizes, int64_t count, int32_t max_n, int32_t* cursor, int32_t* order) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        int s = sizes[i];
        s = s < 0 ? 0 : (s > max_n ? max_n : s);
        const int slot = atomicAdd(&cursor[s], 1);
        order[slot] = (int32_t)i;
    }
}

// Register-kernel bucket edges per dtype (the capacities of pf_reg.hip).
int bucket_edges(int dt, int32_t* e) {
    switch (dt) {
        case PFH_F64: e[0] = 32; e[1] = 64; e[2] = 128; e[3] = 192; e[4] = 256; return 5;
        case PFH_F32: e[0] = 32; e[1] = 64; e[2] = 128; e[3] = 256; return 4;
        case PFH_C64: e[0] = 32; e[1] = 64; e[2] = 128; e[3] = 192; return 4;
        case PFH_C128: e[0] = 32; e[1] = 64; e[2] = 128; e[3] = 192; e[4] = 256; return 5;
    }
    return 0;
}

struct RaggedLayout {
    size_t hist, cursor, order, begin, cnt, edges, queue, spill, bat, slabs, total;
    int glb_grid;
    int64_t slab_elems;
    int64_t bat_elems;   // doubles of slab pairs for the batched blocked kernel
};

int glb_grid() { return 2 * device_cu_count(); }

RaggedLayout ragged_layout(int dt, int64_t count, int32_t max_n) {
    RaggedLayout L{};
    size_t off = 0;
    L.hist = off; off = align_up(off + sizeof(int32_t) * (max_n + 1));
    L.cursor = off; off = align_up(off + sizeof(int32_t) * (max_n + 1));
    L.order = off; off = align_up(off + sizeof(int32_t) * (count > 0 ? count : 1));
    L.begin = off; off = align_up(off + sizeof(int32_t) * (kMaxBuckets + 1));
    L.cnt = off; off = align_up(off + sizeof(int32_t) * (kMaxBuckets + 1));
    L.edges = off; off = align_up(off + sizeof(int32_t) * (kMaxBuckets + 1));
    L.queue = off; off = align_up(off + sizeof(int32_t) * (kMaxBuckets + 1));
    L.spill = off;
    if (dt == PFH_F64 && max_n > lds_max_n())   // spill slabs of the LDS kernel's 192 < n <= 256 bucket
        off = align_up(off + sizeof(double) * (size_t)lds_spill_elems() * lds_spill_slabs());
    L.bat = off;
    L.bat_elems = 0;
    if ((dt == PFH_F64 || dt == PFH_C128) && max_n > 64) {   // slab pairs of the batched blocked kernel
        L.bat_elems = bat_ws_elems(dt, max_n, count);
        off = align_up(off + sizeof(double) * (size_t)L.bat_elems);
    }
    L.slabs = off;
    L.glb_grid = 0;
    L.slab_elems = 0;
    if (max_n > reg_max_n(dt)) {
        // one slab per workgroup of the global-slab kernel, which strides
        // over the bucket: no more workgroups than matrices
        const int64_t g = glb_grid();
        L.glb_grid = (int)(count < g ? (count > 0 ? count : 1) : g);
        L.slab_elems = (int64_t)max_n * max_n;
        off = align_up(off + dtype_bytes(dt) * (size_t)L.slab_elems * L.glb_grid);
    }
    L.total = off;
    return L;
}

__global__ void rg_set_edges(int32_t* dst, int32_t e0, int32_t e1, int32_t e2, int32_t e3, int32_t e4, int32_t e5) {
    if (threadIdx.x == 0) { dst[0] = e0; dst[1] = e1; dst[2] = e2; dst[3] = e3; dst[4] = e4; dst[5] = e5; }
}

template <typename T>
int ragged_typed(int dt, const void* data, const int64_t* offsets, const int32_t* sizes, int64_t count,
                 int32_t max_n, double* log_out, void* phase_out, char* ws, const RaggedLayout& L,
                 hipStream_t s) {
    int32_t* hist = (int32_t*)(ws + L.hist);
    int32_t* cursor = (int32_t*)(ws + L.cursor);
    int32_t* order = (int32_t*)(ws + L.order);
    int32_t* bbegin = (int32_t*)(ws + L.begin);
    int32_t* bcount = (int32_t*)(ws + L.cnt);
    int32_t* edges = (int32_t*)(ws + L.edges);
    int32_t* queue = (int32_t*)(ws + L.queue);

    int32_t e[kMaxBuckets];
    int nb = bucket_edges(dt, e);
    const int rmax = reg_max_n(dt);
    int ne = nb;
    if (max_n > rmax) e[ne++] = max_n;   // the global-slab bucket
    for (int i = ne; i < 6; ++i) e[i] = e[ne - 1];
    hipLaunchKernelGGL(rg_set_edges, dim3(1), dim3(64), 0, s, edges, e[0], e[1], e[2], e[3], e[4], e[5]);
    hipLaunchKernelGGL(rg_zero, dim3(1), dim3(256), 0, s, hist, max_n + 1, queue, kMaxBuckets + 1);
    const int g = (int)((count + 255) / 256 > 1024 ? 1024 : (count + 255) / 256);
    hipLaunchKernelGGL(rg_hist, dim3(g), dim3(256), 0, s, sizes, count, max_n, hist);
    hipLaunchKernelGGL(rg_scan, dim3(1), dim3(256), 0, s, hist, max_n, edges, ne, cursor, bbegin, bcount);
    hipLaunchKernelGGL(rg_scatter, dim3(g), dim3(256), 0, s, sizes, count, max_n, cursor, order);
    int err = (int)hipGetLastError();
    if (err) return err;

    // Largest buckets first so their long tails overlap the small ones'
    // launches as little as possible on the one stream.
    for (int b = ne - 1; b >= 0; --b) {
        const int lo = b == 0 ? 0 : e[b - 1];
        if (lo >= max_n) continue;          // bucket entirely above max_n: empty
        BatchDesc d{};
        d.data = data; d.offsets = offsets; d.sizes = sizes;
        d.order = order; d.begin_dev = bbegin + b; d.count_dev = bcount + b;
        d.count = count; d.dense_n = 0;
        d.log_out = log_out; d.phase_out = phase_out;
        if (dt == PFH_F64 && max_n > lds_max_n()) {
            d.slab = (double*)(ws + L.spill);
            d.slab_elems = lds_spill_elems();
        }
        if ((dt == PFH_F64 || dt == PFH_C128) && max_n > 64) {
            d.bslab = (double*)(ws + L.bat);
            d.bslab_elems = L.bat_elems;
            d.queue = queue + b;            // used by the batched blocked kernel
        }
        if (b >= nb) {
            err = launch_glb<T>(d, (T*)(ws + L.slabs), L.slab_elems, L.glb_grid, s);
        } else {
            err = launch_reg(dt, MODE_SLOG, d, e[b], s);
        }
        if (err) return err;
    }
    return 0;
}

int check_dense(int64_t batch, int32_t n) {
    if (batch < 0 || n < 0) return PFH_EINVAL;
    return 0;
}

template <typename T>
int pfaffian_dense(int dt, const void* A, int64_t batch, int32_t n, void* out, hipStream_t s) {
    if (check_dense(batch, n)) return PFH_EINVAL;
    if (n > 64) return PFH_ETOOBIG;
    if (batch == 0) return 0;
    if (!A || !out) return PFH_EINVAL;
    BatchDesc d{};
    d.data = A; d.count = batch; d.dense_n = n; d.prod_out = out;
    return launch_reg(dt, MODE_PROD, d, n < 2 ? 2 : n, s);
}

// Workspace of an f64 slog of a dense batch at n <= 256: the batched
// blocked kernel's slab pairs (n > 64), then the LDS kernel's spill slabs
// (n > 192).
size_t dense_bat_bytes(int dt, int64_t batch, int32_t n) {
    if ((dt != PFH_F64 && dt != PFH_C128) || n <= 64 || n > reg_max_n(dt)) return 0;
    return align_up(sizeof(double) * (size_t)bat_ws_elems(dt, n, batch));
}
size_t dense_spill_bytes(int dt, int32_t n) {
    if (dt != PFH_F64 || n <= lds_max_n() || n > reg_max_n(dt)) return 0;
    return sizeof(double) * (size_t)lds_spill_elems() * lds_spill_slabs();
}
size_t dense_ws_bytes(int dt, int64_t batch, int32_t n) { return dense_bat_bytes(dt, batch, n) + dense_spill_bytes(dt, n); }

// The automatic method's choice for a dense call: the batched kernels up to
// their capacity, except that a call of fewer matrices than the device has
// CUs (which leaves most CUs idle in the batched kernels) takes the blocked
// path above PFH_TUNE_SWITCH_N (built-in: PFH_SWITCH_N_DEFAULT).
#ifndef PFH_SWITCH_N_DEFAULT
#define PFH_SWITCH_N_DEFAULT 256
#endif
bool dense_uses_batched(int dt, int64_t batch, int32_t n) {
    if (n > reg_max_n(dt)) return false;
    if (n <= 64 || batch >= device_cu_count()) return true;
    const int tv = tune_value(PFH_TUNE_SWITCH_N);
    return n <= (tv >= 64 ? tv : PFH_SWITCH_N_DEFAULT);
}

size_t slog_ws_bytes(int dt, int64_t batch, int32_t n, int method) {
    if (method == PFH_METHOD_AUTO && dense_uses_batched(dt, batch, n)) return dense_ws_bytes(dt, batch, n);
    const size_t es = dtype_bytes(dt);
    if (method == PFH_METHOD_BLOCKED || method == PFH_METHOD_MIXED || method == PFH_METHOD_AUTO)
        return blocked_workspace_bytes(dt, n, method == PFH_METHOD_MIXED);
    return align_up(es * (size_t)n * n) + align_up(sizeof(LargeState<c128>));
}

template <typename T>
int slog_dense(int dt, const T* A, int64_t batch, int32_t n, int method, double* log_out, T* phase_out,
               void* ws_in, size_t ws_bytes, hipStream_t s) {
    if (check_dense(batch, n)) return PFH_EINVAL;
    if (method < PFH_METHOD_AUTO || method > PFH_METHOD_MIXED) return PFH_EINVAL;
    if (batch == 0) return 0;
    if (!A || !log_out || !phase_out) return PFH_EINVAL;
    if (method == PFH_METHOD_AUTO && dense_uses_batched(dt, batch, n)) {
        BatchDesc d{};
        d.data = A; d.count = batch; d.dense_n = n; d.log_out = log_out; d.phase_out = phase_out;
        const size_t sb = dense_ws_bytes(dt, batch, n);
        if (sb) {
            void* ws = nullptr;
            int err = resolve_ws(ws_in, ws_bytes, sb, &ws);
            if (err) return err;
            const size_t bb = dense_bat_bytes(dt, batch, n);
            if (bb) {
                d.bslab = (double*)ws;
                d.bslab_elems = bat_ws_elems(dt, n, batch);
            }
            if (dense_spill_bytes(dt, n)) {
                d.slab = (double*)((char*)ws + bb);
                d.slab_elems = lds_spill_elems();
            }
        }
        return launch_reg(dt, MODE_SLOG, d, n < 2 ? 2 : n, s);
    }
    if (n == 0 || (n & 1)) {
        // odd/empty: one register launch writes (-inf, 0)
        BatchDesc d{};
        d.data = A; d.count = batch; d.dense_n = n; d.log_out = log_out; d.phase_out = phase_out;
        return launch_reg(dt, MODE_SLOG, d, 2, s);
    }
    void* ws = nullptr;
    const size_t need = slog_ws_bytes(dt, batch, n, method);
    int err = resolve_ws(ws_in, ws_bytes, need, &ws);
    if (err) return err;
    const bool blocked = (method != PFH_METHOD_UNBLOCKED);
    const int nb = (method == PFH_METHOD_MIXED && dt == PFH_F64) ? 32 : blocked_panel_width(dt);
    for (int64_t b = 0; b < batch; ++b) {
        const T* Ab = A + b * (int64_t)n * n;
        if (blocked) {
            err = blocked_slog<T>(Ab, n, nb, method == PFH_METHOD_MIXED, ws, log_out + b, phase_out + b, s);
        } else {
            T* W = (T*)ws;
            LargeState<T>* st = (LargeState<T>*)((char*)ws + align_up(sizeof(T) * (size_t)n * n));
            err = large_slog<T>(Ab, n, W, st, log_out + b, phase_out + b, s);
        }
        if (err) return err;
    }
    return 0;
}

// Matrices past the register kernels: the blocked sweep when the device
// takes its cooperative panel, else the unblocked one.
size_t grad_ws_bytes(int dt, int32_t n) {
    if (n <= reg_grad_max_n(dt)) return 0;
    if (blocked_grad_enabled(n)) return blocked_grad_workspace_bytes(dt, n);
    return align_up(dtype_bytes(dt) * (size_t)n * n) + align_up(sizeof(int) * (size_t)n) +
           align_up(sizeof(LargeState<c128>));
}

template <typename T>
int grad_dense(int dt, const T* A, int64_t batch, int32_t n, double* log_out, T* phase_out, T* G,
               void* ws_in, size_t ws_bytes, hipStream_t s) {
    if (check_dense(batch, n)) return PFH_EINVAL;
    if (batch == 0) return 0;
    if (!A || !log_out || !phase_out || !G) return PFH_EINVAL;
    if (n <= reg_grad_max_n(dt) || (n & 1)) {
        if (n > reg_grad_max_n(dt)) return PFH_ETOOBIG;   // odd and too large: nothing to sweep
        BatchDesc d{};
        d.data = A; d.count = batch; d.dense_n = n; d.log_out = log_out; d.phase_out = phase_out;
        d.grad_out = G;
        return launch_reg(dt, MODE_GRAD, d, n < 2 ? 2 : n, s);
    }
    void* ws = nullptr;
    int err = resolve_ws(ws_in, ws_bytes, grad_ws_bytes(dt, n), &ws);
    if (err) return err;
    if (blocked_grad_enabled(n)) {
        for (int64_t i = 0; i < batch; ++i) {
            const int64_t off = i * (int64_t)n * n;
            err = blocked_grad<T>(A + off, n, ws, log_out + i, phase_out + i, G + off, s);
            if (err) return err;
        }
        return 0;
    }
    char* b = (char*)ws;
    T* W = (T*)b;
    int* perm = (int*)(b + align_up(sizeof(T) * (size_t)n * n));
    LargeState<T>* st = (LargeState<T>*)(b + align_up(sizeof(T) * (size_t)n * n) + align_up(sizeof(int) * (size_t)n));
    for (int64_t i = 0; i < batch; ++i) {
        const int64_t off = i * (int64_t)n * n;
        err = large_grad<T>(A + off, n, W, perm, st, log_out + i, phase_out + i, G + off, s);
        if (err) return err;
    }
    return 0;
}

// Device buffers of one host-pointer call, freed on every return path.
struct DevBufs {
    std::vector<void*> p;
    ~DevBufs() { for (void* q : p) (void)hipFree(q); }
    int alloc(size_t bytes, void** out) {
        void* q = nullptr;
        const hipError_t e = hipMalloc(&q, bytes ? bytes : 1);
        if (e != hipSuccess) return (int)e;
        p.push_back(q);
        *out = q;
        return 0;
    }
};

// ---------------------------------------------------------------------------
// Allocation-free ragged slog: the workspace is a fixed device array of the
// library's own, so a call makes no hipMalloc (for code under graph capture
// or loops that must not allocate).  Each device holds its own copy from the
// time the library's code is loaded on it until the process ends.  Host
// code never passes `scratch` itself to a kernel (in host code the name is a
// stand-in, not a device address): the address comes from
// hipGetSymbolAddress, once per device, with that device current.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = (size_t)256 << 20;   // PFH_NOALLOC_SCRATCH_BYTES
constexpr int32_t kNoallocMaxN = 256;                  // PFH_NOALLOC_MAX_N
constexpr int64_t kNoallocMaxCount = (int64_t)1 << 22; // PFH_NOALLOC_MAX_COUNT
static_assert(kScratchBytes == ((size_t)256 << 20), "the header states 256 MiB");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::atomic<void*> g_scratch_ptr[kMaxDevices];

int scratch_address(void** out) {
    const int dev = current_device();
    void* p = g_scratch_ptr[dev].load(std::memory_order_acquire);
    if (!p) {
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        if (!p) return (int)hipErrorInvalidSymbol;
        g_scratch_ptr[dev].store(p, std::memory_order_release);
    }
    *out = p;
    return 0;
}

// The ragged layout with the batched blocked kernel's slab pairs cut to the
// bytes left in `cap` (its grid shrinks to the pairs it is given): false when
// even the fixed part, or one slab pair of the widest bucket, does not fit.
bool ragged_layout_capped(int dt, int64_t count, int32_t max_n, size_t cap, RaggedLayout* out) {
    RaggedLayout L = ragged_layout(dt, count, max_n);
    if (L.glb_grid != 0) return false;                 // sizes past the register kernels: n^2 slabs
    if (L.bat_elems > 0) {
        if (L.bat > cap) return false;
        const int64_t room = (int64_t)((cap - L.bat) / sizeof(double));
        const int64_t np = max_n > 192 ? 256 : (max_n > 128 ? 192 : 128);   // the widest bucket's slabs
        const int64_t need = 2LL * np * np * (dt == PFH_C128 ? 2 : 1);      // one pair of them
        if (room < need) return false;
        if (L.bat_elems > room) L.bat_elems = room;
        L.slabs = align_up(L.bat + sizeof(double) * (size_t)L.bat_elems);
        L.total = L.slabs;
    }
    if (L.total > cap) return false;
    *out = L;
    return true;
}

}  // namespace

int pfh::tune_value(int key) {
    if (key < 0 || key >= PFH_TUNE_NKEYS) return -1;
    return g_tune[key].load(std::memory_order_relaxed) - 1;
}

extern "C" {

int pfh_set_tuning(int key, int value) {
    if (key < 0 || key >= PFH_TUNE_NKEYS || !tune_value_ok(key, value)) return PFH_EINVAL;
    g_tune[key].store(value + 1, std::memory_order_relaxed);
    return 0;
}

int pfh_get_tuning(int key, int* value) {
    if (key < 0 || key >= PFH_TUNE_NKEYS || !value) return PFH_EINVAL;
    *value = tune_value(key);
    return 0;
}

int pfh_slog_pfaffian_grad_workspace_size(int dtype, int64_t batch, int32_t n, size_t* bytes) {
    if (!bytes || dtype_bytes(dtype) == 0 || batch < 0 || n < 0) return PFH_EINVAL;
    *bytes = grad_ws_bytes(dtype, n);
    return 0;
}

int pfh_slog_pfaffian_grad(int dtype, const void* A, int64_t batch, int32_t n, double* log_out, void* phase_out,
                           void* grad_out, void* ws, size_t ws_bytes, hipStream_t s) {
    switch (dtype) {
        case PFH_F32: return grad_dense<float>(dtype, (const float*)A, batch, n, log_out, (float*)phase_out, (float*)grad_out, ws, ws_bytes, s);
        case PFH_F64: return grad_dense<double>(dtype, (const double*)A, batch, n, log_out, (double*)phase_out, (double*)grad_out, ws, ws_bytes, s);
        case PFH_C64: return grad_dense<c64>(dtype, (const c64*)A, batch, n, log_out, (c64*)phase_out, (c64*)grad_out, ws, ws_bytes, s);
        case PFH_C128: return grad_dense<c128>(dtype, (const c128*)A, batch, n, log_out, (c128*)phase_out, (c128*)grad_out, ws, ws_bytes, s);
    }
    return PFH_EINVAL;
}


int pfh_pfaffian_host(int dtype, const void* h_A, int32_t n, void* h_out) {
    const size_t es = dtype_bytes(dtype);
    if (!es || n < 0 || !h_A || !h_out) return PFH_EINVAL;
    if (n > 64) return PFH_ETOOBIG;
    DevBufs b;
    void *dA = nullptr, *dOut = nullptr;
    int err = b.alloc(es * (size_t)n * n, &dA);
    if (!err) err = b.alloc(es, &dOut);
    if (err) return err;
    hipError_t e = hipMemcpy(dA, h_A, es * (size_t)n * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) return (int)e;
    switch (dtype) {
        case PFH_F32: err = pfaffian_dense<float>(dtype, dA, 1, n, dOut, nullptr); break;
        case PFH_F64: err = pfaffian_dense<double>(dtype, dA, 1, n, dOut, nullptr); break;
        case PFH_C64: err = pfaffian_dense<c64>(dtype, dA, 1, n, dOut, nullptr); break;
        default: err = pfaffian_dense<c128>(dtype, dA, 1, n, dOut, nullptr); break;
    }
    if (err) return err;
    e = hipMemcpy(h_out, dOut, es, hipMemcpyDeviceToHost);
    return (int)e;
}

int pfh_slog_pfaffian_host(int dtype, const void* h_A, int32_t n, double* h_log, void* h_phase) {
    const size_t es = dtype_bytes(dtype);
    if (!es || n < 0 || !h_A || !h_log || !h_phase) return PFH_EINVAL;
    DevBufs b;
    void *dA = nullptr, *dLog = nullptr, *dPh = nullptr, *ws = nullptr;
    const size_t wsb = slog_ws_bytes(dtype, 1, n, PFH_METHOD_AUTO);
    int err = b.alloc(es * (size_t)n * n, &dA);
    if (!err) err = b.alloc(sizeof(double), &dLog);
    if (!err) err = b.alloc(es, &dPh);
    if (!err && wsb) err = b.alloc(wsb, &ws);
    if (err) return err;
    hipError_t e = hipMemcpy(dA, h_A, es * (size_t)n * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) return (int)e;
    err = pfh_slog_pfaffian_method(dtype, dA, 1, n, PFH_METHOD_AUTO, (double*)dLog, dPh, ws, wsb, nullptr);
    if (err) return err;
    e = hipMemcpy(h_log, dLog, sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(h_phase, dPh, es, hipMemcpyDeviceToHost);
    return (int)e;
}

const char* pfh_status_string(int status) {
    switch (status) {
        case PFH_OK: return "success";
        case PFH_EINVAL: return "invalid argument";
        case PFH_EWORKSPACE: return "workspace too small";
        case PFH_ETOOBIG: return "matrix size beyond this entry point";
        case PFH_ETIMEOUT: return "a blocked panel gave up waiting for its blocks (that call's results are NaN)";
    }
    return hipGetErrorString((hipError_t)status);
}

int pfh_pfaffian_max_n(void) { return 64; }

int pfh_stream_status(hipStream_t s) {
    int timed_out = 0;
    const int err = blocked_take_status(s, &timed_out);
    if (err) return err;
    return timed_out ? PFH_ETIMEOUT : PFH_OK;
}

int pfh_pfaffian_f32(const float* A, int64_t batch, int32_t n, float* out, hipStream_t s) {
    return pfaffian_dense<float>(PFH_F32, A, batch, n, out, s);
}
int pfh_pfaffian_f64(const double* A, int64_t batch, int32_t n, double* out, hipStream_t s) {
    return pfaffian_dense<double>(PFH_F64, A, batch, n, out, s);
}
int pfh_pfaffian_c64(const void* A, int64_t batch, int32_t n, void* out, hipStream_t s) {
    return pfaffian_dense<c64>(PFH_C64, A, batch, n, out, s);
}
int pfh_pfaffian_c128(const void* A, int64_t batch, int32_t n, void* out, hipStream_t s) {
    return pfaffian_dense<c128>(PFH_C128, A, batch, n, out, s);
}

int pfh_slog_pfaffian_workspace_size(int dtype, int64_t batch, int32_t n, int method, size_t* bytes) {
    if (!bytes || dtype_bytes(dtype) == 0 || batch < 0 || n < 0) return PFH_EINVAL;
    *bytes = slog_ws_bytes(dtype, batch, n, method);
    return 0;
}

int pfh_slog_pfaffian_method(int dtype, const void* A, int64_t batch, int32_t n, int method,
                             double* log_out, void* phase_out, void* ws, size_t ws_bytes, hipStream_t s) {
    switch (dtype) {
        case PFH_F32: return slog_dense<float>(dtype, (const float*)A, batch, n, method, log_out, (float*)phase_out, ws, ws_bytes, s);
        case PFH_F64: return slog_dense<double>(dtype, (const double*)A, batch, n, method, log_out, (double*)phase_out, ws, ws_bytes, s);
        case PFH_C64: return slog_dense<c64>(dtype, (const c64*)A, batch, n, method, log_out, (c64*)phase_out, ws, ws_bytes, s);
        case PFH_C128: return slog_dense<c128>(dtype, (const c128*)A, batch, n, method, log_out, (c128*)phase_out, ws, ws_bytes, s);
    }
    return PFH_EINVAL;
}

int pfh_slog_pfaffian_f32(const float* A, int64_t batch, int32_t n, double* log_out, float* sign_out,
                          void* ws, size_t ws_bytes, hipStream_t s) {
    return pfh_slog_pfaffian_method(PFH_F32, A, batch, n, PFH_METHOD_AUTO, log_out, sign_out, ws, ws_bytes, s);
}
int pfh_slog_pfaffian_f64(const double* A, int64_t batch, int32_t n, double* log_out, double* sign_out,
                          void* ws, size_t ws_bytes, hipStream_t s) {
    return pfh_slog_pfaffian_method(PFH_F64, A, batch, n, PFH_METHOD_AUTO, log_out, sign_out, ws, ws_bytes, s);
}
int pfh_slog_pfaffian_c64(const void* A, int64_t batch, int32_t n, double* log_out, void* phase_out,
                          void* ws, size_t ws_bytes, hipStream_t s) {
    return pfh_slog_pfaffian_method(PFH_C64, A, batch, n, PFH_METHOD_AUTO, log_out, phase_out, ws, ws_bytes, s);
}
int pfh_slog_pfaffian_c128(const void* A, int64_t batch, int32_t n, double* log_out, void* phase_out,
                           void* ws, size_t ws_bytes, hipStream_t s) {
    return pfh_slog_pfaffian_method(PFH_C128, A, batch, n, PFH_METHOD_AUTO, log_out, phase_out, ws, ws_bytes, s);
}

int pfh_slog_pfaffian_ragged_workspace_size(int dtype, int64_t count, int32_t max_n, size_t* bytes) {
    if (!bytes || dtype_bytes(dtype) == 0 || count < 0 || max_n < 0) return PFH_EINVAL;
    *bytes = ragged_layout(dtype, count, max_n).total;
    return 0;
}

int pfh_slog_pfaffian_ragged(int dtype, const void* data, const int64_t* offsets, const int32_t* sizes,
                             int64_t count, int32_t max_n, double* log_out, void* phase_out,
                             void* ws_in, size_t ws_bytes, hipStream_t s) {
    if (dtype_bytes(dtype) == 0 || count < 0 || max_n < 0) return PFH_EINVAL;
    if (count == 0) return 0;
    if (count > 0x7fffffff) return PFH_EINVAL;
    if (!data || !offsets || !sizes || !log_out || !phase_out) return PFH_EINVAL;
    const RaggedLayout L = ragged_layout(dtype, count, max_n);
    void* ws = nullptr;
    int err = resolve_ws(ws_in, ws_bytes, L.total, &ws);
    if (err) return err;
    switch (dtype) {
        case PFH_F32: return ragged_typed<float>(dtype, data, offsets, sizes, count, max_n, log_out, phase_out, (char*)ws, L, s);
        case PFH_F64: return ragged_typed<double>(dtype, data, offsets, sizes, count, max_n, log_out, phase_out, (char*)ws, L, s);
        case PFH_C64: return ragged_typed<c64>(dtype, data, offsets, sizes, count, max_n, log_out, phase_out, (char*)ws, L, s);
        case PFH_C128: return ragged_typed<c128>(dtype, data, offsets, sizes, count, max_n, log_out, phase_out, (char*)ws, L, s);
    }
    return PFH_EINVAL;
}

int pfh_slog_pfaffian_ragged_noalloc(int dtype, const void* data, const int64_t* offsets, const int32_t* sizes,
                                     int64_t count, int32_t max_n, double* log_out, void* phase_out,
                                     hipStream_t s) {
    if (dtype_bytes(dtype) == 0 || count < 0 || max_n < 0) return PFH_EINVAL;
    if (count == 0) return 0;
    if (!data || !offsets || !sizes || !log_out || !phase_out) return PFH_EINVAL;
    if (max_n > kNoallocMaxN || count > kNoallocMaxCount) return PFH_ETOOBIG;
    RaggedLayout L{};
    if (!ragged_layout_capped(dtype, count, max_n, kScratchBytes, &L)) return PFH_ETOOBIG;
    void* ws = nullptr;
    const int err = scratch_address(&ws);
    if (err) return err;
    switch (dtype) {
        case PFH_F32: return ragged_typed<float>(dtype, data, offsets, sizes, count, max_n, log_out, phase_out, (char*)ws, L, s);
        case PFH_F64: return ragged_typed<double>(dtype, data, offsets, sizes, count, max_n, log_out, phase_out, (char*)ws, L, s);
        case PFH_C64: return ragged_typed<c64>(dtype, data, offsets, sizes, count, max_n, log_out, phase_out, (char*)ws, L, s);
        case PFH_C128: return ragged_typed<c128>(dtype, data, offsets, sizes, count, max_n, log_out, phase_out, (char*)ws, L, s);
    }
    return PFH_EINVAL;
}

}  // extern "C"
