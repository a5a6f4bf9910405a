// TEST INFRASTRUCTURE ONLY. This file is the CPU oracle that checks the HIP
// product path; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it. Nothing in pfcuda_amd/ links or calls it.
//
// A C++ restatement of the reference's FP64 CPU Pfaffian
// (src/pfaffian_cpu.cpp:7-106): left-looking Parlett-Reid elimination on the
// upper triangle of a row-major matrix, pairs of rows eliminated per step,
// the eliminated rows reused as storage for the factors.
//
// Departures from the reference, each needed for its use as an oracle:
//   * the result is also produced as (log|Pf|, sign/phase), because the
//     product form overflows past |Pf| ~ 1e308 (src/pfaffian_cpu.cpp:83);
//     log|p| is accumulated with a Neumaier compensated sum;
//   * the scalar type is a template: double, long double (an extended
//     precision "truth" run) and complex over both (the reference is real
//     only and force-casts, bindings/pybind_bindings.cpp:8);
//   * it works on a copy, never on the caller's buffer;
//   * `stale` selects the reference's pivot rule bit for bit: the pivot
//     search starts from |A[k][k+1]| read before row k is updated
//     (src/pfaffian_cpu.cpp:29-30 vs :47-52). With stale=0 the search runs
//     over the updated row, as pfcuda/pfaffian_py.py:32 and the GPU
//     reference do. Ties keep the lowest column (strict '>' as :49).
//   * for complex inputs the pivot magnitude is |z| (std::abs), as np.abs
//     in pfcuda/pfaffian_py.py:32.
// The zero-pivot test is exact (== 0), as src/pfaffian_cpu.cpp:79 and
// pfcuda/pfaffian_py.py:34.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>
#include <atomic>
#include <algorithm>
#include <functional>

namespace {

template <typename S> struct Real { using type = S; };
template <typename R> struct Real<std::complex<R>> { using type = R; };

template <typename R> inline R mag(R x) { return std::fabs(x); }
template <typename R> inline R mag(const std::complex<R>& z) { return std::abs(z); }

template <typename R> inline R unit_phase(R x) { return x > 0 ? R(1) : R(-1); }
template <typename R> inline std::complex<R> unit_phase(const std::complex<R>& z) {
    R m = std::abs(z);
    return std::complex<R>(z.real() / m, z.imag() / m);
}

template <typename R> struct Neumaier {
    R s = 0, c = 0;
    void add(R x) {
        R t = s + x;
        if (std::fabs(s) >= std::fabs(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    R value() const { return s + c; }
};

template <typename S> struct PfResult {
    using R = typename Real<S>::type;
    S prod;       // product form, as src/pfaffian_cpu.cpp returns
    R logabs;     // log|Pf|
    S phase;      // sign (real) or unit phase (complex); 0 when Pf == 0
    int info;     // 0, or 1 + the pair step k at which an exact zero pivot stopped it
};

// Row-major element access, as RM() in include/pfaffian_cpu.h:3.
#define ORM(A, n, i, j) ((A)[(size_t)(i) * (size_t)(n) + (size_t)(j)])

// Row `row` (k or k+1) at columns [c0, c1) gets the corrections of the
// stored pairs r < k (src/pfaffian_cpu.cpp:32-45 and :85-98):
//   A[row][c] += A[r][c] * A[r+1][row] - A[r+1][c] * A[r][row].
// Every column is independent, so a column range can go to its own thread
// with each element's operations in the same order as one thread's.
// Complex values are multiplied as (ac - bd, ad + bc), the value the
// compiler's complex product returns for finite operands, written out so
// the loop avoids its NaN-recovery call.
template <typename S>
inline void lazy_update(S* A, int n, int k, int row, int c0, int c1) {
    S* v = &ORM(A, n, row, 0);
    for (int r = 0; r < k; r += 2) {
        const S* s0 = &ORM(A, n, r, 0);      // x_r / a_r, scaled
        const S* s1 = &ORM(A, n, r + 1, 0);  // y_r
        const S f0 = s1[row], f1 = s0[row];
        if constexpr (std::is_floating_point<S>::value) {
            for (int c = c0; c < c1; ++c) v[c] += s0[c] * f0 - s1[c] * f1;
        } else {
            using R = typename Real<S>::type;
            const R f0r = f0.real(), f0i = f0.imag(), f1r = f1.real(), f1i = f1.imag();
            R* vp = reinterpret_cast<R*>(v);
            const R* p0 = reinterpret_cast<const R*>(s0);
            const R* p1 = reinterpret_cast<const R*>(s1);
            for (int c = c0; c < c1; ++c) {
                const R ar = p0[2 * c], ai = p0[2 * c + 1], br = p1[2 * c], bi = p1[2 * c + 1];
                const R t1r = ar * f0r - ai * f0i, t1i = ar * f0i + ai * f0r;
                const R t2r = br * f1r - bi * f1i, t2i = br * f1i + bi * f1r;
                vp[2 * c] = vp[2 * c] + (t1r - t2r);
                vp[2 * c + 1] = vp[2 * c + 1] + (t1i - t2i);
            }
        }
    }
}

template <typename S>
inline void scale_row(S* v, S inv, int c0, int c1) {
    if constexpr (std::is_floating_point<S>::value) {
        for (int c = c0; c < c1; ++c) v[c] *= inv;
    } else {
        using R = typename Real<S>::type;
        const R ir = inv.real(), ii = inv.imag();
        R* vp = reinterpret_cast<R*>(v);
        for (int c = c0; c < c1; ++c) {
            const R a = vp[2 * c], b = vp[2 * c + 1];
            vp[2 * c] = a * ir - b * ii;
            vp[2 * c + 1] = a * ii + b * ir;
        }
    }
}

// Runs fn(c0, c1) over [a, b) on the calling thread.
struct SerialFor {
    template <typename F> void operator()(int a, int b, F&& fn) const { if (a < b) fn(a, b); }
};

// A fixed set of threads that split a column range [a, b) into contiguous
// chunks, one per thread, and return when every chunk is done.  Each column
// is handled by exactly one thread with the same operations as the serial
// loop, so a run on this pool is bitwise equal to a serial run.
class ColumnPool {
public:
    explicit ColumnPool(int threads) : T_(threads < 1 ? 1 : threads) {
        for (int id = 1; id < T_; ++id) th_.emplace_back([this, id] { worker(id); });
    }
    ~ColumnPool() {
        stop_.store(true, std::memory_order_relaxed);
        gen_.fetch_add(1, std::memory_order_release);
        for (auto& t : th_) t.join();
    }
    template <typename F> void operator()(int a, int b, F&& fn) {
        if (a >= b) return;
        if (T_ == 1 || b - a < 256) { fn(a, b); return; }
        std::function<void(int, int)> f = fn;
        job_ = &f; a_ = a; b_ = b;
        done_.store(0, std::memory_order_relaxed);
        gen_.fetch_add(1, std::memory_order_release);
        run(0);
        while (done_.load(std::memory_order_acquire) != T_ - 1) pause();
    }
private:
    static void pause() { std::this_thread::yield(); }
    void run(int id) {
        const int len = b_ - a_;
        const int per = ((len + T_ - 1) / T_ + 7) & ~7;   // chunks of whole 64-byte lines
        const int c0 = a_ + id * per, c1 = std::min(b_, c0 + per);
        if (c0 < c1) (*job_)(c0, c1);
    }
    void worker(int id) {
        int seen = 0;
        for (;;) {
            int g;
            while ((g = gen_.load(std::memory_order_acquire)) == seen) pause();
            seen = g;
            if (stop_.load(std::memory_order_relaxed)) return;
            run(id);
            done_.fetch_add(1, std::memory_order_release);
        }
    }
    int T_;
    std::vector<std::thread> th_;
    std::atomic<int> gen_{0}, done_{0};
    std::atomic<bool> stop_{false};
    const std::function<void(int, int)>* job_ = nullptr;
    int a_ = 0, b_ = 0;
};

template <typename S, typename For = SerialFor>
PfResult<S> pfaffian_left_looking(S* A, int n, bool stale, For&& pfor = For()) {
    using R = typename Real<S>::type;
    PfResult<S> out{S(0), -INFINITY, S(0), 0};
    if (n == 0 || (n & 1)) return out;            // src/pfaffian_cpu.cpp:8-9

    S result = S(1);
    Neumaier<R> logsum;
    S phase = S(1);

    for (int k = 0; k < n - 1; k += 2) {
        S* v = &ORM(A, n, k, 0);       // row k, the pivot row
        int kp = k + 1;
        R kp_val = mag(v[kp]);         // stale start value (src/pfaffian_cpu.cpp:29-30)

        // Lazy update of row k from the stored pairs (src/pfaffian_cpu.cpp:32-45).
        pfor(k + 1, n, [&](int c0, int c1) { lazy_update<S>(A, n, k, k, c0, c1); });
        if (!stale) kp_val = mag(v[kp]);
        for (int c = k + 2; c < n; ++c) {      // src/pfaffian_cpu.cpp:47-52
            R cv = mag(v[c]);
            if (cv > kp_val) { kp = c; kp_val = cv; }
        }
        if (kp != k + 1) {                     // upper-triangle symmetric swap, :54-75
            const int i = k + 1, j = kp;
            for (int r = 0; r < i; ++r) std::swap(ORM(A, n, r, i), ORM(A, n, r, j));
            for (int c = j + 1; c < n; ++c) std::swap(ORM(A, n, i, c), ORM(A, n, j, c));
            for (int m = i + 1; m < j; ++m) {
                S t = ORM(A, n, i, m);
                ORM(A, n, i, m) = -ORM(A, n, m, j);
                ORM(A, n, m, j) = -t;
            }
            ORM(A, n, i, j) = -ORM(A, n, i, j);
            result *= S(-1);
            phase = -phase;
        }
        const S pivot = v[k + 1];
        if (pivot == S(0)) {                   // exact zero test, :79-80
            out.prod = S(0);
            out.info = k + 1;
            return out;
        }
        result *= pivot;
        logsum.add(std::log(mag(pivot)));
        phase *= unit_phase(pivot);
        if constexpr (!std::is_floating_point<S>::value) {
            R m = std::abs(phase);             // keep the running phase on the unit circle
            phase = S(phase.real() / m, phase.imag() / m);
        }
        const S inv = S(1) / pivot;

        // Lazy update of row k+1 (src/pfaffian_cpu.cpp:85-98), then scale row k (:100-102).
        pfor(k + 2, n, [&](int c0, int c1) {
            lazy_update<S>(A, n, k, k + 1, c0, c1);
            scale_row<S>(v, inv, c0, c1);
        });
    }
    out.prod = result;
    out.logabs = logsum.value();
    out.phase = phase;
    return out;
}

// The closed forms the reference uses for n = 2 and 4 (src/pfaffian_cpu.cpp:11-20).
template <typename S>
bool closed_form(const S* A, int n, S* prod) {
    if (n == 2) { *prod = ORM(A, n, 0, 1); return true; }
    if (n == 4) {
        *prod = ORM(A, n, 0, 1) * ORM(A, n, 2, 3) - ORM(A, n, 0, 2) * ORM(A, n, 1, 3)
              + ORM(A, n, 0, 3) * ORM(A, n, 1, 2);
        return true;
    }
    return false;
}

template <typename S>
PfResult<S> run_copy(const S* A, int n, bool stale, bool use_closed_forms) {
    std::vector<S> work((size_t)n * n);
    std::memcpy(work.data(), A, sizeof(S) * work.size());
    PfResult<S> r = pfaffian_left_looking<S>(work.data(), n, stale);
    S p;
    if (use_closed_forms && closed_form<S>(A, n, &p)) r.prod = p;
    return r;
}

template <typename S, typename Out>
void batch_run(const S* data, const int64_t* offsets, const int32_t* sizes, int64_t count,
               int threads, Out* logabs, Out* phase_out, int phase_stride) {
    std::atomic<int64_t> next{0};
    auto worker = [&]() {
        for (;;) {
            int64_t i = next.fetch_add(1);
            if (i >= count) break;
            PfResult<S> r = run_copy<S>(data + offsets[i], sizes[i], false, false);
            logabs[i] = (Out)r.logabs;
            if constexpr (std::is_floating_point<S>::value) {
                phase_out[i] = (Out)r.phase;
            } else {
                phase_out[2 * i] = (Out)r.phase.real();
                phase_out[2 * i + 1] = (Out)r.phase.imag();
            }
        }
    };
    if (threads <= 1) { worker(); return; }
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
}

}  // namespace

extern "C" {

// Real FP64. out3 = {prod, logabs, sign}. Returns info (0 or 1 + zero-pivot step).
int oracle_pf_f64(const double* A, int n, int stale, int closed_forms, double* out3) {
    PfResult<double> r = run_copy<double>(A, n, stale != 0, closed_forms != 0);
    out3[0] = r.prod; out3[1] = r.logabs; out3[2] = r.phase;
    return r.info;
}

// Extended precision (x87 long double) run of the same elimination on the
// same FP64 input. out3 = {prod, logabs, sign} rounded to double.
int oracle_pf_f64_ld(const double* A, int n, double* out3) {
    std::vector<long double> w((size_t)n * n);
    for (size_t i = 0; i < w.size(); ++i) w[i] = A[i];
    PfResult<long double> r = pfaffian_left_looking<long double>(w.data(), n, false);
    out3[0] = (double)r.prod; out3[1] = (double)r.logabs; out3[2] = (double)r.phase;
    return r.info;
}

// Complex128, interleaved (re, im). out5 = {prod.re, prod.im, logabs, phase.re, phase.im}.
int oracle_pf_c128(const double* A, int n, int stale, int closed_forms, double* out5) {
    using C = std::complex<double>;
    PfResult<C> r = run_copy<C>(reinterpret_cast<const C*>(A), n, stale != 0, closed_forms != 0);
    out5[0] = r.prod.real(); out5[1] = r.prod.imag(); out5[2] = r.logabs;
    out5[3] = r.phase.real(); out5[4] = r.phase.imag();
    return r.info;
}

int oracle_pf_c128_ld(const double* A, int n, double* out5) {
    using C = std::complex<long double>;
    std::vector<C> w((size_t)n * n);
    for (size_t i = 0; i < w.size(); ++i) w[i] = C(A[2 * i], A[2 * i + 1]);
    PfResult<C> r = pfaffian_left_looking<C>(w.data(), n, false);
    out5[0] = (double)r.prod.real(); out5[1] = (double)r.prod.imag(); out5[2] = (double)r.logabs;
    out5[3] = (double)r.phase.real(); out5[4] = (double)r.phase.imag();
    return r.info;
}

// The same eliminations with each pair step's column loops split across
// `threads` host threads (bitwise equal to the calls above; for the large
// single-matrix checks).
int oracle_pf_f64_mt(const double* A, int n, int threads, double* out3) {
    std::vector<double> w(A, A + (size_t)n * n);
    ColumnPool pool(threads);
    PfResult<double> r = pfaffian_left_looking<double>(w.data(), n, false, pool);
    out3[0] = r.prod; out3[1] = r.logabs; out3[2] = r.phase;
    return r.info;
}

int oracle_pf_c128_mt(const double* A, int n, int threads, double* out5) {
    using C = std::complex<double>;
    std::vector<C> w(reinterpret_cast<const C*>(A), reinterpret_cast<const C*>(A) + (size_t)n * n);
    ColumnPool pool(threads);
    PfResult<C> r = pfaffian_left_looking<C>(w.data(), n, false, pool);
    out5[0] = r.prod.real(); out5[1] = r.prod.imag(); out5[2] = r.logabs;
    out5[3] = r.phase.real(); out5[4] = r.phase.imag();
    return r.info;
}

// Ragged batches: matrix i starts at element offsets[i] and is sizes[i]^2
// row-major. One matrix per task across `threads` host threads.
void oracle_slog_ragged_f64(const double* data, const int64_t* offsets, const int32_t* sizes,
                            int64_t count, int threads, double* logabs, double* sign) {
    batch_run<double, double>(data, offsets, sizes, count, threads, logabs, sign, 1);
}

void oracle_slog_ragged_c128(const double* data, const int64_t* offsets, const int32_t* sizes,
                             int64_t count, int threads, double* logabs, double* phase2) {
    // offsets are in complex elements
    using C = std::complex<double>;
    batch_run<C, double>(reinterpret_cast<const C*>(data), offsets, sizes, count, threads,
                         logabs, phase2, 2);
}

}  // extern "C"
