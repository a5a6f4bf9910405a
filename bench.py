"""Benchmark: batched slog_pfaffian throughput on an uneven batch (one GPU).

Workload (the headline metric "batched Pf/s (N 16-256)"): 16384 float64
skew-symmetric matrices, N drawn uniformly from the even numbers 16..256,
entries N(0,1) skew-symmetrised A = (B - B^T)/2, seeded, generated on the GPU
and resident in HBM before timing.  One step = one pfh_slog_pfaffian_ragged
call over the whole batch (device-side size bucketing + one persistent launch
per bucket), the matrices packed back to back in input order.  Prints one
JSON line.

Also reported: the small batched product path (1024 x 32x32 f64, eager and
replayed from a captured graph, and the same kernel over 2^20 matrices); the
dominant kernel's roofline fraction (HIP events on the launch stream) with
its HBM bytes per launch from the committed rocprofv3 counter summary; the
whole batch with one production-kernel launch per matrix; single large
matrices (N = 4096 f64 and c128, plain vs blocked, against the threaded
oracle; the N = 8192 precision matrix Q diag(s_i J) Q^T, FP64 blocked and
split-precision, against the oracle and its closed form); the gradient
1/2 A^-T (512 x 128^2 and 2048^2); the accuracy of a sample against the FP64
CPU oracle and the oracle's own throughput over the whole batch on every
host core this process may use (cpu_baseline, about 10 s), with the oracle's
single-core time for one N = 4096 matrix.  --profile runs only the ragged
steps plus a 1 GiB device copy, the workload of the committed profiles.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK = 78.6e12      # MI355X FP64 (vector = matrix) FLOP/s
HBM_PEAK = 8.0e12        # MI355X HBM3E bytes/s


def gen_ragged(torch, count, lo, hi, seed, dtype, dev):
    """Sizes uniform over even lo..hi; matrix i at the next offset in input
    order (sizes mixed, as a caller's uneven batch comes), generated on the
    device one size at a time and scattered to its place."""
    rng = np.random.default_rng(seed)
    sizes = (2 * rng.integers(lo // 2, hi // 2 + 1, size=count)).astype(np.int32)
    offsets = np.zeros(count, np.int64)
    offsets[1:] = np.cumsum(sizes.astype(np.int64) ** 2)[:-1]
    total = int(offsets[-1] + int(sizes[-1]) ** 2)
    data = torch.empty(total, dtype=dtype, device=dev)
    g = torch.Generator(device=dev)
    for n in np.unique(sizes):
        idx = np.nonzero(sizes == n)[0]
        k = len(idx)
        g.manual_seed(seed * 1000003 + int(n))
        if dtype.is_complex:
            rdt = torch.float64 if dtype == torch.complex128 else torch.float32
            B = torch.complex(torch.randn(k, n, n, generator=g, device=dev, dtype=rdt),
                              torch.randn(k, n, n, generator=g, device=dev, dtype=rdt)) / math.sqrt(2)
        else:
            B = torch.randn(k, n, n, generator=g, device=dev, dtype=dtype)
        pos = (torch.from_numpy(offsets[idx]).to(dev)[:, None]
               + torch.arange(int(n) * int(n), device=dev)[None, :]).reshape(-1)
        data[pos] = ((B - B.transpose(1, 2)) / 2).reshape(-1)
        del B, pos
    return data, offsets, sizes


def host_cpu():
    """(threads this process may use, nproc, CPU model) of the host."""
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return usable, os.cpu_count() or usable, model


class Ragged:
    """Direct C-ABI caller with preallocated workspace and outputs."""

    def __init__(self, torch, lib, data, offsets, sizes, dt):
        self.torch, self.lib = torch, lib
        dev = data.device
        self.data = data
        self.off = torch.from_numpy(offsets).to(dev)
        self.siz = torch.from_numpy(sizes).to(dev)
        self.count = len(sizes)
        self.max_n = int(sizes.max())
        self.dt = dt
        self.log = torch.empty(self.count, dtype=torch.float64, device=dev)
        self.ph = torch.empty(self.count, dtype=data.dtype, device=dev)
        import ctypes
        need = ctypes.c_size_t(0)
        lib.pfh_slog_pfaffian_ragged_workspace_size(dt, self.count, self.max_n, need)
        self.ws = torch.empty(max(int(need.value), 1), dtype=torch.uint8, device=dev)
        self.stream = torch.cuda.current_stream(dev).cuda_stream

    def __call__(self):
        # the current stream at call time, so that a graph capture records it
        st = self.lib.pfh_slog_pfaffian_ragged(
            self.dt, self.data.data_ptr(), self.off.data_ptr(), self.siz.data_ptr(), self.count,
            self.max_n, self.log.data_ptr(), self.ph.data_ptr(), self.ws.data_ptr(), self.ws.numel(),
            self.torch.cuda.current_stream().cuda_stream)
        if st != 0:
            raise RuntimeError(f"pfh_slog_pfaffian_ragged status {st}")


def committed_traffic(kernel_key):
    """HBM bytes per launch of the dominant kernel, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (calibrated against a copy of
    known size); None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "ragged16384_f64.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch " + kernel_key + ":"):
                    return float(line.split(":")[-1])
    except OSError:
        pass
    return None


def time_events(torch, fn, reps):
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / 1e3 / reps


def calibration_copy(torch, dev):
    """A plain 1 GiB device copy: the byte counters' calibration dispatch."""
    src = torch.empty(1 << 27, dtype=torch.float64, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def profile_other(args, torch, pf, _lib, dev):
    """The --profile large / c1 workloads (no timing, no checks)."""
    g = torch.Generator(device=dev)
    if args.profile == "large":
        for n, dt in ((4096, torch.float64), (4096, torch.complex128)):
            g.manual_seed(args.seed + 7)
            if dt.is_complex:
                X = torch.complex(torch.randn(n, n, generator=g, device=dev, dtype=torch.float64),
                                  torch.randn(n, n, generator=g, device=dev, dtype=torch.float64)) / math.sqrt(2)
            else:
                X = torch.randn(n, n, generator=g, device=dev, dtype=torch.float64)
            A = ((X - X.T) / 2).contiguous()
            del X
            pf.slog_pfaffian(A, _lib.PFH_METHOD_BLOCKED)
            torch.cuda.synchronize()
            del A
        # the same construction built on the host (NumPy): rocBLAS, behind
        # torch's QR and products, does not initialize under counter collection
        rng = np.random.default_rng(args.seed + 8)
        Q, R = np.linalg.qr(rng.standard_normal((8192, 8192)))
        Q *= np.sign(np.diag(R))
        del R
        sv = 10.0 ** (rng.random(4096) * 6 - 3)
        QD = np.empty_like(Q)
        QD[:, 0::2] = -Q[:, 1::2] * sv
        QD[:, 1::2] = Q[:, 0::2] * sv
        M = QD @ Q.T
        del Q, QD
        A8 = torch.from_numpy((M - M.T) / 2).to(dev)
        del M
        pf.slog_pfaffian(A8, _lib.PFH_METHOD_BLOCKED)
        pf.slog_pfaffian(A8, _lib.PFH_METHOD_MIXED)
        torch.cuda.synchronize()
        del A8
        from pfcuda_amd.adjoint import slog_grad
        g.manual_seed(args.seed + 5)
        X = torch.randn(2048, 2048, generator=g, device=dev, dtype=torch.float64)
        slog_grad(((X - X.T) / 2).contiguous())
        torch.cuda.synchronize()
    else:
        for count in (1024, 1 << 20):
            g.manual_seed(args.seed + 1)
            B = torch.randn(count, 32, 32, generator=g, device=dev, dtype=torch.float64)
            A = ((B - B.transpose(1, 2)) / 2).contiguous()
            del B
            pf.pfaffian(A)
            torch.cuda.synchronize()
            del A
    calibration_copy(torch, dev)
    print(f"# profile mode {args.profile}: done, then one 1073741824-byte copy", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--count", type=int, default=16384)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile", nargs="?", const="ragged", choices=["ragged", "large", "c1"],
                    help="a short workload for profiler runs, then one 1 GiB device copy (the byte "
                         "counters' calibration): ragged = the timed ragged steps; large = one blocked "
                         "solve each of N = 4096 f64 and c128 and of the N = 8192 precision matrix (FP64 "
                         "and split precision), and one N = 2048 gradient; c1 = pfaffian() of 1024 and "
                         "2^20 32x32 f64 matrices")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the batch would shard across GPUs with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch
    import pfcuda_amd
    from pfcuda_amd import _lib

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    lib = _lib.lib
    strm = torch.cuda.current_stream(dev).cuda_stream
    if args.profile in ("large", "c1"):
        profile_other(args, torch, pfcuda_amd, _lib, dev)
        return
    t0 = time.time()
    data, offsets, sizes = gen_ragged(torch, args.count, 16, 256, args.seed, torch.float64, dev)
    torch.cuda.synchronize()
    print(f"# generated {args.count} matrices ({data.numel() * 8 / 1e9:.2f} GB) in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)

    run = Ragged(torch, lib, data, offsets, sizes, _lib.PFH_F64)
    for _ in range(args.warmup):
        run()
    torch.cuda.synchronize()
    ms = time_events(torch, run, args.steps) * 1e3
    value = args.count / (ms / 1e3)
    print(f"# ragged: {ms:.3f} ms/step", file=sys.stderr, flush=True)

    if args.profile:
        calibration_copy(torch, dev)
        print(f"# profile mode ragged: {args.steps} ragged steps and one 1073741824-byte copy done",
              file=sys.stderr, flush=True)
        return

    # the same call captured in a graph and replayed (checked against eager)
    want_log = run.log.clone()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        run()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        run()
    graph.replay()
    torch.cuda.synchronize()
    graph_ok = bool(torch.equal(run.log, want_log))
    ms_graph = time_events(torch, graph.replay, args.steps) * 1e3
    print(f"# ragged graph replay: {ms_graph:.3f} ms/step, equal to eager: {graph_ok}", file=sys.stderr, flush=True)
    del graph

    flops = float(np.sum(sizes.astype(np.float64) ** 3) / 3.0)
    tri_bytes = float(np.sum(sizes.astype(np.float64) * (sizes - 1) / 2.0) * 8)

    # dominant kernel: the largest bucket (192 < N <= 256), one pf_bat_kernel launch
    # plus the few-microsecond bucketing kernels of the ragged call
    big = np.nonzero(sizes > 192)[0]
    sub = Ragged(torch, lib, data, offsets[big], sizes[big], _lib.PFH_F64)
    sub()
    torch.cuda.synchronize()
    t_big = time_events(torch, sub, 5)
    big_flops = float(np.sum(sizes[big].astype(np.float64) ** 3) / 3.0)
    achieved = big_flops / t_big

    # per-matrix launches of the production kernels: the whole batch, one
    # pfh_slog_pfaffian_f64 call (batch 1, the automatic method) per matrix,
    # as a caller without the ragged entry point would run it
    import ctypes
    need = ctypes.c_size_t(0)
    lib.pfh_slog_pfaffian_workspace_size(_lib.PFH_F64, 1, int(sizes.max()), _lib.PFH_METHOD_AUTO, need)
    ws_pm = torch.empty(max(int(need.value), 1), dtype=torch.uint8, device=dev)
    lo_pm = torch.empty(args.count, dtype=torch.float64, device=dev)
    ph_pm = torch.empty(args.count, dtype=torch.float64, device=dev)
    base_ptr = data.data_ptr()
    calls = [(base_ptr + 8 * int(offsets[i]), int(sizes[i]), lo_pm.data_ptr() + 8 * i, ph_pm.data_ptr() + 8 * i)
             for i in range(args.count)]
    fn_pm = lib.pfh_slog_pfaffian_f64

    def per_matrix():
        for a_ptr, n_i, l_ptr, p_ptr in calls:
            st = fn_pm(a_ptr, 1, n_i, l_ptr, p_ptr, ws_pm.data_ptr(), ws_pm.numel(), strm)
            if st != 0:
                raise RuntimeError(f"per-matrix slog status {st}")
    per_matrix()
    torch.cuda.synchronize()
    t_pm = time_events(torch, per_matrix, 1)
    per_matrix_res = {"ms": t_pm * 1e3, "pf_per_s": args.count / t_pm,
                      "speedup_of_ragged_call": t_pm / (ms / 1e3),
                      "equal_to_ragged": bool(torch.equal(lo_pm, run.log) and torch.equal(ph_pm, run.ph)),
                      "note": "one pfh_slog_pfaffian_f64 launch per matrix (batch 1, automatic method), "
                              "the same kernels as the ragged call"}
    print(f"# per-matrix launches: {per_matrix_res}", file=sys.stderr, flush=True)
    del calls, ws_pm, lo_pm, ph_pm

    # the same batch shape in complex128
    datac, offc, sizc = gen_ragged(torch, args.count, 16, 256, args.seed, torch.complex128, dev)
    runc = Ragged(torch, lib, datac, offc, sizc, _lib.PFH_C128)
    runc()
    torch.cuda.synchronize()
    t_c = time_events(torch, runc, 3)
    c128_res = {"ms": t_c * 1e3, "pf_per_s": args.count / t_c,
                "tflops": float(np.sum(4.0 * sizc.astype(np.float64) ** 3 / 3.0)) / t_c / 1e12}
    if not args.no_cpu:
        # a sample gathered on the device and checked against the FP64 oracle
        import oracle
        kc = min(256, args.count)
        idc = np.random.default_rng(args.seed + 4).choice(args.count, size=kc, replace=False)
        parts = [datac[int(offc[i]):int(offc[i]) + int(sizc[i]) ** 2] for i in idc]
        hostc = torch.cat(parts).cpu().numpy()
        offs_c = np.cumsum([0] + [int(sizc[i]) ** 2 for i in idc[:-1]]).astype(np.int64)
        wl, wp = oracle.slog_ragged(hostc, offs_c, sizc[idc], threads=host_cpu()[0])
        gl = runc.log.cpu().numpy()[idc]
        gp = runc.ph.cpu().numpy()[idc]
        c128_res["max_abs_dlog"] = float(np.max(np.abs(gl - wl)))
        c128_res["max_abs_dphase"] = float(np.max(np.abs(gp - wp)))
        c128_res["checked"] = int(kc)
        del hostc, parts
    print(f"# c128 ragged: {c128_res}", file=sys.stderr, flush=True)
    del datac, runc

    # small batched product path: 1024 x 32x32 f64
    g = torch.Generator(device=dev)
    g.manual_seed(args.seed + 1)
    B = torch.randn(1024, 32, 32, generator=g, device=dev, dtype=torch.float64)
    A1 = ((B - B.transpose(1, 2)) / 2).contiguous()
    out1 = torch.empty(1024, dtype=torch.float64, device=dev)

    def c1():
        # the current stream at call time, so that a graph capture records it
        lib.pfh_pfaffian_f64(A1.data_ptr(), 1024, 32, out1.data_ptr(), torch.cuda.current_stream().cuda_stream)
    c1()
    torch.cuda.synchronize()
    t_c1 = time_events(torch, c1, 20)
    # the 1024-matrix call captured in a graph and replayed (launch cost
    # amortized; checked against eager)
    want1 = out1.clone()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        c1()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g1 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g1):
        c1()
    out1.zero_()
    g1.replay()
    torch.cuda.synchronize()
    c1_graph_ok = bool(torch.equal(out1, want1))
    t_c1g = time_events(torch, g1.replay, 200)
    del g1

    # single large matrices: plain (unblocked) vs blocked Parlett-Reid
    def large(n, dtype, seed, reps=1):
        gg = torch.Generator(device=dev)
        gg.manual_seed(seed)
        if dtype.is_complex:
            X = torch.complex(torch.randn(n, n, generator=gg, device=dev, dtype=torch.float64),
                              torch.randn(n, n, generator=gg, device=dev, dtype=torch.float64)) / math.sqrt(2)
        else:
            X = torch.randn(n, n, generator=gg, device=dev, dtype=dtype)
        A = ((X - X.T) / 2).contiguous()
        del X
        return A

    def time_slog(A, method, reps):
        dt = {torch.float64: _lib.PFH_F64, torch.complex128: _lib.PFH_C128}[A.dtype]
        n = A.shape[0]
        import ctypes
        need = ctypes.c_size_t(0)
        lib.pfh_slog_pfaffian_workspace_size(dt, 1, n, method, need)
        ws = torch.empty(max(int(need.value), 1), dtype=torch.uint8, device=dev)
        lo = torch.empty(1, dtype=torch.float64, device=dev)
        ph = torch.empty(1, dtype=A.dtype, device=dev)

        def f():
            st = lib.pfh_slog_pfaffian_method(dt, A.data_ptr(), 1, n, method, lo.data_ptr(), ph.data_ptr(),
                                              ws.data_ptr(), ws.numel(), strm)
            if st != 0:
                raise RuntimeError(f"slog status {st}")
        f()
        torch.cuda.synchronize()
        t = time_events(torch, f, reps)
        return t, float(lo.item()), complex(ph.item())

    large_res = {}
    cpu_single = None
    for n, dtype, tag in ((4096, torch.float64, "c2_f64_n4096"), (4096, torch.complex128, "c2_c128_n4096")):
        A = large(n, dtype, args.seed + 7)
        tb, lb, pb = time_slog(A, _lib.PFH_METHOD_BLOCKED, 3)
        tu, lu, pu = time_slog(A, _lib.PFH_METHOD_UNBLOCKED, 1)
        fl = (4 if dtype.is_complex else 1) * n ** 3 / 3
        large_res[tag] = {"blocked_ms": tb * 1e3, "unblocked_ms": tu * 1e3, "speedup": tu / tb,
                          "blocked_tflops": fl / tb / 1e12, "dlog_blocked_vs_unblocked": abs(lb - lu),
                          "sign_agree": abs(pb - pu) < 1e-9}
        if not args.no_cpu:
            import oracle
            Ah = A.cpu().numpy()
            l0, p0 = oracle.slog(Ah, threads=host_cpu()[0])
            large_res[tag]["dlog_blocked_vs_oracle"] = abs(lb - l0)
            large_res[tag]["dlog_unblocked_vs_oracle"] = abs(lu - l0)
            large_res[tag]["phase_agree_with_oracle"] = abs(pb - complex(p0)) < 1e-9
            if not dtype.is_complex:
                # the reference's CPU algorithm on one core, one N = 4096 matrix
                t0 = time.time()
                l1, _ = oracle.slog(Ah, threads=1)
                t1 = time.time() - t0
                cpu_single = {"workload": "one N = 4096 f64 slog (the C2 matrix)", "seconds": t1, "threads": 1,
                              "kind": "port", "gpu_blocked_ms": tb * 1e3, "gpu_speedup": t1 / tb,
                              "equal_to_threaded_oracle": l1 == l0}
            del Ah
        print(f"# {tag}: {large_res[tag]}", file=sys.stderr, flush=True)
        del A
    # C3-shaped input: A = Q diag-block([0, s_i; -s_i, 0]) Q^T, s_i log-uniform in [1e-3, 1e3]
    n8 = 8192
    gg = torch.Generator(device=dev)
    gg.manual_seed(args.seed + 8)
    Q, R = torch.linalg.qr(torch.randn(n8, n8, generator=gg, device=dev, dtype=torch.float64))
    Q = Q * torch.sign(torch.diagonal(R))
    del R
    svals = 10.0 ** (torch.rand(n8 // 2, generator=gg, device=dev, dtype=torch.float64) * 6 - 3)
    D = torch.zeros(n8, n8, dtype=torch.float64, device=dev)
    ii = torch.arange(0, n8, 2, device=dev)
    D[ii, ii + 1] = svals
    D[ii + 1, ii] = -svals
    M = Q @ D @ Q.T
    del D
    A8 = ((M - M.T) / 2).contiguous()
    del M
    sign_det_q = float(torch.sign(torch.linalg.det(Q)).item())
    del Q
    closed = float(torch.log(svals).sum().item())
    tb8, lb8, pb8 = time_slog(A8, _lib.PFH_METHOD_BLOCKED, 2)
    tu8, lu8, pu8 = time_slog(A8, _lib.PFH_METHOD_UNBLOCKED, 1)
    tm8, lm8, pm8 = time_slog(A8, _lib.PFH_METHOD_MIXED, 2)
    l08 = None
    if not args.no_cpu:
        import oracle
        l08, s08 = oracle.slog(A8.cpu().numpy(), threads=host_cpu()[0])
    large_res["c3_f64_n8192"] = {"blocked_ms": tb8 * 1e3, "unblocked_ms": tu8 * 1e3, "speedup": tu8 / tb8,
                                 "dlog_blocked_vs_unblocked": abs(lb8 - lu8),
                                 "dlog_blocked_vs_oracle": None if l08 is None else abs(lb8 - l08),
                                 "dlog_blocked_vs_closed_form": abs(lb8 - closed),
                                 "sign_blocked": pb8.real, "sign_closed_form": sign_det_q,
                                 "note": "closed form log|Pf| = sum log s_i is exact only to ~kappa*u of the rounded Q"}
    print(f"# c3: {large_res['c3_f64_n8192']}", file=sys.stderr, flush=True)
    # the split-precision method on the same matrix: FP64 panel, trailing
    # update as bf16 digit products (exact in fp32) recombined in FP64
    slices = ctypes.c_int(-1)
    lib.pfh_get_tuning(_lib.TUNE_KEYS["mixed_slices"], ctypes.byref(slices))
    large_res["mixed_vs_fp64_n8192"] = {
        "mixed_ms": tm8 * 1e3, "fp64_blocked_ms": tb8 * 1e3, "mixed_over_fp64_time": tm8 / tb8,
        "slices": slices.value if slices.value > 0 else "built-in (9)",
        "dlog_mixed_vs_oracle": None if l08 is None else abs(lm8 - l08),
        "dlog_fp64_vs_oracle": None if l08 is None else abs(lb8 - l08),
        "dlog_mixed_vs_fp64": abs(lm8 - lb8),
        "sign_mixed_equals_oracle": None if l08 is None else pm8.real == s08,
        "within_1e-9": None if l08 is None else abs(lm8 - l08) <= 1e-9,
        "note": "MIXED: FP64 panel; trailing update on v_mfma_f32_16x16x32_bf16 with 7-bit digit slices, "
                "exact fp32 sums, FP64 recombination"}
    print(f"# mixed: {large_res['mixed_vs_fp64_n8192']}", file=sys.stderr, flush=True)
    del A8

    # C5: gradient 1/2 A^-T through the HIP block sweep
    from pfcuda_amd.adjoint import slog_grad
    gg.manual_seed(args.seed + 5)
    X = torch.randn(512, 128, 128, generator=gg, device=dev, dtype=torch.float64)
    Ab = ((X - X.transpose(1, 2)) / 2).contiguous()
    slog_grad(Ab)
    torch.cuda.synchronize()
    t_gb = time_events(torch, lambda: slog_grad(Ab), 3)
    Gb = slog_grad(Ab)
    ref = 0.5 * torch.linalg.inv(Ab[:4]).transpose(1, 2)
    err_gb = float(((Gb[:4] - ref).abs().amax() / ref.abs().amax()).item())
    X = torch.randn(2048, 2048, generator=gg, device=dev, dtype=torch.float64)
    A2 = ((X - X.T) / 2).contiguous()
    slog_grad(A2)
    torch.cuda.synchronize()
    t_g2 = time_events(torch, lambda: slog_grad(A2), 5)
    G2 = slog_grad(A2)
    ref2 = 0.5 * torch.linalg.inv(A2).T
    err_g2 = float(((G2 - ref2).abs().amax() / ref2.abs().amax()).item())
    grad_res = {"batch512_n128_ms": t_gb * 1e3, "batch512_n128_rel_err_vs_inv": err_gb,
                "n2048_ms": t_g2 * 1e3, "n2048_rel_err_vs_inv": err_g2}
    print(f"# c5: {grad_res}", file=sys.stderr, flush=True)

    # C1's kernel over 2^20 matrices (8.6 GB resident), last among the GPU
    # measurements: the large transient allocation is not in the way of the others
    big = 1 << 20
    g.manual_seed(args.seed + 11)
    Bb = torch.randn(big, 32, 32, generator=g, device=dev, dtype=torch.float64)
    Abig = ((Bb - Bb.transpose(1, 2)) / 2).contiguous()
    del Bb
    outb = torch.empty(big, dtype=torch.float64, device=dev)

    def c1big():
        lib.pfh_pfaffian_f64(Abig.data_ptr(), big, 32, outb.data_ptr(), strm)
    c1big()
    torch.cuda.synchronize()
    c1_big_finite = bool(torch.isfinite(outb).all())
    t_c1b = time_events(torch, c1big, 5)
    del Abig, outb
    torch.cuda.empty_cache()
    print(f"# c1: 1024 x 32^2 {t_c1 * 1e3:.4f} ms (graph replay {t_c1g * 1e3:.4f} ms, equal {c1_graph_ok}); "
          f"2^20 x 32^2 {t_c1b * 1e3:.3f} ms", file=sys.stderr, flush=True)
    # accuracy of a sample against the FP64 CPU oracle, and its CPU time
    res = {}
    cpu_baseline = None
    if not args.no_cpu:
        import oracle
        k = min(args.cpu_sample, args.count)
        idx = np.arange(k)
        host = data.cpu().numpy()
        sub_sizes = sizes[idx]
        threads, nproc, model = host_cpu()
        # correctness sample: the first k matrices against the oracle
        wl, ws = oracle.slog_ragged(host, offsets[idx], sub_sizes, threads=threads)
        gl = run.log.cpu().numpy()[idx]
        gs = run.ph.cpu().numpy()[idx]
        res = {"max_abs_dlog": float(np.max(np.abs(gl - wl))),
               "sign_mismatches": int(np.sum(gs != ws)), "checked": int(k)}
        # CPU throughput: whole-batch passes of the oracle until >= 10 s of wall time
        passes, tc = 0, time.time()
        while True:
            oracle.slog_ragged(host, offsets, sizes, threads=threads)
            passes += 1
            tcpu = time.time() - tc
            if tcpu >= 10.0:
                break
        cpu_baseline = {"value": passes * args.count / tcpu, "unit": "Pf/s", "cores": threads, "kind": "port",
                        "nproc": nproc, "cpu_model": model,
                        "sample": f"{passes} pass(es) over the full {args.count}-matrix batch (N 16-256 f64), "
                                  f"{threads} threads (every core this process may use), one matrix per task, "
                                  f"{tcpu:.1f} s",
                        "single_matrix": cpu_single}

    line = {
        "metric": "batched Pf/s (N 16-256); slog_pf N=8192 ms; mixed vs FP64, |dlog|<=1e-9",
        "value": value, "unit": "Pf/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded N(0,1), A=(B-B^T)/2, generated on GPU)",
        "config": {"workload": "ragged slog_pfaffian: 16384 f64 skew matrices, even N uniform in [16,256]",
                   "count": args.count, "seed": args.seed, "parallelism": "single GPU",
                   "packing": "one buffer, matrices back to back in input order (sizes mixed); "
                              "the call buckets them by size on the device"},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": FP64_PEAK / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK,
                     "traffic": committed_traffic("pf_bat_kernel<256,12>"),
                     "kernel": "pf_bat_kernel<256,12> (bucket 192<N<=256)",
                     "launch_ms": t_big * 1e3, "algorithmic_flop": big_flops},
        "whole_batch": {"flop": flops, "tflops": flops / (ms / 1e3) / 1e12,
                        "frac_fp64_peak": flops / (ms / 1e3) / FP64_PEAK, "triangle_bytes": tri_bytes},
        "graph_replay": {"ms_per_step": ms_graph, "pf_per_s": args.count / (ms_graph / 1e3),
                         "equal_to_eager": graph_ok},
        "per_matrix_launches_f64": per_matrix_res,
        "ragged_c128": c128_res,
        "small_batched_c1": {"workload": "pfaffian() 1024 x 32x32 f64", "ms": t_c1 * 1e3,
                             "pf_per_s": 1024 / t_c1,
                             "hbm_frac": (1024 * 32 * 31 / 2 * 8) / t_c1 / HBM_PEAK,
                             "graph_replay_ms": t_c1g * 1e3, "graph_replay_pf_per_s": 1024 / t_c1g,
                             "graph_equal_to_eager": c1_graph_ok,
                             "batch_2p20_ms": t_c1b * 1e3, "batch_2p20_pf_per_s": big / t_c1b,
                             "batch_2p20_hbm_frac": (big * 32 * 31 / 2 * 8) / t_c1b / HBM_PEAK,
                             "batch_2p20_read_frac": (big * 32 * 32 * 8) / t_c1b / HBM_PEAK,
                             "batch_2p20_all_finite": c1_big_finite,
                             "note": "hbm_frac counts the strict upper triangle (algorithmic bytes); "
                                     "read_frac the full square the kernel reads"},
        "large_single_matrix": large_res,
        "gradient_c5": grad_res,
        "accuracy_vs_oracle": res,
        "cpu_baseline": cpu_baseline,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
