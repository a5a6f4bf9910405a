"""Autotuner for the library's run-time choices (include/pfaffian_hip.h,
PFH_TUNE_*): which batched kernel serves each slog size bucket, which form of
the trailing update the blocked paths use, which sweep computes large
gradients, the panel widths, and the size at which single matrices leave the
batched kernels for the blocked path.

``tune()`` times every candidate of every key on the current device, on
seeded synthetic workloads of the shape the key governs, checks that each
candidate's results agree with the built-in choice's (|dlog| <= 1e-9, same
sign or phase within 1e-9), keeps the fastest, and writes a JSON cache keyed
by the device architecture.  ``load()`` applies a cache written for the same
architecture; importing ``pfcuda_amd`` with PFH_TUNE_FILE set does the same.
The panel widths (pairs per panel) of the batched blocked kernel's buckets
and of the blocked factorization choose among the widths the library
instantiates, and the switch point picks the largest n that a dense call of
few matrices keeps on the batched kernels (larger ones run blocked).

    python -m pfcuda_amd.autotune [--out PATH] [--quick]
"""
from __future__ import annotations

import argparse
import ctypes
import datetime
import json
import math
import os
import sys

from . import _lib
from ._lib import TUNE_KEYS, check, lib

# candidate values per key (the value lists of include/pfaffian_hip.h)
CANDIDATES = {
    "f64_bucket_128": [0, 2, 3], "f64_bucket_192": [0, 2, 3], "f64_bucket_256": [0, 2, 3],
    "c128_bucket_128": [0, 3],
    "trail_f64": [0, 1], "trail_c128": [0, 1], "trail_sweep": [0, 1],
    "panel_launch": [0], "grad_blocked": [0, 1],
    # panel widths (pairs per panel) and the single-matrix switch point
    "bat_nb_128": [6, 8, 12], "bat_nb_192": [6, 8, 12], "bat_nb_256": [8, 10, 12],
    "bat_nb_c128_128": [4, 6, 8], "bat_nb_c128_192": [4, 6, 8], "bat_nb_c128_256": [6, 8],
    "blocked_nb": [16, 24, 32], "blocked_nb_c128": [8, 16], "grad_nb": [16, 32],
    "switch_n": [96, 128, 160, 192, 224, 256],
    "small_kernel": [0, 1],
}
BUCKETS = {"f64_bucket_128": (64, 128), "f64_bucket_192": (128, 192), "f64_bucket_256": (192, 256),
           "c128_bucket_128": (64, 128),
           "bat_nb_128": (64, 128), "bat_nb_192": (128, 192), "bat_nb_256": (192, 256),
           "bat_nb_c128_128": (64, 128), "bat_nb_c128_192": (128, 192), "bat_nb_c128_256": (192, 256)}
SWITCH_SIZES = (96, 128, 160, 192, 224, 256)


def set_tuning(name: str, value: int) -> None:
    check(lib.pfh_set_tuning(TUNE_KEYS[name], int(value)), f"set_tuning({name}, {value})")


def get_tuning(name: str) -> int:
    v = ctypes.c_int(0)
    check(lib.pfh_get_tuning(TUNE_KEYS[name], ctypes.byref(v)), f"get_tuning({name})")
    return v.value


def device_arch() -> str:
    import torch
    return torch.cuda.get_device_properties(0).gcnArchName


def default_path(arch: str) -> str:
    safe = "".join(ch if ch.isalnum() or ch in "-_" else "_" for ch in arch)
    return os.path.join(os.path.expanduser("~"), ".cache", "pfcuda_amd", f"tune-{safe}.json")


def _time(torch, fn, reps):
    fn()
    torch.cuda.synchronize()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / reps


def _skew(torch, shape, complex_, gen, dev):
    if complex_:
        B = torch.complex(torch.randn(*shape, generator=gen, device=dev, dtype=torch.float64),
                          torch.randn(*shape, generator=gen, device=dev, dtype=torch.float64)) / math.sqrt(2)
    else:
        B = torch.randn(*shape, generator=gen, device=dev, dtype=torch.float64)
    return ((B - B.transpose(-1, -2)) / 2).contiguous()


def _bucket_workload(torch, name, count, dev):
    """A ragged batch of `count` matrices, n even and uniform over the bucket."""
    from .api import pack_ragged, slog_pfaffian_ragged
    lo, hi = BUCKETS[name]
    gen = torch.Generator(device=dev)
    gen.manual_seed(1234 + hi)
    sizes = [lo + 2 + 2 * int(i) for i in torch.randint(0, (hi - lo) // 2, (count,), generator=torch.Generator().manual_seed(hi))]
    cplx = "c128" in name
    mats = [_skew(torch, (n, n), cplx, gen, dev) for n in sizes]
    data, offsets, sz = pack_ragged(mats)
    offsets, szd, mx = offsets.to(dev), sz.to(dev), int(sz.max())   # no host work per call

    def run():
        return slog_pfaffian_ragged(data, offsets, szd, max_n=mx)
    return run


def _switch_workload(torch, dev):
    """Single-matrix solves (the automatic method) at n = 96 .. 256, one each:
    the switch point's candidates differ in which of these run blocked."""
    from .api import slog_pfaffian
    gen = torch.Generator(device=dev)
    gen.manual_seed(5678)
    mats = [_skew(torch, (n, n), False, gen, dev) for n in SWITCH_SIZES]

    def run():
        outs = [slog_pfaffian(A) for A in mats]
        return torch.stack([o[0] for o in outs]), torch.stack([o[1] for o in outs])
    return run


def _small_workload(torch, dev):
    """pfaffian() of 2^16 seeded 32 x 32 f64 matrices."""
    from .api import pfaffian
    gen = torch.Generator(device=dev)
    gen.manual_seed(3232)
    A = _skew(torch, (1 << 16, 32, 32), False, gen, dev)

    def run():                                    # the kernel's output alone: nothing else is timed
        return (pfaffian(A),)
    return run


def _single_workload(torch, name, n, dev):
    """One n x n solve (blocked slog) or gradient, through the host API."""
    from .api import slog_pfaffian
    from .adjoint import slog_grad
    gen = torch.Generator(device=dev)
    gen.manual_seed(4321 + n)
    cplx = name in ("trail_c128", "blocked_nb_c128")
    A = _skew(torch, (n, n), cplx, gen, dev)
    if name in ("trail_sweep", "grad_blocked", "grad_nb"):
        def run():
            G, la, ph = slog_grad(A, return_slog=True)
            return la, ph, G
    else:
        def run():
            return slog_pfaffian(A, method=_lib.PFH_METHOD_BLOCKED)
    return run


def _agree(torch, ref, got, tol=1e-9):
    """log|Pf| within tol, sign / phase within tol, and a gradient (when the
    workload returns one) within tol relative to its largest entry; a lone
    Pfaffian (the product form) within tol relative to each value."""
    if len(ref) == 1:
        return bool(((ref[0] - got[0]).abs() <= tol * ref[0].abs()).all())
    l0, p0, l1, p1 = ref[0], ref[1], got[0], got[1]
    dl = (l0 - l1).abs()
    dl = torch.where(torch.isinf(l0) & (l0 == l1), torch.zeros_like(dl), dl)
    ok = bool((dl <= tol).all()) and bool(((p0 - p1).abs() <= tol).all())
    if ok and len(ref) > 2:
        g0, g1 = ref[2], got[2]
        ok = float((g0 - g1).abs().max()) <= tol * float(g0.abs().max())
    return ok


def tune(out: str | None = None, quick: bool = False, verbose: bool = True, keys=None) -> dict:
    """Time every candidate of each key on the current device; set the
    fastest that agrees with the built-in choice, and write the cache."""
    import torch
    dev = torch.device("cuda:0")
    arch = device_arch()
    count = 256 if quick else 2048
    n_big = 1024 if quick else 2048
    reps = 2 if quick else 5
    choices, timings = {}, {}
    for name in (keys or list(CANDIDATES)):   # the run-time choices (not the panel's spin limit)
        set_tuning(name, -1)
        if name in BUCKETS:
            run = _bucket_workload(torch, name, count, dev)
        elif name == "switch_n":
            run = _switch_workload(torch, dev)
        elif name == "small_kernel":
            run = _small_workload(torch, dev)
        else:
            run = _single_workload(torch, name, n_big, dev)
        ref = run()
        torch.cuda.synchronize()
        best, best_ms, ms_of = -1, math.inf, {}
        for v in CANDIDATES[name]:
            set_tuning(name, v)
            got = run()
            torch.cuda.synchronize()
            if not _agree(torch, ref, got):
                ms_of[str(v)] = None              # disagrees with the built-in choice: never chosen
                continue
            ms = _time(torch, run, reps)
            ms_of[str(v)] = ms
            if ms < best_ms:
                best, best_ms = v, ms
        set_tuning(name, best)
        choices[name], timings[name] = best, ms_of
        if verbose:
            print(f"[autotune] {name}: {ms_of} -> {best}", file=sys.stderr, flush=True)
    result = {"arch": arch, "library": "pfcuda_amd", "created": datetime.datetime.now().isoformat(timespec="seconds"),
              "workload": {"bucket_count": count, "single_n": n_big, "reps": reps},
              "choices": choices, "timings_ms": timings}
    path = out or default_path(arch)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(result, f, indent=1)
    if verbose:
        print(f"[autotune] wrote {path}", file=sys.stderr, flush=True)
    return result


def apply(choices: dict) -> None:
    """Set every key named in `choices` (unknown names are an error)."""
    for name, v in choices.items():
        if name not in CANDIDATES:
            raise KeyError(f"unknown tuning key {name!r}")
        set_tuning(name, int(v))


def load(path: str | None = None, arch: str | None = None) -> dict:
    """Apply a cache written by tune() for this device's architecture.  A
    cache for another architecture is not applied (returns {})."""
    arch = arch or device_arch()
    path = path or default_path(arch)
    with open(path) as f:
        data = json.load(f)
    if data.get("arch") != arch:
        print(f"[autotune] {path} was tuned on {data.get('arch')!r}, not {arch!r}: not applied",
              file=sys.stderr, flush=True)
        return {}
    apply(data.get("choices", {}))
    return data["choices"]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=None, help="cache path (default: ~/.cache/pfcuda_amd/tune-<arch>.json)")
    ap.add_argument("--quick", action="store_true", help="small workloads (seconds)")
    args = ap.parse_args(argv)
    res = tune(out=args.out, quick=args.quick)
    print(json.dumps(res["choices"]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
