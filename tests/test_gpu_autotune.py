"""The autotuner on the GPU: a quick tune of some keys picks values from each
key's list, the settings it leaves behind give results that agree with the
oracle, and the cache it writes loads back."""
import math

import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


def test_quick_tune_and_reload(cuda, tmp_path):
    import pfcuda_amd as pf
    from pfcuda_amd import autotune
    keys = ["f64_bucket_128", "c128_bucket_128", "trail_f64", "grad_blocked"]
    out = str(tmp_path / "tune.json")
    try:
        res = autotune.tune(out=out, quick=True, verbose=False, keys=keys)
        for k in keys:
            assert res["choices"][k] in autotune.CANDIDATES[k], (k, res)
            assert autotune.get_tuning(k) == res["choices"][k]
        # results under the tuned settings: the oracle
        rng = np.random.default_rng(70)
        mats = [oracle.skew_gaussian(rng, int(n)) for n in (80, 100, 128, 600)]
        for A in mats:
            la, s = pf.slog_pfaffian(torch.from_numpy(A).to(cuda))
            l0, p0 = oracle.slog(A)
            assert abs(float(la) - l0) < 1e-9 and float(s) == p0
        for name in autotune.CANDIDATES:
            autotune.set_tuning(name, -1)
        assert autotune.load(out) == res["choices"]
    finally:
        for name in autotune.CANDIDATES:
            autotune.set_tuning(name, -1)


@pytest.mark.parametrize("key,values,sizes,cplx", [
    ("bat_nb_128", [6, 8, 12], (66, 100, 128), False),
    ("bat_nb_192", [6, 8, 12], (130, 160, 192), False),
    ("bat_nb_256", [8, 10, 12], (194, 230, 256), False),
    ("bat_nb_c128_128", [4, 6, 8], (66, 128), True),
    ("bat_nb_c128_192", [4, 6, 8], (130, 192), True),
    ("bat_nb_c128_256", [6, 8], (194, 256), True),
])
def test_every_batched_panel_width_vs_oracle(cuda, key, values, sizes, cplx):
    # each instantiated width of each bucket, through a ragged call of 40
    # matrices (more than a panel's worth of pair steps at every size)
    import pfcuda_amd as pf
    from pfcuda_amd import autotune
    rng = np.random.default_rng(71)
    ns = [sizes[i % len(sizes)] for i in range(40)]
    mats = [oracle.skew_gaussian(rng, n, complex_=cplx) for n in ns]
    data = np.concatenate([m.reshape(-1) for m in mats])
    offs = np.cumsum([0] + [n * n for n in ns[:-1]])
    want = [oracle.slog(m) for m in mats]
    try:
        for v in values:
            autotune.set_tuning(key, v)
            gl, gp = pf.slog_pfaffian_ragged(torch.from_numpy(data).to(cuda), offs, ns)
            gl, gp = gl.cpu().numpy(), gp.cpu().numpy()
            for i, (l0, p0) in enumerate(want):
                assert abs(gl[i] - l0) < 1e-9 and abs(complex(gp[i]) - complex(p0)) < 1e-9, (key, v, ns[i])
    finally:
        autotune.set_tuning(key, -1)


@pytest.mark.parametrize("key,values,cplx", [("blocked_nb", [16, 24, 32], False), ("blocked_nb_c128", [8, 16], True)])
def test_every_blocked_panel_width_vs_oracle(cuda, key, values, cplx):
    import pfcuda_amd as pf
    from pfcuda_amd import autotune
    rng = np.random.default_rng(72)
    mats = [oracle.skew_gaussian(rng, n, complex_=cplx) for n in (258, 700, 1030)]
    try:
        for v in values:
            autotune.set_tuning(key, v)
            for A in mats:
                la, ph = pf.slog_pfaffian(torch.from_numpy(A).to(cuda), pf.PFH_METHOD_BLOCKED)
                l0, p0 = oracle.slog(A)
                assert abs(float(la) - l0) < 1e-9 and abs(complex(ph.item()) - complex(p0)) < 1e-9, (key, v)
            pf.check_status()
    finally:
        autotune.set_tuning(key, -1)


def test_switch_point_routes_single_matrices_vs_oracle(cuda):
    # with the switch at 96, single matrices of 98 .. 256 run blocked, and a
    # batch of many matrices of the same size stays on the batched kernels;
    # both agree with the oracle
    import pfcuda_amd as pf
    from pfcuda_amd import autotune
    rng = np.random.default_rng(73)
    try:
        autotune.set_tuning("switch_n", 96)
        for n in (98, 160, 256):
            A = oracle.skew_gaussian(rng, n)
            la, s = pf.slog_pfaffian(torch.from_numpy(A).to(cuda))
            l0, s0 = oracle.slog(A)
            assert abs(float(la) - l0) < 1e-9 and float(s) == s0, n
        B = oracle.skew_gaussian(rng, 160, batch=300)
        la, s = pf.slog_pfaffian(torch.from_numpy(B).to(cuda))
        for a, l1, s1 in zip(B[:: 37], la.cpu().numpy()[:: 37], s.cpu().numpy()[:: 37]):
            l0, s0 = oracle.slog(a)
            assert abs(l1 - l0) < 1e-9 and s1 == s0
    finally:
        autotune.set_tuning("switch_n", -1)


def test_quick_tune_of_widths_and_switch(cuda, tmp_path):
    from pfcuda_amd import autotune
    keys = ["bat_nb_256", "bat_nb_c128_128", "blocked_nb", "blocked_nb_c128", "switch_n"]
    try:
        res = autotune.tune(out=str(tmp_path / "t.json"), quick=True, verbose=False, keys=keys)
        for k in keys:
            assert res["choices"][k] in autotune.CANDIDATES[k], (k, res)
            assert all(v is not None for v in res["timings_ms"][k].values()), (k, res)   # every candidate agreed
    finally:
        for name in autotune.CANDIDATES:
            autotune.set_tuning(name, -1)
