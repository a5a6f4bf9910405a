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
