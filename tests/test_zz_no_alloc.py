"""The allocation-free ragged entry point against the ordinary one, through
the C ABI: the same kernels on the same matrices, so the results must be
bitwise equal (each matrix is factored by one workgroup, whichever it is).
Two calls in a row, the second smaller, so the second runs on what the
first left in the shared scratch array."""
import ctypes

import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib(cuda):
    from pfcuda_amd import _lib
    return _lib


def _batch(rng, sizes, cplx, dev):
    mats = [oracle.skew_gaussian(rng, n, complex_=cplx) for n in sizes]
    data = torch.from_numpy(np.concatenate([m.reshape(-1) for m in mats])).to(dev)
    offs = torch.tensor(np.cumsum([0] + [n * n for n in sizes[:-1]]), dtype=torch.int64, device=dev)
    sz = torch.tensor(sizes, dtype=torch.int32, device=dev)
    return data, offs, sz


def _ragged(L, dt, data, offs, sz, max_n, noalloc):
    dev = data.device
    count = sz.numel()
    lo = torch.full((count,), 7.0, dtype=torch.float64, device=dev)
    ph = torch.full((count,), 7.0, dtype=data.dtype, device=dev)
    strm = torch.cuda.current_stream(dev).cuda_stream
    if noalloc:
        st = L.lib.pfh_slog_pfaffian_ragged_noalloc(dt, data.data_ptr(), offs.data_ptr(), sz.data_ptr(), count,
                                                    max_n, lo.data_ptr(), ph.data_ptr(), strm)
    else:
        need = ctypes.c_size_t(0)
        L.check(L.lib.pfh_slog_pfaffian_ragged_workspace_size(dt, count, max_n, ctypes.byref(need)), "ws")
        ws = torch.empty(max(need.value, 1), dtype=torch.uint8, device=dev)
        st = L.lib.pfh_slog_pfaffian_ragged(dt, data.data_ptr(), offs.data_ptr(), sz.data_ptr(), count, max_n,
                                            lo.data_ptr(), ph.data_ptr(), ws.data_ptr(), ws.numel(), strm)
    L.check(st, "ragged")
    torch.cuda.synchronize()
    return lo.cpu().numpy(), ph.cpu().numpy()


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128, torch.float32])
def test_noalloc_matches_ragged_twice(lib, cuda, dtype):
    dt = {torch.float64: lib.PFH_F64, torch.complex128: lib.PFH_C128, torch.float32: lib.PFH_F32}[dtype]
    cplx = dtype.is_complex
    rng = np.random.default_rng(606)
    # first call: every bucket, 1500 matrices in the 192 < n <= 256 bucket
    # (more than the slab pairs the scratch array holds, so its workgroups
    # re-claim), odd and empty sizes
    sizes1 = ([int(x) for x in 2 * rng.integers(97, 129, size=1500)]
              + [int(x) for x in 2 * rng.integers(8, 97, size=1500)] + [0, 7, 2, 4])
    rng.shuffle(sizes1)
    # second call: smaller batch and max_n
    sizes2 = [int(x) for x in 2 * rng.integers(8, 65, size=300)] + [5]
    for sizes, max_n in ((sizes1, 256), (sizes2, 128)):
        data, offs, sz = _batch(rng, sizes, cplx, cuda)
        data = data.to(dtype)
        want = _ragged(lib, dt, data, offs, sz, max_n, noalloc=False)
        got = _ragged(lib, dt, data, offs, sz, max_n, noalloc=True)
        assert np.array_equal(got[0], want[0])
        assert np.array_equal(got[1], want[1])
        # and the values are the oracle's (f64 / c128 bar)
        if dtype != torch.float32:
            for i in range(0, len(sizes), 97):
                n = sizes[i]
                a = data[offs[i]:offs[i] + n * n].reshape(n, n).cpu().numpy()
                l0, p0 = oracle.slog(a)
                assert (got[0][i] == l0) or abs(got[0][i] - l0) <= 1e-9, (i, n)


def test_noalloc_rejects_beyond_its_limits(lib, cuda):
    dev = cuda
    data = torch.zeros(300 * 300, dtype=torch.float64, device=dev)
    offs = torch.zeros(1, dtype=torch.int64, device=dev)
    sz = torch.tensor([300], dtype=torch.int32, device=dev)
    lo = torch.empty(1, dtype=torch.float64, device=dev)
    ph = torch.empty(1, dtype=torch.float64, device=dev)
    strm = torch.cuda.current_stream(dev).cuda_stream
    f = lib.lib.pfh_slog_pfaffian_ragged_noalloc
    assert f(lib.PFH_F64, data.data_ptr(), offs.data_ptr(), sz.data_ptr(), 1, 300, lo.data_ptr(), ph.data_ptr(),
             strm) == -3                                        # max_n past 256
    assert f(lib.PFH_F64, data.data_ptr(), offs.data_ptr(), sz.data_ptr(), (1 << 22) + 1, 64, lo.data_ptr(),
             ph.data_ptr(), strm) == -3                         # count past 2^22 (checked before any access)
    assert f(lib.PFH_F64, None, offs.data_ptr(), sz.data_ptr(), 1, 64, lo.data_ptr(), ph.data_ptr(), strm) == -1
