"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle,
the golden fixtures and the reference's own known-answer tests.

Tolerances: product form rel 1e-12 (f64/c128), log|Pf| abs 1e-9 with exact
sign (f64/c128); float32/complex64 are checked at rel 1e-3 / abs 1e-3
against the FP64 oracle."""
import math
import os

import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def pf(cuda):
    import pfcuda_amd
    return pfcuda_amd


def unpack(data, sizes):
    off = 0
    for n in sizes:
        yield data[off:off + n * n].reshape(n, n)
        off += n * n


def test_c1_batch_1024x32_f64(pf, cuda):
    # Plain method, small: 1024 x (32x32) f64, A = (B - B^T)/2, rel 1e-12
    rng = np.random.default_rng(0)
    A = oracle.skew_gaussian(rng, 32, batch=1024)
    got = pf.pfaffian(torch.from_numpy(A).to(cuda)).cpu().numpy()
    want = np.array([oracle.pf_f64(a)[0] for a in A])
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_two_matrices_per_wave_batches(pf, cuda, dtype):
    # Real batches of 8192 and more n <= 32 matrices run two per wave: an odd
    # count (the last wave holds one matrix), ragged sizes with odd and zero
    # n, integer entries (many exact ties in the pivot search) and a NaN row.
    # Product form: bit for bit the one-per-wave kernel (the batch in chunks
    # below the switch) and rel 1e-12 against the oracle (f64); slog of the
    # ragged batch against the oracle.
    rng = np.random.default_rng(11)
    cnt = 8193
    A = oracle.skew_gaussian(rng, 32, batch=cnt)
    B = rng.integers(-2, 3, size=(600, 32, 32)).astype(np.float64)
    A[:600] = np.triu(B, 1) - np.triu(B, 1).transpose(0, 2, 1)
    A[700, 3, 5:] = np.nan
    A[700, 5:, 3] = np.nan
    At = torch.from_numpy(A).to(dtype).to(cuda)
    got = pf.pfaffian(At)
    parts = torch.cat([pf.pfaffian(At[i:i + 4096]) for i in range(0, cnt, 4096)])
    assert torch.equal(torch.nan_to_num(got, nan=7.0), torch.nan_to_num(parts, nan=7.0))
    got = got.cpu().numpy()
    assert math.isnan(got[700])
    lo = dtype == torch.float32
    for i in list(range(0, 600, 7)) + list(range(600, cnt, 97)) + [cnt - 1]:
        if i == 700:
            continue
        a = A[i].astype(np.float32).astype(np.float64) if lo else A[i]
        want = oracle.pf_f64(a)[0]
        tol = (1e-3 * abs(want) + 1e-5) if lo else 1e-12 * max(abs(want), 1.0)
        assert abs(got[i] - want) <= tol, (i, got[i], want)

    sizes = [int(x) for x in rng.integers(0, 33, size=cnt)]
    sizes[:6] = [0, 1, 2, 31, 32, 3]
    mats = [oracle.skew_gaussian(rng, n) if n >= 2 else np.zeros((n, n)) for n in sizes]
    data = np.concatenate([m.reshape(-1) for m in mats]) if sizes else np.zeros(0)
    offs = np.cumsum([0] + [n * n for n in sizes[:-1]])
    gl, gp = pf.slog_pfaffian_ragged(torch.from_numpy(data).to(dtype).to(cuda), offs, sizes)
    gl, gp = gl.cpu().numpy(), gp.cpu().numpy()
    src = data.astype(np.float32).astype(np.float64) if lo else data
    wl, wp = oracle.slog_ragged(src, offs, sizes, threads=8)
    for i, n in enumerate(sizes):
        if math.isinf(wl[i]):
            assert math.isinf(gl[i]) and gl[i] < 0 and gp[i] == 0, (i, n)
        else:
            assert abs(gl[i] - wl[i]) < (1e-3 if lo else 1e-9), (i, n, gl[i], wl[i])
            assert gp[i] == wp[i], (i, n, gp[i], wp[i])


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128, torch.float32, torch.complex64])
def test_pfaffian_all_sizes(pf, cuda, dtype):
    rng = np.random.default_rng(1)
    cplx = dtype.is_complex
    lo = dtype in (torch.float32, torch.complex64)
    for n in range(6, 65, 2):
        A = oracle.skew_gaussian(rng, n, complex_=cplx, batch=5)
        got = pf.pfaffian(torch.from_numpy(A).to(dtype).to(cuda)).cpu().numpy()
        for a, g in zip(A, got):
            want = oracle.pf_c128(a)[0] if cplx else oracle.pf_f64(a)[0]
            if lo:
                assert abs(g - want) <= 1e-3 * abs(want) + 1e-5, (n, g, want)
            else:
                assert abs(g - want) <= 1e-12 * abs(want) * max(1.0, n / 16), (n, g, want)


def test_pfaffian_golden_reference_values(pf, cuda):
    g = np.load(os.path.join(GOLD, "ref_product.npz"))
    for A, want in zip(unpack(g["data"], g["sizes"]), g["pf"]):
        got = float(pf.pfaffian(torch.from_numpy(A.copy()).to(cuda)).item())
        assert got == pytest.approx(want, rel=1e-12, abs=1e-300)


def test_reference_known_answers_pfaffian(pf, cuda):
    # test/pfaffian_test.py: zero matrix, odd dims, 2x2, det relation, identity
    for n in range(2, 20, 2):
        assert float(pf.pfaffian(torch.zeros(n, n, dtype=torch.float64, device=cuda))) == 0.0
    rng = np.random.default_rng(0)
    for n in [1, 3, 5, 11, 15]:
        B = rng.normal(size=(n, n))
        assert abs(float(pf.pfaffian(torch.from_numpy(B - B.T).to(cuda)))) <= 1e-12
    for a in [0.0, 1.0, -3.5, 1e-12, 1e10]:
        A = torch.tensor([[0, a], [-a, 0]], dtype=torch.float64, device=cuda)
        assert float(pf.pfaffian(A)) == pytest.approx(a)
    rng = np.random.default_rng(1)
    for n in range(2, 33, 2):
        B = rng.normal(size=(n, n))
        A = B - B.T
        p = float(pf.pfaffian(torch.from_numpy(A).to(cuda)))
        assert np.isclose(p * p, np.linalg.det(A), rtol=1e-8, atol=1e-10)
    rng = np.random.default_rng(seed=0)
    for n in range(2, 33, 2):
        for _ in range(10):
            M, want = oracle.tridiag_identity_case(rng, n)
            got = float(pf.pfaffian(torch.from_numpy(M).to(cuda)))
            assert np.isclose(want, got, rtol=1e-10, atol=1e-12), n


def test_reference_known_answers_slog(pf, cuda):
    # test/slog_pfaffian_test.py: zero, odd, det relation, identity
    for n in range(34, 68, 2):
        la, s = pf.slog_pfaffian(torch.zeros(n, n, dtype=torch.float64, device=cuda))
        assert float(s) == 0 and math.isinf(float(la)) and float(la) < 0
    rng = np.random.default_rng(0)
    for n in [35, 37, 39]:
        B = rng.normal(size=(n, n))
        la, s = pf.slog_pfaffian(torch.from_numpy(B - B.T).to(cuda))
        assert float(s) == 0 and float(la) == -math.inf
    rng = np.random.default_rng(1)
    for n in range(34, 66, 2):
        B = rng.normal(size=(n, n))
        A = B - B.T
        la, s = pf.slog_pfaffian(torch.from_numpy(A).to(cuda))
        assert np.isclose(np.exp(2 * float(la)), np.linalg.det(A), rtol=1e-8, atol=1e-10)
    rng = np.random.default_rng(seed=0)
    for n in range(40, 300, 10):
        for _ in range(10):
            M, want = oracle.tridiag_identity_case(rng, n)
            la, s = pf.slog_pfaffian(torch.from_numpy(M).to(cuda))
            got = float(s) * math.exp(float(la))
            assert np.isclose(want, got, rtol=1e-10, atol=1e-12), n


@pytest.mark.parametrize("cplx", [False, True])
def test_slog_vs_oracle_and_mpmath(pf, cuda, cplx):
    kind = "complex" if cplx else "real"
    g = np.load(os.path.join(GOLD, "mp_slog.npz"))
    for A, la, ph in zip(unpack(g[f"{kind}_data"], g[f"{kind}_sizes"]), g[f"{kind}_log"], g[f"{kind}_phase"]):
        gl, gp = pf.slog_pfaffian_batched(torch.from_numpy(A.copy()).to(cuda))
        assert abs(float(gl) - la) < 1e-10
        assert abs(complex(gp.item()) - ph) < 1e-9
    rng = np.random.default_rng(2)
    for n in (34, 64, 100, 128, 130, 192, 200, 256):
        A = oracle.skew_gaussian(rng, n, complex_=cplx, batch=3)
        gl, gp = pf.slog_pfaffian(torch.from_numpy(A).to(cuda))
        for a, l1, p1 in zip(A, gl.cpu().numpy(), gp.cpu().numpy()):
            l0, p0 = oracle.slog(a)
            assert abs(l1 - l0) < 1e-9, (n, l1, l0)
            assert abs(p1 - p0) < 1e-9, (n, p1, p0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
def test_ragged_vs_oracle(pf, cuda, dtype):
    rng = np.random.default_rng(3)
    cplx = dtype.is_complex
    sizes = [int(x) for x in 2 * rng.integers(8, 129, size=300)] + [2, 4, 0, 7, 33, 256, 256, 40]
    mats = [oracle.skew_gaussian(rng, n, complex_=cplx) for n in sizes]
    mats[-1][:] = 0                                   # a zero matrix: (-inf, 0)
    data = np.concatenate([m.reshape(-1) for m in mats])
    offs = np.cumsum([0] + [int(n) * int(n) for n in sizes[:-1]])
    gl, gp = pf.slog_pfaffian_ragged(torch.from_numpy(data).to(cuda), offs, sizes)
    gl, gp = gl.cpu().numpy(), gp.cpu().numpy()
    wl, wp = oracle.slog_ragged(data, offs, sizes, threads=8)
    for i, n in enumerate(sizes):
        if math.isinf(wl[i]):
            assert math.isinf(gl[i]) and gp[i] == 0, (i, n)
        else:
            assert abs(gl[i] - wl[i]) < 1e-9, (i, n, gl[i], wl[i])
            assert abs(gp[i] - wp[i]) < 1e-9, (i, n, gp[i], wp[i])


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
def test_ragged_above_256_and_beyond_max_n(pf, cuda, dtype):
    # Sizes past the batched kernels go to the global-slab bucket of the
    # same call; an odd one gives (-inf, 0).  A size past the caller's max_n
    # (which sizes the workspace) gives NaN and touches nothing, and the
    # other matrices of the call are unaffected.
    rng = np.random.default_rng(256)
    cplx = dtype.is_complex
    sizes = [258, 300, 400, 301, 16, 130, 256, 64]
    mats = [oracle.skew_gaussian(rng, n, complex_=cplx) for n in sizes]
    data = np.concatenate([m.reshape(-1) for m in mats])
    offs = np.cumsum([0] + [n * n for n in sizes[:-1]])
    gl, gp = pf.slog_pfaffian_ragged(torch.from_numpy(data).to(cuda), offs, sizes)
    gl, gp = gl.cpu().numpy(), gp.cpu().numpy()
    wl, wp = oracle.slog_ragged(data, offs, sizes, threads=8)
    for i, n in enumerate(sizes):
        if n & 1:
            assert gl[i] == -math.inf and gp[i] == 0, (i, n)
        else:
            assert abs(gl[i] - wl[i]) < 1e-9 and abs(gp[i] - wp[i]) < 1e-9, (i, n, gl[i], wl[i])
    sizes2 = [100, 300, 320]
    mats2 = [oracle.skew_gaussian(rng, n, complex_=cplx) for n in sizes2]
    data2 = np.concatenate([m.reshape(-1) for m in mats2])
    offs2 = np.cumsum([0] + [n * n for n in sizes2[:-1]])
    gl2, gp2 = pf.slog_pfaffian_ragged(torch.from_numpy(data2).to(cuda), offs2, sizes2, max_n=300)
    gl2, gp2 = gl2.cpu().numpy(), gp2.cpu().numpy()
    for i in (0, 1):
        l0, p0 = oracle.slog(mats2[i])
        assert abs(gl2[i] - l0) < 1e-9 and abs(gp2[i] - p0) < 1e-9, (i, gl2[i], l0)
    assert math.isnan(gl2[2]) and np.isnan(gp2[2])


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
def test_panel_poll_limit_gives_nan_then_recovers(pf, cuda, dtype):
    # The panel's wait for the other blocks' candidate records is bounded:
    # past its limit it sets the panel's error flag, the output is NaN,
    # pfh_stream_status reports PFH_ETIMEOUT, and nothing hangs.  The limit is forced to zero from the host
    # (polls give up before reading), for the solve and for the gradient's
    # block sweep; the next calls, at the default limit, are exact again.
    from pfcuda_amd import _lib
    from pfcuda_amd.adjoint import slog_grad
    rng = np.random.default_rng(77)
    A = oracle.skew_gaussian(rng, 520, complex_=dtype.is_complex)
    At = torch.from_numpy(A).to(cuda)
    l0, p0 = oracle.slog(A)
    spin, gb = _lib.TUNE_KEYS["panel_spin"], _lib.TUNE_KEYS["grad_blocked"]
    pf.check_status()                          # nothing pending from earlier tests
    try:
        _lib.check(_lib.lib.pfh_set_tuning(gb, 1), "grad_blocked")
        _lib.check(_lib.lib.pfh_set_tuning(spin, 0), "panel_spin")
        la, ph = pf.slog_pfaffian(At, pf.PFH_METHOD_BLOCKED)
        assert math.isnan(float(la)) and torch.isnan(torch.as_tensor(ph)).all()
        # the status call reports the timeout once, then it is cleared
        with pytest.raises(_lib.PfaffianLibError, match="status -4"):
            pf.check_status()
        pf.check_status()
        G = slog_grad(At)
        assert torch.isnan(G).all()
        with pytest.raises(_lib.PfaffianLibError, match="status -4"):
            pf.check_status()
    finally:
        _lib.lib.pfh_set_tuning(spin, -1)
    try:
        la, ph = pf.slog_pfaffian(At, pf.PFH_METHOD_BLOCKED)
        assert abs(float(la) - l0) < 1e-9 and abs(complex(ph) - complex(p0)) < 1e-9
        G = slog_grad(At).cpu().numpy()
        Gi = 0.5 * np.linalg.inv(A).T
        assert np.abs(G - Gi).max() <= 1e-9 * np.abs(Gi).max()
        pf.check_status()                      # exact calls report nothing
    finally:
        _lib.lib.pfh_set_tuning(gb, -1)


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
@pytest.mark.parametrize("n", [300, 512])
def test_unblocked_large_vs_oracle(pf, cuda, dtype, n):
    rng = np.random.default_rng(n)
    A = oracle.skew_gaussian(rng, n, complex_=dtype.is_complex)
    gl, gp = pf.slog_pfaffian(torch.from_numpy(A).to(cuda), method=pf.PFH_METHOD_UNBLOCKED)
    l0, p0 = oracle.slog(A)
    assert abs(float(gl) - l0) < 1e-9
    assert abs(complex(gp.item()) - p0) < 1e-9


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
@pytest.mark.parametrize("n", [258, 300, 640, 1000])
def test_blocked_large_vs_oracle(pf, cuda, dtype, n):
    rng = np.random.default_rng(n + 1)
    A = oracle.skew_gaussian(rng, n, complex_=dtype.is_complex)
    gl, gp = pf.slog_pfaffian(torch.from_numpy(A).to(cuda), method=pf.PFH_METHOD_BLOCKED)
    l0, p0 = oracle.slog(A)
    assert abs(float(gl) - l0) < 1e-9, (float(gl), l0)
    assert abs(complex(gp.item()) - p0) < 1e-9


def test_blocked_zero_and_identity(pf, cuda):
    la, s = pf.slog_pfaffian(torch.zeros(300, 300, dtype=torch.float64, device=cuda),
                             method=pf.PFH_METHOD_BLOCKED)
    assert float(s) == 0 and float(la) == -math.inf
    rng = np.random.default_rng(seed=0)
    for n in (260, 290, 330):
        M, want = oracle.tridiag_identity_case(rng, n)
        la, s = pf.slog_pfaffian(torch.from_numpy(M).to(cuda), method=pf.PFH_METHOD_BLOCKED)
        assert np.isclose(want, float(s) * math.exp(float(la)), rtol=1e-10, atol=1e-12), n


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
def test_blocked_dies_mid_panel(pf, cuda, dtype):
    # 100 zero rows/columns scattered by a permutation: the first 250 pair
    # steps have nonzero pivots, then the trailing block is exactly zero, so
    # the factorization stops inside a panel with (-inf, 0).
    rng = np.random.default_rng(77)
    n = 600
    A = oracle.skew_gaussian(rng, n, complex_=dtype.is_complex)
    A[500:, :] = 0
    A[:, 500:] = 0
    perm = rng.permutation(n)
    A = np.ascontiguousarray(A[perm][:, perm])
    la, s = pf.slog_pfaffian(torch.from_numpy(A).to(cuda), method=pf.PFH_METHOD_BLOCKED)
    assert float(la) == -math.inf and complex(s.item()) == 0
    l0, p0 = oracle.slog(A)
    assert l0 == -math.inf and p0 == 0


def test_pfaffian_graph_replay(pf, cuda):
    # The dense product call (one kernel launch) captured in a graph: every
    # replay equals the eager result bit for bit, with eager work between.
    from pfcuda_amd import _lib
    rng = np.random.default_rng(13)
    A = oracle.skew_gaussian(rng, 32, batch=1024)
    At = torch.from_numpy(A).to(cuda)
    out = torch.empty(1024, dtype=torch.float64, device=cuda)

    def call():
        _lib.check(_lib.lib.pfh_pfaffian_f64(At.data_ptr(), 1024, 32, out.data_ptr(),
                                             torch.cuda.current_stream().cuda_stream), "pfaffian")
    call()
    torch.cuda.synchronize()
    want = out.clone()
    np.testing.assert_allclose(want[:16].cpu().numpy(), [oracle.pf_f64(a)[0] for a in A[:16]], rtol=1e-12)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        call()
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        call()
    for _ in range(3):
        out.fill_(0.0)
        (At @ At).sum().item()                 # eager work between replays
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, want)


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
def test_blocked_and_grad_graph_replay(pf, cuda, dtype):
    # The blocked slog and the large gradient enqueue a device-to-device copy
    # and plain kernel launches only (workspace from the caller), so both can
    # be captured into one graph: every replay equals the eager result bit
    # for bit, with eager work between the replays.
    import ctypes
    from pfcuda_amd import _lib
    lib = _lib.lib
    dt = _lib.PFH_C128 if dtype.is_complex else _lib.PFH_F64
    rng = np.random.default_rng(63)
    n = 600
    A0 = oracle.skew_gaussian(rng, n, complex_=dtype.is_complex)
    A = torch.from_numpy(A0).to(cuda)
    need, gneed = ctypes.c_size_t(0), ctypes.c_size_t(0)
    _lib.check(lib.pfh_slog_pfaffian_workspace_size(dt, 1, n, _lib.PFH_METHOD_BLOCKED, need), "workspace")
    _lib.check(lib.pfh_slog_pfaffian_grad_workspace_size(dt, 1, n, gneed), "grad workspace")
    ws = torch.empty(max(int(need.value), 1), dtype=torch.uint8, device=cuda)
    gws = torch.empty(max(int(gneed.value), 1), dtype=torch.uint8, device=cuda)
    lo = torch.empty(1, dtype=torch.float64, device=cuda)
    ph = torch.empty(1, dtype=dtype, device=cuda)
    glo = torch.empty(1, dtype=torch.float64, device=cuda)
    gph = torch.empty(1, dtype=dtype, device=cuda)
    G = torch.empty_like(A)

    def call():
        s = torch.cuda.current_stream().cuda_stream
        _lib.check(lib.pfh_slog_pfaffian_method(dt, A.data_ptr(), 1, n, _lib.PFH_METHOD_BLOCKED, lo.data_ptr(),
                                                ph.data_ptr(), ws.data_ptr(), ws.numel(), s), "slog")
        _lib.check(lib.pfh_slog_pfaffian_grad(dt, A.data_ptr(), 1, n, glo.data_ptr(), gph.data_ptr(), G.data_ptr(),
                                              gws.data_ptr(), gws.numel(), s), "grad")
    call()
    torch.cuda.synchronize()
    want = (lo.clone(), ph.clone(), G.clone())
    l0, p0 = oracle.slog(A0)
    assert abs(float(want[0]) - l0) < 1e-9 and abs(complex(want[1].item()) - p0) < 1e-9
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        call()
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        call()
    for _ in range(3):
        lo.fill_(0.0); ph.fill_(0.0); G.fill_(0.0)
        (A @ A).abs().sum().item()             # eager work between replays
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(lo, want[0]) and torch.equal(ph, want[1]) and torch.equal(G, want[2])


def test_ragged_graph_replay(pf, cuda):
    # The ragged call (device-side bucketing plus one launch per bucket, no
    # memsets, no host round trip) captured in a graph: every replay gives
    # the eager result bit for bit, with eager work between the replays.
    import ctypes
    from pfcuda_amd import _lib
    rng = np.random.default_rng(12)
    for dt, tdt in ((_lib.PFH_F64, torch.float64), (_lib.PFH_C128, torch.complex128)):
        sizes = (2 * rng.integers(8, 129, size=160)).astype(np.int32)
        mats = [oracle.skew_gaussian(rng, int(n), complex_=tdt.is_complex) for n in sizes]
        data = torch.from_numpy(np.concatenate([m.reshape(-1) for m in mats])).to(cuda)
        offs = torch.from_numpy(np.cumsum([0] + [int(n) ** 2 for n in sizes[:-1]]).astype(np.int64)).to(cuda)
        siz = torch.from_numpy(sizes).to(cuda)
        count, max_n = len(sizes), int(sizes.max())
        need = ctypes.c_size_t(0)
        _lib.check(_lib.lib.pfh_slog_pfaffian_ragged_workspace_size(dt, count, max_n, need), "workspace")
        ws = torch.empty(max(int(need.value), 1), dtype=torch.uint8, device=cuda)
        lo = torch.empty(count, dtype=torch.float64, device=cuda)
        ph = torch.empty(count, dtype=tdt, device=cuda)

        def call():
            _lib.check(_lib.lib.pfh_slog_pfaffian_ragged(
                dt, data.data_ptr(), offs.data_ptr(), siz.data_ptr(), count, max_n, lo.data_ptr(),
                ph.data_ptr(), ws.data_ptr(), ws.numel(), torch.cuda.current_stream().cuda_stream), "ragged")

        call()
        torch.cuda.synchronize()
        want_l, want_p = lo.clone(), ph.clone()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            call()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            call()
        for rep in range(3):
            lo.fill_(0)
            ph.fill_(0)
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(lo, want_l) and torch.equal(ph, want_p), (tdt, rep)
            call()
            torch.cuda.synchronize()
        wl, wp = oracle.slog_ragged(data.cpu().numpy(), offs.cpu().numpy(), sizes, threads=8)
        assert np.max(np.abs(want_l.cpu().numpy() - wl)) < 1e-9


def test_batched_dies_at_first_and_mid_panel(pf, cuda):
    # f64 slog at 64 < n <= 256 (the batched blocked kernel): an all-zero
    # matrix dies at the first pair step; 50 zero rows/columns scattered in an
    # n = 200 matrix leave 150 nonzero pivots, so it dies inside a panel;
    # a healthy matrix in the same launch is unaffected.
    rng = np.random.default_rng(91)
    z = np.zeros((130, 130))
    B = oracle.skew_gaussian(rng, 200)
    B[150:, :] = 0
    B[:, 150:] = 0
    perm = rng.permutation(200)
    B = np.ascontiguousarray(B[perm][:, perm])
    C = oracle.skew_gaussian(rng, 256)
    mats, sizes = [z, B, C], [130, 200, 256]
    data = np.concatenate([m.reshape(-1) for m in mats])
    offs = np.cumsum([0] + [n * n for n in sizes[:-1]])
    gl, gp = pf.slog_pfaffian_ragged(torch.from_numpy(data).to(cuda), offs, sizes)
    gl, gp = gl.cpu().numpy(), gp.cpu().numpy()
    assert gl[0] == -math.inf and gp[0] == 0
    assert gl[1] == -math.inf and gp[1] == 0
    l0, p0 = oracle.slog(C)
    assert abs(gl[2] - l0) < 1e-9 and gp[2] == p0
    # the dense entry point on the same kind of batch
    la, s = pf.slog_pfaffian(torch.from_numpy(np.stack([B, B])).to(cuda))
    assert (la.cpu().numpy() == -math.inf).all() and (s.cpu().numpy() == 0).all()


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
@pytest.mark.parametrize("n", [8, 32, 64, 128, 200, 300])
def test_grad_is_half_inverse_transpose(pf, cuda, dtype, n):
    from pfcuda_amd.adjoint import slog_grad
    if dtype.is_complex and n == 200:
        pytest.skip("covered by 300 (global sweep) and 128 (register sweep)")
    rng = np.random.default_rng(100 + n)
    A = oracle.skew_gaussian(rng, n, complex_=dtype.is_complex, batch=2)
    G, la, ph = slog_grad(torch.from_numpy(A).to(cuda), return_slog=True)
    G = G.cpu().numpy()
    for b in range(2):
        want = 0.5 * np.linalg.inv(A[b]).T
        err = np.abs(G[b] - want).max() / np.abs(want).max()
        assert err < 1e-10, (n, err)
        l0, p0 = oracle.slog(A[b])
        assert abs(float(la[b]) - l0) < 1e-9 and abs(complex(ph[b].item()) - p0) < 1e-9


def test_grad_autograd_and_finite_differences(pf, cuda):
    # C5 check: central finite differences (h = 1e-6) along skew directions
    rng = np.random.default_rng(55)
    n = 40
    A0 = oracle.skew_gaussian(rng, n)
    A = torch.from_numpy(A0).to(cuda).requires_grad_(True)
    la, _ = pf.slog_pfaffian(A)
    (G,) = torch.autograd.grad(la, A)
    G = G.cpu().numpy()
    h = 1e-6
    for (i, j) in [(0, 1), (3, 17), (10, 39), (25, 26)]:
        E = np.zeros((n, n)); E[i, j] = 1.0; E[j, i] = -1.0
        fp = oracle.slog(A0 + h * E)[0]
        fm = oracle.slog(A0 - h * E)[0]
        fd = (fp - fm) / (2 * h)
        assert abs((G[i, j] - G[j, i]) - fd) < 1e-6 * max(1.0, abs(fd)), (i, j, G[i, j] - G[j, i], fd)


@pytest.mark.parametrize("blocked", ["1", "0"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
@pytest.mark.parametrize("n", [520, 530, 1000])
def test_grad_large_sweeps_vs_inverse(pf, cuda, monkeypatch, dtype, n, blocked):
    # The blocked block sweep (groups of 32 pairs, 16 for complex128; the last
    # group is partial, of 4, 9 and 20 / 4 pairs) and the unblocked one
    # (PFH_GRAD_BLOCKED=0) against numpy's inverse and the oracle's slog.
    from pfcuda_amd.adjoint import slog_grad
    monkeypatch.setenv("PFH_GRAD_BLOCKED", blocked)
    rng = np.random.default_rng(300 + n)
    A = oracle.skew_gaussian(rng, n, complex_=dtype.is_complex)
    G, la, ph = slog_grad(torch.from_numpy(A).to(cuda), return_slog=True)
    want = 0.5 * np.linalg.inv(A).T
    err = np.abs(G.cpu().numpy() - want).max() / np.abs(want).max()
    assert err < 1e-10, (n, blocked, err)
    l0, p0 = oracle.slog(A)
    assert abs(float(la) - l0) < 1e-9 and abs(complex(ph.item()) - p0) < 1e-9


@pytest.mark.parametrize("dtype", [torch.float32, torch.complex64])
def test_grad_large_single_precision(pf, cuda, dtype):
    # the blocked sweep's plain trailing update (no matrix cores)
    from pfcuda_amd.adjoint import slog_grad
    rng = np.random.default_rng(31)
    n = 300
    A = oracle.skew_gaussian(rng, n, complex_=dtype.is_complex)
    G = slog_grad(torch.from_numpy(A).to(cuda).to(dtype)).cpu().numpy()
    want = 0.5 * np.linalg.inv(A).T
    assert np.abs(G - want).max() / np.abs(want).max() < 1e-2


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
@pytest.mark.parametrize("zero_from", [0, 500])
def test_grad_large_singular(pf, cuda, dtype, zero_from):
    # an exactly singular matrix: the sweep stops at its zero pivot (in the
    # first group, or after 250 pair steps) and returns (-inf, 0) and a NaN
    # gradient, as the unblocked sweep does
    from pfcuda_amd.adjoint import slog_grad
    rng = np.random.default_rng(78)
    n = 600
    A = oracle.skew_gaussian(rng, n, complex_=dtype.is_complex)
    A[zero_from:, :] = 0
    A[:, zero_from:] = 0
    perm = rng.permutation(n)
    A = np.ascontiguousarray(A[perm][:, perm])
    G, la, ph = slog_grad(torch.from_numpy(A).to(cuda), return_slog=True)
    assert float(la) == -math.inf and complex(ph.item()) == 0
    assert torch.isnan(G).all()


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
def test_trailing_forms_bitwise_equal(pf, cuda, monkeypatch, dtype):
    # The LDS-staged and register forms of the trailing updates issue the
    # same products in the same order, so they agree bit for bit (blocked
    # factorization at n = 600, blocked sweep at n = 520).
    from pfcuda_amd.adjoint import slog_grad
    rng = np.random.default_rng(61)
    A = oracle.skew_gaussian(rng, 600, complex_=dtype.is_complex)
    B = oracle.skew_gaussian(rng, 520, complex_=dtype.is_complex)
    At, Bt = torch.from_numpy(A).to(cuda), torch.from_numpy(B).to(cuda)
    res = {}
    for form in ("1", "0"):
        monkeypatch.setenv("PFH_TRAIL_STAGED", form)
        la, s = pf.slog_pfaffian(At, method=pf.PFH_METHOD_BLOCKED)
        G = slog_grad(Bt)
        res[form] = (float(la), complex(s.item()), G.cpu().numpy())
    assert res["1"][0] == res["0"][0] and res["1"][1] == res["0"][1]
    assert np.array_equal(res["1"][2], res["0"][2])
    l0, p0 = oracle.slog(A)
    assert abs(res["1"][0] - l0) < 1e-9 and abs(res["1"][1] - p0) < 1e-9


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128, torch.float32])
def test_panel_exchange_forms_bitwise_equal(pf, cuda, dtype):
    # The panel's participants (one per wave, the default, or one per block
    # of 256 slots) exchange candidates through one XCD's L2, or (blocks)
    # device-wide; the arithmetic and its order are the same, so the three
    # forms agree bit for bit (blocked factorization at n = 1000, whose
    # panels cross several blocks, and the blocked sweep at n = 600).
    from pfcuda_amd import _lib
    from pfcuda_amd.adjoint import slog_grad
    rng = np.random.default_rng(62)
    A = oracle.skew_gaussian(rng, 1000, complex_=dtype.is_complex)
    B = oracle.skew_gaussian(rng, 600, complex_=dtype.is_complex)
    At, Bt = torch.from_numpy(A).to(cuda).to(dtype), torch.from_numpy(B).to(cuda).to(dtype)
    key = _lib.TUNE_KEYS["panel_xcd"]
    res = {}
    try:
        for form in (2, 1, 0):
            _lib.check(_lib.lib.pfh_set_tuning(key, form), "panel_xcd")
            la, s = pf.slog_pfaffian(At, method=pf.PFH_METHOD_BLOCKED)
            G = slog_grad(Bt) if dtype != torch.float32 else torch.zeros(1)
            res[form] = (float(la), complex(s.item()), G.cpu().numpy())
        pf.check_status()
    finally:
        _lib.lib.pfh_set_tuning(key, -1)
    for form in (1, 0):
        assert res[2][0] == res[form][0] and res[2][1] == res[form][1], form
        assert np.array_equal(res[2][2], res[form][2]), form
    if dtype != torch.float32:
        l0, p0 = oracle.slog(A)
        assert abs(res[2][0] - l0) < 1e-9 and abs(res[2][1] - p0) < 1e-9


def test_panel_forms_switch_within_one_solve(pf, cuda):
    # complex128 at n = 8320: the per-wave panel holds at most 128 waves
    # (8192 columns), so the first panels run in the per-block form and the
    # later ones per wave, each form with its own epoch count; the solve
    # equals, bit for bit, the one with every panel in the per-block forms
    # (PFH_PANEL_XCD = 1: XCD-local while the blocks fit one XCD, else
    # device-wide), and agrees with the unblocked path.
    from pfcuda_amd import _lib
    rng = np.random.default_rng(83)
    n = 8320
    A = oracle.skew_gaussian(rng, n, complex_=True)
    At = torch.from_numpy(A).to(cuda)
    key = _lib.TUNE_KEYS["panel_xcd"]
    res = {}
    try:
        for form in (2, 1):
            _lib.check(_lib.lib.pfh_set_tuning(key, form), "panel_xcd")
            la, ph = pf.slog_pfaffian(At, method=pf.PFH_METHOD_BLOCKED)
            res[form] = (float(la), complex(ph.item()))
        pf.check_status()
    finally:
        _lib.lib.pfh_set_tuning(key, -1)
    assert res[2] == res[1]
    lu, pu = pf.slog_pfaffian(At, method=pf.PFH_METHOD_UNBLOCKED)
    assert abs(res[2][0] - float(lu)) < 1e-8 and abs(res[2][1] - complex(pu.item())) < 1e-8


def test_grad_n2048_residual(pf, cuda):
    # C5 at full size: A G^T = I/2 (G = 1/2 A^-T), and the slog against the oracle
    from pfcuda_amd.adjoint import slog_grad
    rng = np.random.default_rng(2048)
    n = 2048
    A = oracle.skew_gaussian(rng, n)
    At = torch.from_numpy(A).to(cuda)
    G, la, ph = slog_grad(At, return_slog=True)
    R = At @ G.T - 0.5 * torch.eye(n, dtype=torch.float64, device=cuda)
    rel = float(R.abs().max()) / (float(At.abs().max()) * float(G.abs().max()) * n)
    assert rel < 1e-13, rel
    l0, p0 = oracle.slog(A)
    assert abs(float(la) - l0) < 1e-8 and float(ph) == p0


@pytest.mark.parametrize("key,lo,hi,values,cplx", [
    ("f64_bucket_128", 64, 128, (0, 2, 3), False),
    ("f64_bucket_192", 128, 192, (0, 2, 3), False),
    ("f64_bucket_256", 192, 256, (0, 2, 3), False),
    ("c128_bucket_128", 64, 128, (0, 3), True),
])
def test_every_bucket_kernel_vs_oracle(pf, cuda, key, lo, hi, values, cplx):
    # Each kernel a size bucket can run (column per lane, LDS-resident with
    # its spill slab above 192, batched blocked), chosen through the tuning
    # key, on a ragged batch of that bucket: the oracle, and the same
    # results whichever kernel runs.
    from pfcuda_amd import autotune
    rng = np.random.default_rng(hi + (7 if cplx else 0))
    sizes = [int(s) for s in 2 * rng.integers(lo // 2 + 1, hi // 2 + 1, size=24)] + [hi, lo + 2]
    mats = [oracle.skew_gaussian(rng, n, complex_=cplx) for n in sizes]
    data, offsets, sz = pf.pack_ragged([torch.from_numpy(m) for m in mats])
    want = [oracle.slog(m) for m in mats]
    try:
        for v in values:
            autotune.set_tuning(key, v)
            la, ph = pf.slog_pfaffian_ragged(data.to(cuda), offsets, sz)
            la, ph = la.cpu().numpy(), ph.cpu().numpy()
            for i, (l0, p0) in enumerate(want):
                assert abs(la[i] - l0) < 1e-9 and abs(complex(ph[i]) - p0) < 1e-9, (key, v, sizes[i])
    finally:
        autotune.set_tuning(key, -1)


def test_per_step_blocked_path_and_unblocked_sweep(cuda):
    # With the panel kernel off (PFH_BLOCKED_COOP=0, read once per process,
    # hence a child process): the blocked factorization launches once per
    # pair step and the large gradient uses the unblocked sweep; both against
    # the oracle and numpy's inverse.
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = r"""
import sys, numpy as np, torch
sys.path.insert(0, ROOT); sys.path.insert(0, ROOT + "/tests")
import oracle, pfcuda_amd as pf
from pfcuda_amd.adjoint import slog_grad
rng = np.random.default_rng(91)
for cplx in (False, True):
    A = oracle.skew_gaussian(rng, 600, complex_=cplx)
    la, s = pf.slog_pfaffian(torch.from_numpy(A).cuda(), method=pf.PFH_METHOD_BLOCKED)
    l0, p0 = oracle.slog(A)
    assert abs(float(la) - l0) < 1e-9 and abs(complex(s.item()) - p0) < 1e-9, (cplx, float(la), l0)
    B = oracle.skew_gaussian(rng, 300, complex_=cplx)
    G = slog_grad(torch.from_numpy(B).cuda()).cpu().numpy()
    want = 0.5 * np.linalg.inv(B).T
    assert np.abs(G - want).max() / np.abs(want).max() < 1e-10
print("per-step ok", flush=True)
""".replace("ROOT", repr(root))
    env = dict(os.environ, PFH_BLOCKED_COOP="0")
    r = subprocess.run([sys.executable, "-c", script], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "per-step ok" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]


def test_blocked_panel_wider_than_the_panel_kernel(pf, cuda):
    # n = 16448: the first panel spans 65 blocks of 256 columns, one more than
    # the panel kernel takes, so it runs as per-step launches, and the panels
    # after it on the panel kernel.  Against the unblocked path (independent
    # kernels) on the same matrix; the oracle would take minutes here.
    n = 16448
    g = torch.Generator(device=cuda)
    g.manual_seed(16448)
    X = torch.randn(n, n, generator=g, device=cuda, dtype=torch.float64)
    A = ((X - X.T) / 2).contiguous()
    del X
    lb, sb = pf.slog_pfaffian(A, method=pf.PFH_METHOD_BLOCKED)
    lu, su = pf.slog_pfaffian(A, method=pf.PFH_METHOD_UNBLOCKED)
    assert math.isfinite(float(lb)) and abs(float(lb) - float(lu)) < 1e-8, (float(lb), float(lu))
    assert float(sb) == float(su)
