"""The operators under torch's function transforms, compilation and graph
capture: the torch counterpart of the reference's jit / vmap / grad use of
its FFI targets (pfcuda/cuda_api.py:10-107).

What pins each result:
- vmap against the same batch through the plain call (bitwise: one launch
  over the same matrices either way) and against the oracle;
- first derivatives against central finite differences along random skew
  directions (h = 1e-6, rel 1e-6) and against 1/2 A^-T from an FP64 inverse;
- second derivatives (through the adjoint operator's own derivative) against
  central finite differences of the first derivative (rel 1e-5);
- torch.compile (fullgraph=True for the forward) and captured graphs
  against eager (bitwise).
"""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu

H = 1e-6


@pytest.fixture(scope="module")
def pf(cuda):
    import pfcuda_amd
    return pfcuda_amd


def skew(rng, n, batch=None, complex_=False):
    return torch.from_numpy(oracle.skew_gaussian(rng, n, complex_=complex_, batch=batch))


def skew_dir(rng, n, complex_):
    V = skew(rng, n, complex_=complex_)
    return V / V.abs().max()


def directional(grad, V):
    """Directional derivative of a real function from torch's gradient of it
    (the conjugate convention for complex inputs): Re sum(conj(grad) * V)."""
    return float((grad.conj() * V).sum().real)


def logpf(pf):
    return lambda A: pf.slog_pfaffian(A)[0]


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
def test_vmap_slog_is_one_batched_launch(pf, cuda, dtype):
    rng = np.random.default_rng(10)
    A = skew(rng, 40, batch=6, complex_=dtype.is_complex).to(dtype).to(cuda)
    want_l, want_p = pf.slog_pfaffian(A)
    got_l, got_p = torch.func.vmap(pf.slog_pfaffian)(A)
    assert torch.equal(got_l, want_l) and torch.equal(got_p, want_p)
    # a vmapped dimension that is not the leading one
    got_l1, got_p1 = torch.func.vmap(pf.slog_pfaffian, in_dims=1)(A.movedim(0, 1))
    assert torch.equal(got_l1, want_l) and torch.equal(got_p1, want_p)
    # nested: (2, 3) batch of matrices
    got_l2, _ = torch.func.vmap(torch.func.vmap(pf.slog_pfaffian))(A.reshape(2, 3, 40, 40))
    assert torch.equal(got_l2.reshape(6), want_l)
    for a, g in zip(A.cpu().numpy(), got_l.cpu().numpy()):
        assert abs(g - oracle.slog(a)[0]) <= 1e-9


def test_vmap_pfaffian(pf, cuda):
    rng = np.random.default_rng(11)
    A = skew(rng, 32, batch=8).to(cuda)
    assert torch.equal(torch.func.vmap(pf.pfaffian)(A), pf.pfaffian(A))
    # the closed forms (n = 4) are torch operations and map as well
    B = skew(rng, 4, batch=5).to(cuda)
    assert torch.allclose(torch.func.vmap(pf.pfaffian)(B), pf.pfaffian(B), rtol=0, atol=0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
def test_func_grad_against_fd_and_inverse(pf, cuda, dtype):
    rng = np.random.default_rng(12)
    cplx = dtype.is_complex
    A = skew(rng, 48, complex_=cplx).to(cuda)
    f = logpf(pf)
    G = torch.func.grad(f)(A)
    Gi = 0.5 * torch.linalg.inv(A).mT
    want = Gi.conj() if cplx else Gi
    assert torch.allclose(G, want, rtol=1e-9, atol=1e-9 * float(want.abs().max()))
    for _ in range(2):
        V = skew_dir(rng, 48, cplx).to(cuda)
        fd = (float(f(A + H * V)) - float(f(A - H * V))) / (2 * H)
        d = directional(G, V)
        assert abs(d - fd) <= 1e-6 * max(1.0, abs(fd)), (d, fd)


def test_complex_phase_reverse_mode_against_fd(pf, cuda):
    # a real function of the phase alone, and one of log|Pf| and the phase
    rng = np.random.default_rng(19)
    A = skew(rng, 44, complex_=True).to(cuda)
    w = complex(rng.normal(), rng.normal())

    def psi(X):
        la, ph = pf.slog_pfaffian(X)
        return (w * ph).real

    def chi(X):
        la, ph = pf.slog_pfaffian(X)
        return la * 0.3 + (w * ph).imag

    for fn in (psi, chi):
        G = torch.func.grad(fn)(A)
        Ar = A.clone().requires_grad_()
        (g2,) = torch.autograd.grad(fn(Ar), Ar)
        assert torch.equal(G, g2)
        for _ in range(2):
            V = skew_dir(rng, 44, True).to(cuda)
            fd = (float(fn(A + H * V)) - float(fn(A - H * V))) / (2 * H)
            d = directional(G, V)
            assert abs(d - fd) <= 1e-6 * max(1.0, abs(fd)), (fn.__name__, d, fd)


def test_vmap_of_grad_per_matrix(pf, cuda):
    rng = np.random.default_rng(13)
    A = skew(rng, 40, batch=5).to(cuda)
    G = torch.func.vmap(torch.func.grad(logpf(pf)))(A)
    from pfcuda_amd.adjoint import slog_grad
    assert torch.equal(G, slog_grad(A))
    Ar = A.clone().requires_grad_()
    (g_ref,) = torch.autograd.grad(pf.slog_pfaffian(Ar)[0].sum(), Ar)
    assert torch.equal(G, g_ref)


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
def test_func_jvp_against_fd(pf, cuda, dtype):
    rng = np.random.default_rng(14)
    cplx = dtype.is_complex
    A = skew(rng, 40, complex_=cplx).to(cuda)
    V = skew_dir(rng, 40, cplx).to(cuda)
    (l, p), (dl, dp) = torch.func.jvp(pf.slog_pfaffian, (A,), (V,))
    lp, pp = pf.slog_pfaffian(A + H * V)
    lm, pm = pf.slog_pfaffian(A - H * V)
    fd = (float(lp) - float(lm)) / (2 * H)
    assert abs(float(dl) - fd) <= 1e-6 * max(1.0, abs(fd))
    if cplx:
        fdp = (complex(pp) - complex(pm)) / (2 * H)
        assert abs(complex(dp) - fdp) <= 1e-6 * max(1.0, abs(fdp))
    else:
        assert dp is None or float(dp) == 0.0
    # vmap of jvp: tangents of a batch in one launch of each operator
    Ab = skew(rng, 40, batch=3, complex_=cplx).to(cuda)
    Vb = torch.stack([skew_dir(rng, 40, cplx) for _ in range(3)]).to(cuda)
    dlb = torch.func.vmap(lambda a, v: torch.func.jvp(pf.slog_pfaffian, (a,), (v,))[1][0])(Ab, Vb)
    for i in range(3):
        _, (dli, _) = torch.func.jvp(pf.slog_pfaffian, (Ab[i],), (Vb[i],))
        assert torch.allclose(dlb[i], dli, rtol=1e-12, atol=0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
def test_second_derivative_against_fd(pf, cuda, dtype):
    rng = np.random.default_rng(15)
    cplx = dtype.is_complex
    n = 40
    A = skew(rng, n, complex_=cplx).to(cuda)
    W = skew(rng, n, complex_=cplx).to(cuda)
    f = logpf(pf)

    def phi(X):                       # a real scalar of the gradient
        return (W * torch.func.grad(f)(X)).sum().real

    Hphi = torch.func.grad(phi)(A)    # reverse over reverse
    V0 = skew_dir(rng, n, cplx).to(cuda)
    _, dphi = torch.func.jvp(phi, (A,), (V0,))   # forward over reverse
    assert abs(float(dphi) - directional(Hphi, V0)) <= 1e-10 * max(1.0, abs(float(dphi)))
    Ar = A.clone().requires_grad_()
    (g1,) = torch.autograd.grad(f(Ar), Ar, create_graph=True)
    (h2,) = torch.autograd.grad((W * g1).sum().real, Ar)
    assert torch.allclose(Hphi, h2, rtol=1e-12, atol=1e-12 * float(h2.abs().max()))
    for _ in range(2):
        V = skew_dir(rng, n, cplx).to(cuda)
        fd = (float(phi(A + H * V)) - float(phi(A - H * V))) / (2 * H)
        d = directional(Hphi, V)
        assert abs(d - fd) <= 1e-5 * max(1.0, abs(fd)), (d, fd)


def test_compile_fullgraph(pf, cuda):
    rng = np.random.default_rng(16)
    A = skew(rng, 40, batch=4).to(cuda)
    B = skew(rng, 32, batch=4).to(cuda)

    def fwd(X, Y):
        la, s = pf.slog_pfaffian(X)
        return la * s, pf.pfaffian(Y)

    cf = torch.compile(fwd, backend="aot_eager", fullgraph=True)
    got = cf(A, B)
    want = fwd(A, B)
    assert torch.equal(got[0], want[0]) and torch.equal(got[1], want[1])

    def loss(X):
        return pf.slog_pfaffian(X)[0].sum()

    # With inputs that require grad, dynamo does not trace an autograd.Function
    # that defines a custom jvp: the compiled loss breaks its graph there and
    # runs the Function eagerly, with the same result.
    Xc = A.clone().requires_grad_()
    torch.compile(loss, backend="aot_eager")(Xc).backward()
    Xe = A.clone().requires_grad_()
    loss(Xe).backward()
    assert torch.equal(Xc.grad, Xe.grad)


def test_opcheck_operators(pf, cuda):
    rng = np.random.default_rng(17)
    A = skew(rng, 40, batch=3).to(cuda)
    B = skew(rng, 32, batch=3).to(cuda)
    torch.library.opcheck(torch.ops.pfcuda.pfaffian.default, (B,))
    torch.library.opcheck(torch.ops.pfcuda.slog_pfaffian.default, (A, 0))
    torch.library.opcheck(torch.ops.pfcuda.slog_pfaffian_grad.default, (A,))
    mats = [skew(rng, n) for n in (16, 70, 130)]
    data, off, siz = pf.pack_ragged(mats)
    torch.library.opcheck(torch.ops.pfcuda.slog_pfaffian_ragged.default,
                          (data.to(cuda), off.to(cuda), siz.to(cuda), 130))


def test_graph_capture_through_public_api(pf, cuda):
    rng = np.random.default_rng(18)
    A = skew(rng, 40, batch=8).to(cuda)
    S = skew(rng, 520).to(cuda)          # the blocked single-matrix path
    want = (pf.slog_pfaffian(A), pf.slog_pfaffian(S))
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        pf.slog_pfaffian(A), pf.slog_pfaffian(S)
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        out = (pf.slog_pfaffian(A), pf.slog_pfaffian(S))
    for r in range(3):
        # eager work and a synchronisation between replays
        A.mul_(1.0)
        junk = torch.randn(1 << 16, device=cuda).sum()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for (gl, gp), (wl, wp) in zip(out, want):
            assert torch.equal(gl, wl) and torch.equal(gp, wp), r
    del junk
