"""pfaffian_cpu and pfaffian_py, the library's versions of the reference's
CPU surface (pfcuda/__init__.py:1-5): against the products the reference's
own CPU code returned (tests/golden/ref_product.npz), the oracle and the
closed forms, and the reference binding's argument handling."""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def unpack(data, sizes):
    off = 0
    for n in sizes:
        yield data[off:off + n * n].reshape(n, n)
        off += n * n


def test_against_reference_products():
    import pfcuda_amd as pf
    g = np.load(os.path.join(GOLD, "ref_product.npz"))
    for A, want in zip(unpack(g["data"], g["sizes"]), g["pf"]):
        n = A.shape[0]
        for f in (pf.pfaffian_cpu, pf.pfaffian_py):
            got = float(f(A.copy()))
            if n <= 4:
                assert got == want, (f.__name__, n, got, want)          # the same closed forms
            else:
                assert abs(got - want) <= 1e-12 * abs(want), (f.__name__, n, got, want)


@pytest.mark.parametrize("n", [6, 20, 64, 130])
def test_against_oracle(n):
    import pfcuda_amd as pf
    rng = np.random.default_rng(n)
    A = oracle.skew_gaussian(rng, n)
    want = oracle.pf_f64(A)[0]
    assert abs(pf.pfaffian_cpu(A.copy()) - want) <= 1e-12 * abs(want)
    assert abs(pf.pfaffian_py(A.copy()) - want) <= 1e-12 * abs(want)
    C = oracle.skew_gaussian(rng, n, complex_=True)
    wc = oracle.pf_c128(C)[0]
    assert abs(pf.pfaffian_py(C.copy()) - wc) <= 1e-12 * abs(wc)
    if abs(want) < 1e30:      # float32's product form overflows past ~3e38, as the reference's does
        F = A.astype(np.float32)
        assert abs(pf.pfaffian_py(F.copy()) - want) <= 1e-3 * abs(want)


def test_argument_handling_and_edge_cases():
    import pfcuda_amd as pf
    with pytest.raises(RuntimeError, match="Input must be 2D"):
        pf.pfaffian_cpu(np.zeros(4))
    with pytest.raises(RuntimeError, match="Matrix must be square"):
        pf.pfaffian_cpu(np.zeros((2, 4)))
    with pytest.raises(RuntimeError, match="n must be even"):
        pf.pfaffian_cpu(np.zeros((3, 3)))
    assert pf.pfaffian_cpu(np.zeros((0, 0))) == 0.0
    assert pf.pfaffian_py(np.zeros((3, 3))) == 0.0 and pf.pfaffian_py(np.zeros((0, 0))) == 0.0
    # exactly singular: the exact zero-pivot test gives 0
    rng = np.random.default_rng(3)
    A = oracle.skew_gaussian(rng, 10)
    A[7, :] = 0
    A[:, 7] = 0
    assert pf.pfaffian_cpu(A.copy()) == 0.0 and pf.pfaffian_py(A.copy()) == 0.0
    # in place on a C-contiguous float64 input (as the reference's binding),
    # on a converted copy otherwise
    B = oracle.skew_gaussian(rng, 8)
    B0 = B.copy()
    pf.pfaffian_cpu(B)
    assert not np.array_equal(B, B0)
    B32 = B0.astype(np.float32)
    B32c = B32.copy()
    pf.pfaffian_cpu(B32)
    assert np.array_equal(B32, B32c)
    # sign under an interchange of two index pairs' roles: Pf(P A P^T) = det(P) Pf(A)
    Ap = B0[[1, 0] + list(range(2, 8))][:, [1, 0] + list(range(2, 8))]
    assert abs(pf.pfaffian_cpu(Ap.copy()) + pf.pfaffian_cpu(B0.copy())) <= 1e-13 * abs(pf.pfaffian_cpu(B0.copy()))
