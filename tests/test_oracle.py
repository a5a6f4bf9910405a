"""CPU pins of the oracle (no GPU).  Each oracle function is checked against
something that does not share its arithmetic: the reference's own compiled
code (oracle/_ref and the committed ref_product.npz), closed forms, brute
force over perfect matchings, mpmath at 50 digits, and the reference's own
known-answer tests (test/pfaffian_test.py, test/slog_pfaffian_test.py)."""
import math
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


# ---- bitwise parity with the reference's CPU code -------------------------

def test_stale_mode_matches_reference_fixture_bitwise():
    g = np.load(os.path.join(GOLD, "ref_product.npz"))
    for A, want in zip(unpack(g["data"], g["sizes"]), g["pf"]):
        got = oracle.pf_f64(A, stale=True, closed_forms=True)[0]
        assert got == want, (A.shape, got, want)


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (reference tree absent)")
def test_stale_mode_matches_compiled_reference_bitwise():
    rng = np.random.default_rng(5)
    for n in list(range(2, 41, 2)) + [64, 100]:
        for _ in range(5):
            A = oracle.skew_gaussian(rng, n)
            assert oracle.pf_f64(A, stale=True, closed_forms=True)[0] == oracle.ref_pfaffian(A)
    # zero matrices and a rank-deficient one go through the exact zero test
    for n in (6, 10):
        Z = np.zeros((n, n))
        assert oracle.pf_f64(Z, stale=True)[0] == oracle.ref_pfaffian(Z) == 0.0


# ---- closed forms and brute force -----------------------------------------

def test_closed_forms_n2_n4():
    rng = np.random.default_rng(1)
    for _ in range(10):
        A = oracle.skew_gaussian(rng, 2)
        assert oracle.pf_f64(A)[0] == pytest.approx(A[0, 1], rel=1e-15)
        A = oracle.skew_gaussian(rng, 4)
        want = A[0, 1] * A[2, 3] - A[0, 2] * A[1, 3] + A[0, 3] * A[1, 2]
        assert oracle.pf_f64(A)[0] == pytest.approx(want, rel=1e-13)


@pytest.mark.parametrize("n", [2, 4, 6, 8, 10])
def test_brute_force_integer(n):
    rng = np.random.default_rng(n)
    for _ in range(4):
        B = rng.integers(-6, 7, size=(n, n))
        A = B - B.T
        exact = oracle.pfaffian_brute(A.tolist())
        assert exact * exact == round(np.linalg.det(A))   # the brute force itself: Pf^2 = det
        got = oracle.pf_f64(A.astype(float))[0]
        assert got == pytest.approx(float(exact), rel=1e-12, abs=1e-9)
        got_py = oracle.pfaffian_py(A.astype(float))
        assert got_py == pytest.approx(float(exact), rel=1e-12, abs=1e-9)


def test_brute_force_complex():
    rng = np.random.default_rng(3)
    for n in (4, 6, 8):
        A = oracle.skew_gaussian(rng, n, complex_=True)
        exact = oracle.pfaffian_brute(A.tolist())
        assert oracle.pf_c128(A)[0] == pytest.approx(exact, rel=1e-12)
        assert oracle.pfaffian_py(A) == pytest.approx(exact, rel=1e-12)


# ---- mpmath --------------------------------------------------------------

@pytest.mark.parametrize("kind", ["real", "complex"])
def test_against_mpmath_fixture(kind):
    g = np.load(os.path.join(GOLD, "mp_slog.npz"))
    for A, la, ph in zip(unpack(g[f"{kind}_data"], g[f"{kind}_sizes"]), g[f"{kind}_log"],
                         g[f"{kind}_phase"]):
        got_la, got_ph = oracle.slog(A)
        assert abs(got_la - la) < 1e-11
        assert abs(got_ph - ph) < 1e-10
        py_la, py_ph = oracle.pfaffian_py(A, return_log=True)
        assert abs(py_la - la) < 1e-11 and abs(py_ph - ph) < 1e-10


@pytest.mark.parametrize("kind", ["real", "complex"])
def test_long_double_runs_against_mpmath_fixture(kind):
    # pf_f64_ld / pf_c128_ld against the committed 50-digit values: tighter
    # than the double runs are held to
    g = np.load(os.path.join(GOLD, "mp_slog.npz"))
    for A, la, ph in zip(unpack(g[f"{kind}_data"], g[f"{kind}_sizes"]), g[f"{kind}_log"],
                         g[f"{kind}_phase"]):
        if kind == "complex":
            _, got_la, got_ph, _ = oracle.pf_c128_ld(A)
        else:
            _, got_la, got_ph, _ = oracle.pf_f64_ld(A)
        assert abs(got_la - la) < 2e-14 * max(1.0, abs(la))
        assert abs(got_ph - ph) < 1e-14


def test_long_double_complex_run_is_closer_to_mpmath():
    rng = np.random.default_rng(12)
    A = oracle.skew_gaussian(rng, 40, complex_=True)
    mla, mph, _ = oracle.pfaffian_mp(A, 60)
    _, la64, ph64, _ = oracle.pf_c128(A)
    _, la80, ph80, _ = oracle.pf_c128_ld(A)
    assert abs(la80 - mla) <= abs(la64 - mla) + 1e-15
    assert abs(la80 - mla) < 1e-14 and abs(ph80 - mph) < 1e-14


def test_long_double_complex_brute_force_gaussian_integers():
    # Gaussian-integer skew matrices: the expansion over perfect matchings is
    # exact, and the extended run must give it to ~1e-17 relative; a dropped
    # term, a sign or a conjugate in the complex path fails by O(1)
    rng = np.random.default_rng(13)
    for n in (4, 6, 8, 10):
        for _ in range(3):
            B = rng.integers(-5, 6, size=(n, n)) + 1j * rng.integers(-5, 6, size=(n, n))
            A = B - B.T
            exact = complex(oracle.pfaffian_brute(A.tolist()))
            prod, la, ph, _ = oracle.pf_c128_ld(A)
            if exact == 0:
                assert abs(prod) < 1e-9
                continue
            assert abs(prod - exact) <= 1e-15 * abs(exact)
            assert abs(la - math.log(abs(exact))) < 1e-15 * max(1.0, abs(math.log(abs(exact))))
            assert abs(ph - exact / abs(exact)) < 1e-15


@pytest.mark.parametrize("cplx", [False, True])
def test_threaded_oracle_is_bitwise_equal(cplx):
    # the threaded form splits each pair step's column loops; every column
    # keeps its operation order, so the results are the same bits
    rng = np.random.default_rng(14)
    for n in (6, 130, 512, 700):
        A = oracle.skew_gaussian(rng, n, complex_=cplx)
        want = oracle.pf_c128(A) if cplx else oracle.pf_f64(A)
        for threads in (1, 3, 8):
            got = oracle.pf_c128_mt(A, threads) if cplx else oracle.pf_f64_mt(A, threads)
            # the product form overflows to NaN/inf past n ~ 600: equal as NaN
            assert np.array_equal(np.array(got, dtype=complex), np.array(want, dtype=complex), equal_nan=True)


def test_long_double_run_is_closer_to_mpmath():
    rng = np.random.default_rng(11)
    A = oracle.skew_gaussian(rng, 40)
    mla, mph, _ = oracle.pfaffian_mp(A, 60)
    _, la64, s64, _ = oracle.pf_f64(A)
    _, la80, s80, _ = oracle.pf_f64_ld(A)
    assert s64 == s80 == mph
    assert abs(la80 - mla) <= abs(la64 - mla) + 1e-15
    assert abs(la80 - mla) < 1e-14


# ---- the reference's known-answer tests, on the oracle --------------------

def test_reference_identity_small():
    # test/pfaffian_test.py:80-110, seed 0, n = 2..32, 10 each
    rng = np.random.default_rng(seed=0)
    for n in range(2, 33, 2):
        for _ in range(10):
            M, want = oracle.tridiag_identity_case(rng, n)
            got = oracle.pf_f64(M)[0]
            assert np.isclose(want, got, rtol=1e-10, atol=1e-12)
            assert np.isclose(want, oracle.pfaffian_py(M), rtol=1e-10, atol=1e-12)


def test_reference_identity_large_log():
    # test/slog_pfaffian_test.py:76-107, seed 0, n = 40..290 step 10 (3 each here)
    rng = np.random.default_rng(seed=0)
    for n in range(40, 300, 10):
        for _ in range(3):
            M, want = oracle.tridiag_identity_case(rng, n)
            la, s = oracle.slog(M)
            assert np.isclose(want, s * math.exp(la), rtol=1e-10, atol=1e-12)


def test_reference_det_relation():
    # test/pfaffian_test.py:55-65 and slog_pfaffian_test.py:47-59, seed 1
    rng = np.random.default_rng(1)
    for n in range(2, 66, 2):
        A = rng.normal(size=(n, n))
        A = A - A.T
        la, _ = oracle.slog(A)
        sd, ld = np.linalg.slogdet(A)
        assert sd > 0
        assert abs(2 * la - ld) < 1e-9


def test_zero_and_odd():
    for n in range(2, 20, 2):
        prod, la, s, info = oracle.pf_f64(np.zeros((n, n)))
        assert prod == 0.0 and la == -math.inf and s == 0.0 and info == 1
    rng = np.random.default_rng(0)
    for n in (1, 3, 5, 11, 15, 35):
        B = rng.normal(size=(n, n))
        prod, la, s, _ = oracle.pf_f64(B - B.T)
        assert prod == 0.0 and la == -math.inf and s == 0.0


def test_transpose_flips_sign_for_n_2_mod_4():
    # Pf(A^T) = (-1)^(n/2) Pf(A): the check that catches a layout mix-up
    rng = np.random.default_rng(2)
    for n in (6, 8, 10, 12):
        A = oracle.skew_gaussian(rng, n)
        assert oracle.pf_f64(A.T)[0] == pytest.approx((-1) ** (n // 2) * oracle.pf_f64(A)[0], rel=1e-12)


def test_planted_sign_error_is_caught():
    # A restatement that dropped the interchange sign flip would fail the
    # identity test above: emulate it by comparing |Pf| with the signed value.
    rng = np.random.default_rng(9)
    M, want = oracle.tridiag_identity_case(rng, 16)
    prod = oracle.pf_f64(M)[0]
    assert np.isclose(prod, want, rtol=1e-10)
    flipped_case = [oracle.tridiag_identity_case(rng, 16) for _ in range(8)]
    assert any(not np.isclose(abs(oracle.pf_f64(m)[0]), w, rtol=1e-10) for m, w in flipped_case)


def test_complex_identity():
    # Pf(B T B^T) = det(B) Pf(T) for any B (transpose, not conjugate transpose)
    rng = np.random.default_rng(4)
    for n in (6, 12, 20, 40):
        t = rng.normal(size=n - 1) + 1j * rng.normal(size=n - 1)
        t[1::2] = 0
        T = np.diag(t, 1) - np.diag(t, -1)
        B = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        want = np.linalg.det(B) * np.prod(t[0::2])
        prod, la, ph, _ = oracle.pf_c128(B @ T @ B.T)
        assert prod == pytest.approx(want, rel=1e-9)
        assert abs(la - math.log(abs(want))) < 1e-9
        assert abs(ph - want / abs(want)) < 1e-9


def test_ragged_driver_matches_single_calls():
    rng = np.random.default_rng(8)
    sizes = [16, 2, 30, 7, 64, 0, 40]
    mats = [oracle.skew_gaussian(rng, n) for n in sizes]
    data = np.concatenate([m.reshape(-1) for m in mats])
    offs = np.cumsum([0] + [n * n for n in sizes[:-1]])
    la, s = oracle.slog_ragged(data, offs, sizes, threads=3)
    for i, m in enumerate(mats):
        l1, s1 = oracle.slog(m)
        assert (la[i] == l1) or (math.isinf(la[i]) and math.isinf(l1))
        assert s[i] == s1
