"""Pins the oracle's numerics: the reference's own compiled Planck, Legendre
and band-LU routines (golden vectors and, where oracle/_ref is built, a
live comparison), closed forms and textbook routines."""
import ctypes as C

import numpy as np
import pytest
from conftest import load_golden

import oracle_lib as ol


@pytest.fixture(scope="module")
def gold():
    return load_golden("numerics_ref.json")


def test_planck_matches_reference_bits(gold):
    # src/Planck.cpp compiled from the reference tree (tests/golden/make_fixtures.py)
    for lo, hi, t, b, db in gold["planck"]:
        assert ol.planck(lo, hi, t) == b, (lo, hi, t)
        assert ol.planck_deriv(lo, hi, t) == pytest.approx(db, rel=1e-14, abs=0.0), (lo, hi, t)


def test_planck_live_reference_grid():
    ref = ol.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.Generator(np.random.Philox(5))
    for _ in range(300):
        lo = rng.uniform(0, 5000)
        hi = lo + rng.choice([0.0, 0.5, 1.0, 30.0, 2000.0])
        t = rng.uniform(1.0, 400.0)
        out = C.c_double()
        assert ref.ref_planck2(lo, hi, t, C.byref(out)) == 0
        assert ol.planck(lo, hi, t) == out.value


def test_planck_closed_forms():
    # Full band integrates to sigma T^4 / pi (power + exponential series branches).
    sigma = 5.67032e-8
    for t in (200.0, 300.0):
        assert ol.planck(0.0, 1e6, t) == pytest.approx(sigma * t**4 / np.pi, rel=1e-5)
    # Narrow band (Simpson branch) against scipy quadrature of the Planck shape.
    from scipy.integrate import quad

    c2 = 1.438786
    for lo, t in ((100.0, 200.0), (1000.0, 250.0), (5000.0, 300.0)):
        f = lambda v: v**3 / np.expm1(v)
        val = quad(f, c2 * lo / t, c2 * (lo + 1) / t, epsabs=0, epsrel=1e-13)[0]
        expect = sigma / np.pi * t**4 * 15 / np.pi**4 * val
        assert ol.planck(lo, lo + 1.0, t) == pytest.approx(expect, rel=2e-6)
    # Derivative against a central difference (tests/test_jacobian.cpp:186-204).
    for lo, hi, t in ((500.0, 1000.0, 250.0), (100.0, 101.0, 230.0)):
        h = 1e-3
        fd = (ol.planck(lo, hi, t + h) - ol.planck(lo, hi, t - h)) / (2 * h)
        assert ol.planck_deriv(lo, hi, t) == pytest.approx(fd, rel=1e-5)


def test_planck_validation():
    with pytest.raises(ol.OracleError):
        ol.planck(-1.0, 2.0, 200.0)
    with pytest.raises(ol.OracleError):
        ol.planck(3.0, 2.0, 200.0)
    assert ol.planck(100.0, 101.0, 5e-5) == 0.0


def test_legendre_matches_reference(gold):
    for rec in gold["legendre"]:
        mu = np.array(rec["mu"])
        m, twonm1 = rec["m"], rec["twonm1"]
        ylm = np.zeros((twonm1 + 2, mu.size))
        for mm in range(0, m + 1):  # the recurrence in m needs rows m-1
            ol.legendre_flat(mu, mm, twonm1, ylm)
        np.testing.assert_array_equal(ylm[m : twonm1 + 1], np.array(rec["ylm"])[m : twonm1 + 1])


def test_legendre_closed_forms():
    mu = np.linspace(-1, 1, 9)
    y = ol.legendre_flat(mu, 0, 4)
    np.testing.assert_allclose(y[0], 1.0)
    np.testing.assert_allclose(y[1], mu)
    np.testing.assert_allclose(y[2], 0.5 * (3 * mu**2 - 1), atol=1e-15)
    from numpy.polynomial import legendre as npl

    for l in range(5):
        np.testing.assert_allclose(y[l], npl.legval(mu, [0] * l + [1]), atol=1e-14)


@pytest.mark.parametrize("nn", [2, 4, 8, 16, 32])
def test_gauss_quadrature_vs_numpy(nn):
    cmu, cwt = ol.gauss(nn)
    x, w = np.polynomial.legendre.leggauss(nn)
    order = np.argsort(-x)
    np.testing.assert_allclose(cmu[:nn], 0.5 * (1 + x[order]), rtol=0, atol=2e-15)
    np.testing.assert_allclose(cwt[:nn], 0.5 * w[order], rtol=5e-13)
    np.testing.assert_array_equal(cmu[nn:], -cmu[:nn])
    assert abs(cwt[:nn].sum() - 1.0) < 1e-13  # (tests/test_solver.cpp:156-171)


def _band_dense(abd, n, ml):
    a = np.zeros((n, n))
    for j in range(n):
        for i in range(max(0, j - ml), min(n, j + ml + 1)):
            a[i, j] = abd[j, 2 * ml + i - j]
    return a


def test_band_lu_matches_reference(gold):
    lib = ol.lib()
    for rec in gold["band"]:
        n, ml = rec["n"], rec["ml"]
        abd = np.array(rec["abd"])
        lda = 3 * ml + 1
        a = _band_dense(abd, n, ml)
        ipvt = np.zeros(n, np.int32)
        work = abd.copy()
        lib.oracle_band_factor(work.ctypes.data, lda, n, ml, ml, ipvt.ctypes.data)
        np.testing.assert_array_equal(ipvt, rec["ipvt"])  # same pivot sequence
        x = np.array(rec["b"])
        lib.oracle_band_solve(work.ctypes.data, lda, n, ml, ml, ipvt.ctypes.data, x.ctypes.data)
        np.testing.assert_array_equal(x, rec["x"])
        np.testing.assert_allclose(x, np.linalg.solve(a, rec["b"]), rtol=1e-9, atol=1e-12)


def test_band_lu_needs_pivoting():
    # A zero on the diagonal forces a row exchange; brute-force dense solve pins it.
    n, ml = 6, 2
    lda = 3 * ml + 1
    rng = np.random.Generator(np.random.Philox(3))
    abd = np.zeros((n, lda))
    for j in range(n):
        for i in range(max(0, j - ml), min(n, j + ml + 1)):
            abd[j, 2 * ml + i - j] = rng.normal()
    abd[0, 2 * ml] = 0.0
    a = _band_dense(abd, n, ml)
    b = rng.normal(size=n)
    ipvt = np.zeros(n, np.int32)
    lib = ol.lib()
    lib.oracle_band_factor(abd.ctypes.data, lda, n, ml, ml, ipvt.ctypes.data)
    assert ipvt[0] != 1
    x = b.copy()
    lib.oracle_band_solve(abd.ctypes.data, lda, n, ml, ml, ipvt.ctypes.data, x.ctypes.data)
    np.testing.assert_allclose(x, np.linalg.solve(a, b), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("n", [2, 5, 8, 16])
def test_jacobi_eigen_vs_numpy(n):
    rng = np.random.Generator(np.random.Philox(n))
    m = rng.normal(size=(n, n))
    a = m @ m.T + n * np.eye(n)
    a[0, 1] = a[1, 0] = a[0, 0] * (1 + 1e-12)  # near-degenerate pair
    ev = np.zeros(n)
    vec = np.zeros((n, n), order="F")
    af = np.asfortranarray(a)
    ol.lib().oracle_sym_eigen(n, af.ctypes.data, ev.ctypes.data, vec.ctypes.data)
    w = np.linalg.eigvalsh(a)
    np.testing.assert_allclose(ev, w, rtol=1e-13, atol=1e-13 * abs(w).max())
    np.testing.assert_allclose(vec.T @ vec, np.eye(n), atol=1e-13)
    np.testing.assert_allclose(a @ vec, vec * ev, atol=1e-12 * abs(w).max())


def test_cholesky_rejects_indefinite():
    a = np.asfortranarray(np.array([[1.0, 2.0], [2.0, 1.0]]))
    l = np.zeros((2, 2), order="F")
    assert ol.lib().oracle_cholesky(2, a.ctypes.data, l.ctypes.data) == 1
    spd = np.asfortranarray(np.array([[4.0, 2.0], [2.0, 3.0]]))
    assert ol.lib().oracle_cholesky(2, spd.ctypes.data, l.ctypes.data) == 0
    np.testing.assert_allclose(np.tril(l) @ np.tril(l).T, spd, atol=1e-15)


def test_chapman_plane_parallel_limit():
    # With a huge planet radius the slant path factor tends to 1/mu0.
    zd = np.array([100.0, 80.0, 60.0, 40.0])
    dtau = np.array([0.1, 0.2, 0.3])
    for mu0 in (0.3, 0.7, 1.0):
        s = ol.lib().oracle_chapman(3, 0.0, zd.ctypes.data, dtau.ctypes.data, 3, 6.371e9, mu0)
        assert s == pytest.approx(dtau.sum() / mu0, rel=1e-6)
