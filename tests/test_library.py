"""The built libraries load and export what include/disort_amd.h declares,
and the host-only entry points behave (no GPU is touched here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "disort_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(disort_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_abi():
    names = declared_functions()
    for expected in ("disort_ctx_create", "disort_flux_solve", "disort_flux_plan_create", "disort_flux_execute",
                     "disort_quadrature", "disort_version"):
        assert expected in names


def test_library_exports_every_declared_symbol():
    from disort_amd import _capi

    lib = C.CDLL(_capi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"libdisort_amd.so lacks {missing}"
    assert _capi.version().startswith("disort_amd")


def test_quadrature_matches_gauss_legendre():
    from disort_amd import _capi

    lib = C.CDLL(_capi.LIB_PATH)
    lib.disort_quadrature.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
    for nstr in (4, 6, 8, 16, 32, 64):
        mu = np.zeros(nstr)
        wt = np.zeros(nstr)
        assert lib.disort_quadrature(nstr, mu.ctypes.data, wt.ctypes.data) == 0
        x, w = np.polynomial.legendre.leggauss(nstr // 2)
        order = np.argsort(-x)  # nodes descend in cos, as the Newton loop finds them
        np.testing.assert_allclose(mu[: nstr // 2], 0.5 * (1 + x[order]), rtol=0, atol=2e-15)
        np.testing.assert_allclose(wt[: nstr // 2], 0.5 * w[order], rtol=5e-13)
        np.testing.assert_array_equal(mu[nstr // 2 :], -mu[: nstr // 2])
        np.testing.assert_array_equal(wt[nstr // 2 :], wt[: nstr // 2])
    assert lib.disort_quadrature(5, mu.ctypes.data, wt.ctypes.data) == 1


def test_planck_band_matches_oracle():
    """The library's host Planck entry point (the kernels' arithmetic) is bit
    for bit the oracle's band integral, itself pinned to the reference's
    src/Planck.cpp; invalid arguments are rejected."""
    import sys

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    from disort_amd import _capi

    t = np.array([0.0, 1e-5, 150.0, 200.0, 255.5, 300.0, 1000.0])
    for lo, hi in ((1000.0, 1001.0), (100.0, 101.0), (10.0, 2500.0), (0.0, 50.0)):
        b = _capi.planck_band(lo, hi, t)
        assert all(b[k] == ol.planck(lo, hi, tk) for k, tk in enumerate(t)), (lo, hi)
    for bad in ((-1.0, 2.0, [200.0]), (5.0, 2.0, [200.0]), (1.0, 2.0, [-3.0])):
        with pytest.raises(_capi.DisortInvalid):
            _capi.planck_band(*bad)


def test_planck_band_deriv_matches_oracle():
    """The host dB/dT entry point agrees with the oracle's derivative (pinned
    to the reference's planckFunctionDeriv2) to rounding, and with a central
    difference of the band integral."""
    import sys

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    from disort_amd import _capi

    t = np.array([150.0, 200.0, 255.5, 300.0, 1000.0])
    for lo, hi in ((1000.0, 1001.0), (100.0, 101.0), (10.0, 2500.0)):
        d = _capi.planck_band_deriv(lo, hi, t)
        ref = np.array([ol.planck_deriv(lo, hi, tk) for tk in t])
        np.testing.assert_allclose(d, ref, rtol=1e-13)
        h = 1e-3
        fd = (_capi.planck_band(lo, hi, t + h) - _capi.planck_band(lo, hi, t - h)) / (2 * h)
        np.testing.assert_allclose(d, fd, rtol=1e-5)
    with pytest.raises(_capi.DisortInvalid):
        _capi.planck_band_deriv(1.0, 2.0, [-3.0])


def test_facade_module_loads():
    import disort_amd.disortpp as d

    for n in (4, 6, 8, 10, 12, 14, 16, 32, 64):
        assert hasattr(d, f"DisortFluxSolver{n}")
    for name in ("DisortConfig", "DisortFluxConfig", "DisortResult", "FluxResult", "DisortSolver",
                 "solve_spectral", "solve_spectral_bands", "solve_flux_spectral", "solve_flux_spectral_bands",
                 "create_flux_solver"):
        assert hasattr(d, name), name
    assert d.__version__ == "1.0"


def test_ctx_without_gpu_fails_loudly():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
    from disort_amd._capi import Context, DisortError

    with pytest.raises(DisortError):
        Context(0)
