"""Garcia-Siewert tabulated phase functions (Haze L, Cloud C.1) in the facade
(phase_function::fillHazeGarciaSiewert / fillCloudGarciaSiewert,
src/PhaseFunction.cpp:44-103) and the oracle solving the reference's
disotest 4a-4c (Haze L, 32 streams) and 5a-5b (Cloud C.1, 48 streams, 299
moments) cases against the Garcia & Siewert benchmark values they quote
(tests/test_solver.cpp:1197-1420, :1572-1700).  Those are values of the
exact transfer solution printed to six digits; the reference test asserts
them at 2 %.  The oracle's 32-/48-stream answer agrees with every one of
them to within one unit of the last printed digit (at most 4e-6 relative),
which is what is asserted here."""
import math

import numpy as np
import pytest

import oracle_lib as ol
from conftest import printed_ulp
from disort_amd import general as G
from disort_amd.batch import OUT_FIELDS, scalars

import disort_amd.disortpp as d

# (name, moments, nstr, nmom, dtau, omega, mu0, user tau or None, {(field, level): value})
CASES = [
    ("disotest_4a", "haze", 32, 0, 1.0, 1.0, 1.0, None,
     {("flux_up", 0): 1.73223e-01, ("flux_down", 1): 1.81264, ("flux_direct_beam", 1): 1.15573}),
    ("disotest_4b", "haze", 32, 0, 1.0, 0.9, 1.0, None,
     {("flux_up", 0): 1.23665e-01, ("flux_down", 1): 1.51554}),
    ("disotest_4c", "haze", 32, 0, 1.0, 0.9, 0.5, None,
     {("flux_up", 0): 2.25487e-01, ("flux_down", 1): 8.03294e-01, ("flux_direct_beam", 1): 2.12584e-01}),
    ("disotest_5a", "cloud", 48, 299, 64.0, 1.0, 1.0, [0.0, 32.0, 64.0],
     {("flux_up", 0): 2.66174, ("flux_down", 1): 2.24768, ("flux_up", 1): 1.76783, ("flux_down", 2): 4.79851e-01}),
    ("disotest_5b", "cloud", 48, 299, 64.0, 0.9, 1.0, [3.2, 12.8, 48.0],
     {("flux_down", 0): 1.74767, ("flux_up", 0): 2.70485e-01, ("flux_down", 1): 2.33975e-01,
      ("flux_up", 1): 3.74252e-02, ("flux_down", 2): 6.38345e-05, ("flux_up", 2): 1.02904e-05}),
]


def facade_moments(kind, nstr, nmom):
    c = d.DisortConfig(1, nstr, nmom)
    c.allocate()
    (c.set_haze_garcia_siewert if kind == "haze" else c.set_cloud_garcia_siewert)()
    return np.array(c.phase_function_moments)


def case_batch(case, n=1):
    name, kind, nstr, nmom, dtau, w, mu0, utau, _ = case
    pm = facade_moments(kind, nstr, nmom)
    flags = G.G_USER_TAU if utau is not None else 0
    rep = lambda a: np.repeat(np.asarray(a, np.float64)[None], n, axis=0)  # noqa: E731
    return G.GeneralBatch(flags=flags, nstr=nstr, dtau=rep([dtau]), ssa=rep([w]), pmom=rep(pm),
                          scal=np.repeat(scalars(1, beam_flux=math.pi, beam_mu=mu0), n, axis=0),
                          tau_user=None if utau is None else rep(utau), phi_user=[0.0], accuracy=1e-4)


def test_fills_are_the_published_expansions():
    haze = facade_moments("haze", 16, 100)[0]
    assert haze[0] == 1.0
    assert haze[1] == 2.41260 / 3 and haze[2] == 3.23047 / 5 and haze[82] == 0.00001 / 165
    assert (haze[83:] == 0.0).all()
    cloud = facade_moments("cloud", 16, 300)[0]
    assert cloud[1] == 2.544 / 3 and cloud[298] == 0.001 / 597
    assert (cloud[299:] == 0.0).all()
    # The Haze-L asymmetry parameter g = chi_1 = 0.8042; moments decrease.
    assert haze[1] == pytest.approx(0.8042, abs=1e-4)
    assert np.all(np.diff(haze[1:40]) < 0)
    # One layer only (lc >= 0) leaves the others untouched.
    c = d.DisortConfig(2, 8, 0)
    c.allocate()
    c.set_isotropic()
    c.set_haze_garcia_siewert(1)
    pm = np.array(c.phase_function_moments)
    assert pm[0, 1] == 0.0 and pm[1, 1] == 2.41260 / 3


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_oracle_garcia_siewert_benchmarks(case):
    gb = case_batch(case)
    flux, _, _, _ = ol.general_one(gb, 0)
    for (field, lev), v in case[-1].items():
        got = flux[OUT_FIELDS.index(field), lev]
        assert abs(got - v) <= printed_ulp(v), (case[0], field, lev, got, v)
