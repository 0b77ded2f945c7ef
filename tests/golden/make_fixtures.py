"""Writes the golden fixtures under tests/golden/ (run in the build container).

* reference_cases.json -- expected fluxes transcribed from the reference's
  own tests (file:line given per case): DisORT++ self-regression values
  (13 significant digits) and C-DISORT disotest tables (7 digits), for the
  configurations the flux-only solver can express (outputs at layer
  boundaries, Lambertian surface).  The Cloud C.1 moment table is the
  published Garcia & Siewert (1985) coefficient list the reference's tests
  use (tests/test_solver.cpp:110-130).
* numerics_ref.json -- input/output vectors of the reference's own
  src/Planck.cpp, src/Legendre.cpp and src/BandSolver.hpp, compiled here
  into oracle/_ref/libref_parts.so by oracle/Makefile.

Usage: python tests/golden/make_fixtures.py   (needs oracle/_ref built and
the reference tree mounted at /root/reference)
"""
from __future__ import annotations

import ctypes as C
import json
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_TREE = "/root/reference"


def cloud_moments():
    """Garcia-Siewert Cloud C.1 coefficients a_k, k=1..298, as listed in
    tests/test_solver.cpp:110-130 of the reference."""
    txt = open(os.path.join(REF_TREE, "tests", "test_solver.cpp")).read()
    i = txt.index("static const double cldmom[299]")
    body = txt[txt.index("{", i) + 1 : txt.index("}", i)]
    vals = [float(x) for x in re.findall(r"[0-9]+\.[0-9]+", body)]
    assert len(vals) >= 298, len(vals)
    return vals[:298]


def hg(g, n):
    return [g**k for k in range(n)]


def reference_cases():
    pi = np.pi
    cases = []

    def add(name, src, tol, L, nstr, P, dtau, ssa, pmom_rows, sc, expect, flags=0, temper=None):
        cases.append(
            dict(
                name=name,
                source=src,
                rtol=tol,
                nlyr=L,
                nstr=nstr,
                pstride=P,
                dtau=list(dtau),
                ssa=list(ssa),
                pmom=[list(r) for r in pmom_rows],
                scal=sc,
                flags=flags,
                temper=temper,
                expect=expect,
            )
        )

    # --- DisORT++ self-regression (13 significant digits) -------------------
    add("deltam_plus_A_D", "tests/test_solver.cpp:5238-5249,5386-5395", 1e-11, 1, 16, 18,
        [1.0], [0.9], [hg(0.85, 18)], dict(beam_flux=1.0, beam_mu=0.5, albedo=0.1),
        {"flux_up": {"0": 8.063437290338e-02, "1": 3.446011784953e-02},
         "flux_direct_beam": {"0": 5.0e-01, "1": 6.766764161831e-02},
         "flux_down": {"1": 2.769335368770e-01}}, flags=4)
    add("deltam_plus_B", "tests/test_solver.cpp:5264-5304", 1e-11, 6, 16, 18,
        [0.1, 0.2, 0.5, 1.0, 2.0, 5.0], [0.9] * 6, [hg(0.9, 18)] * 6,
        dict(beam_flux=1.0, beam_mu=0.5, albedo=0.1),
        {"flux_up": {"0": 7.706455092675e-02, "6": 4.821314169115e-03},
         "flux_direct_beam": {"0": 5.0e-01},
         "flux_down": {"6": 4.821313033092e-02}}, flags=4)
    cld = cloud_moments()
    pm = [1.0] + [cld[k - 1] / (2 * k + 1) for k in range(1, 299)]
    add("deltam_plus_E_cloud", "tests/test_solver.cpp:5400-5434 (omega=1: conservative, 1e-8)", 1e-8,
        1, 16, 299, [5.0], [1.0], [pm], dict(beam_flux=1.0, beam_mu=0.5, albedo=0.0),
        {"flux_up": {"0": 2.317019156027e-01}, "flux_direct_beam": {"0": 5.0e-01},
         "flux_down": {"1": 2.682753843991e-01}}, flags=4)

    # --- C-DISORT tables (disotest) at layer boundaries ---------------------
    iso16 = [1.0] + [0.0] * 16
    add("disotest_6b_beer", "tests/test_solver.cpp:2093-2142", 1e-6, 1, 16, 17, [1.0], [0.0], [iso16],
        dict(beam_flux=200.0, beam_mu=0.5),
        {"flux_direct_beam": {"0": 1.0e2, "1": 1.353353e1}, "flux_down": {"0": 0.0, "1": 0.0},
         "flux_up": {"0": 0.0, "1": 0.0}})
    add("disotest_6c_lambert", "tests/test_solver.cpp:2144-2198", 1e-6, 1, 16, 17, [1.0], [0.0], [iso16],
        dict(beam_flux=200.0, beam_mu=0.5, albedo=0.5),
        {"flux_direct_beam": {"0": 1.0e2, "1": 1.353353e1}, "flux_up": {"0": 1.484504, "1": 6.766764}})
    add("disotest_6a_transparent", "tests/test_solver.cpp:2046-2091", 1e-9, 1, 16, 17, [0.0], [0.0], [iso16],
        dict(beam_flux=200.0, beam_mu=0.5),
        {"flux_direct_beam": {"0": 100.0, "1": 100.0}, "flux_down": {"0": 0.0, "1": 0.0},
         "flux_up": {"0": 0.0, "1": 0.0}})
    iso8 = [1.0] + [0.0] * 8
    dt9 = [float(k + 1) for k in range(6)]
    w9 = [0.65 + 0.05 * k for k in range(6)]
    add("disotest_9a", "tests/test_solver.cpp:1876-1952 (tau 0, 6, 21 = levels 0, 3, 6)", 2e-6, 6, 8, 9,
        dt9, w9, [iso8] * 6, dict(iso_top=1.0 / pi),
        {"flux_down": {"0": 1.0, "3": 6.714448e-03, "6": 6.169678e-07},
         "flux_up": {"0": 2.279734e-01, "3": 2.192911e-03}})
    dgis = [1.0, 2.00916 / 3.0, 1.56339 / 5.0, 0.67407 / 7.0, 0.22215 / 9.0, 0.04725 / 11.0,
            0.00671 / 13.0, 0.00068 / 15.0, 0.00005 / 17.0]
    add("disotest_9b_dgis", "tests/test_solver.cpp:1954-2044 (levels 0, 3, 6)", 2e-6, 6, 8, 9,
        dt9, w9, [dgis] * 6, dict(iso_top=1.0 / pi),
        {"flux_down": {"0": 1.0, "3": 2.764751e-02, "6": 7.418531e-05},
         "flux_up": {"0": 1.000789e-01, "3": 4.160164e-03}})
    ray = [1.0, 0.0, 0.1] + [0.0] * 14
    add("multilayer_1_rayleigh", "tests/test_solver.cpp:1383-1441", 2e-6, 2, 16, 17, [0.1, 0.2], [0.5, 0.8],
        [ray, ray], dict(beam_flux=1.0, beam_mu=0.5),
        {"flux_direct_beam": {"0": 5.0e-01, "1": 4.093654e-01, "2": 2.744058e-01},
         "flux_down": {"1": 2.327019e-02, "2": 6.715049e-02},
         "flux_up": {"0": 6.593572e-02, "1": 5.456810e-02}})
    add("multilayer_2_hg_conservative", "tests/test_solver.cpp:1443-1500 (omega=1)", 2e-6, 2, 16, 33,
        [0.5, 0.5], [1.0, 1.0], [hg(0.75, 33)] * 2, dict(beam_flux=pi, beam_mu=1.0),
        {"flux_direct_beam": {"0": 3.141593, "1": 1.905472, "2": 1.155727},
         "flux_down": {"1": 1.149116, "2": 1.738491},
         "flux_up": {"0": 2.473744e-01, "1": 1.603704e-01}})
    add("multilayer_3_isotropic", "tests/test_solver.cpp:1502-1576", 2e-6, 3, 16, 17, [0.5, 1.0, 0.5],
        [0.5, 0.9, 0.5], [iso16] * 3, dict(beam_flux=1.0, beam_mu=0.5, albedo=0.1),
        {"flux_direct_beam": {"0": 5.0e-01, "1": 1.839397e-01, "2": 2.489353e-02},
         "flux_down": {"1": 6.416506e-02, "2": 8.605320e-02},
         "flux_up": {"0": 1.133613e-01, "1": 1.031752e-01, "2": 1.536433e-02}})
    # Thermal T1-T4 (C-DISORT, intensity_corr off), wavenumbers 0..50000.
    therm = dict(wn_lo=0.0, wn_hi=50000.0, emis_top=1.0)
    add("thermal_T1", "tests/test_solver.cpp:3216-3235", 2e-6, 1, 16, 17, [1.0], [0.0], [iso16],
        dict(temp_bot=300.0, temp_top=0.0, **therm),
        {"flux_up": {"0": 2.729989e2, "1": 4.592959e2}, "flux_down": {"1": 2.571187e2}},
        flags=1, temper=[200.0, 300.0])
    add("thermal_T2", "tests/test_solver.cpp:3237-3250", 2e-6, 1, 16, 17, [1.0], [0.0], [iso16],
        dict(temp_bot=300.0, temp_top=0.0, albedo=0.5, **therm),
        {"flux_up": {"0": 2.508219e2, "1": 3.582073e2}, "flux_down": {"1": 2.571187e2}},
        flags=1, temper=[200.0, 300.0])
    add("thermal_T3", "tests/test_solver.cpp:3252-3271", 2e-6, 1, 16, 17, [1.0], [0.0], [iso16],
        dict(temp_bot=300.0, temp_top=0.0, beam_flux=200.0, beam_mu=0.5, **therm),
        {"flux_direct_beam": {"0": 1.0e2}, "flux_up": {"0": 1.007611e2, "1": 4.592959e2}},
        flags=1, temper=[0.0, 0.0])
    add("thermal_T4", "tests/test_solver.cpp:3273-3298", 2e-6, 1, 16, 17, [1.0], [0.0], [iso16],
        dict(temp_bot=300.0, temp_top=250.0, **therm),
        {"flux_down": {"0": 2.214969e2, "1": 3.416943e2}, "flux_up": {"0": 3.390985e2, "1": 4.592959e2}},
        flags=1, temper=[250.0, 300.0])
    return cases


def numerics_ref():
    lib = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_parts.so"))
    D, PD = C.c_double, C.POINTER(C.c_double)
    lib.ref_planck2.argtypes = [D, D, D, PD]
    lib.ref_planck_deriv2.argtypes = [D, D, D, PD]
    lib.ref_legendre_flat.argtypes = [C.c_int] * 4 + [C.c_void_p, C.c_void_p]
    lib.ref_band_factor.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
    lib.ref_band_solve.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    rng = np.random.Generator(np.random.Philox(99))
    planck = []
    pts = []
    for b in (0, 1, 17, 250, 899, 2500, 5000, 9999):  # C2 bands [100+b, 101+b]
        for t in (150.0, 200.0, 236.5, 300.0, 350.0):
            pts.append((100.0 + b, 101.0 + b, t))
    pts += [(0.0, 50000.0, 300.0), (0.0, 50000.0, 200.0), (500.0, 1000.0, 250.0), (800.0, 1200.0, 310.0),
            (10.0, 11.0, 300.0), (2000.0, 2000.0, 280.0), (1.0, 5.0, 250.0), (3000.0, 9000.0, 260.0),
            (100.0, 101.0, 5e-5), (100.0, 101.0, 0.0)]
    for lo, hi, t in pts:
        b, db = D(), D()
        assert lib.ref_planck2(lo, hi, t, C.byref(b)) == 0
        assert lib.ref_planck_deriv2(lo, hi, t, C.byref(db)) == 0
        planck.append([lo, hi, t, b.value, db.value])
    legendre = []
    for nmu, mmax, twonm1 in ((5, 3, 15), (4, 7, 31)):
        mu = np.sort(rng.uniform(-1, 1, nmu))
        ylm = np.zeros((twonm1 + 2, nmu))
        for m in range(0, mmax + 1):
            lib.ref_legendre_flat(nmu, m, nmu - 1, twonm1, mu.ctypes.data, ylm.ctypes.data)
            legendre.append(dict(mu=mu.tolist(), m=m, twonm1=twonm1, ylm=ylm[: twonm1 + 1].tolist()))
    band = []
    for n, ml in ((12, 3), (40, 5), (48, 23)):
        lda = 3 * ml + 1
        abd = np.zeros((n, lda))  # column-major: abd[col, row]
        for j in range(n):
            for i in range(max(0, j - ml), min(n, j + ml + 1)):
                abd[j, 2 * ml + i - j] = rng.normal()
        b = rng.normal(size=n)
        a0 = abd.copy()
        ipvt = np.zeros(n, np.int32)
        lib.ref_band_factor(abd.ctypes.data, lda, n, ml, ml, ipvt.ctypes.data)
        x = b.copy()
        lib.ref_band_solve(abd.ctypes.data, lda, n, ml, ml, ipvt.ctypes.data, x.ctypes.data)
        band.append(dict(n=n, ml=ml, abd=a0.tolist(), b=b.tolist(), ipvt=ipvt.tolist(), x=x.tolist()))
    return dict(planck=planck, legendre=legendre, band=band)


if __name__ == "__main__":
    with open(os.path.join(HERE, "reference_cases.json"), "w") as f:
        json.dump(reference_cases(), f, indent=1)
    with open(os.path.join(HERE, "numerics_ref.json"), "w") as f:
        json.dump(numerics_ref(), f)
    print("wrote reference_cases.json, numerics_ref.json")
