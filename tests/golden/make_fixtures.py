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


def _hg_rows(g, P):
    """HG moments as the reference fills them (running product g^k)."""
    row, gk = [], 1.0
    for _ in range(P):
        row.append(gk)
        gk *= g
    return row


def _iso_rows(P):
    return [1.0] + [0.0] * (P - 1)


def _ray_rows(P):
    return [1.0 if k == 0 else (0.1 if k == 2 else 0.0) for k in range(P)]


def intensity_cases():
    """General-solver cases transcribed from the reference's tests
    (tests/test_solver.cpp, line ranges per case).  Each expectation is
    (quantity, indices, value, tolerance) with the tolerance the reference
    test uses: relative ('rtol') or absolute ('atol')."""
    pi = np.pi
    cases = []

    def add(name, src, nlyr, nstr, nmom, dtau, ssa, pmom, flags, scal, expect, accuracy=0.0, tau_user=None,
            mu_user=None, phi_user=None, temper=None, alt=None, rpv=None, mu_phase=None, phase=None):
        P = max(nmom, nstr) + 1
        assert all(len(r) == P for r in pmom), name
        case = dict(name=name, source=src, nlyr=nlyr, nstr=nstr, nmom=nmom, pstride=P, dtau=list(dtau),
                    ssa=list(ssa), pmom=[list(r) for r in pmom], flags=flags, scal=scal,
                    accuracy=accuracy, tau_user=tau_user, mu_user=mu_user, phi_user=phi_user,
                    temper=temper, alt=alt, rpv=rpv, expect=expect)
        if mu_phase is not None:
            case.update(mu_phase=list(mu_phase), phase=[list(r) for r in phase])
        cases.append(case)

    def I(iu, lu, j, v, rtol=None, atol=None):
        return dict(q="intensity", iu=iu, lu=lu, j=j, v=v, rtol=rtol, atol=atol)

    def F(field, lu, v, rtol=None, atol=None):
        return dict(q=field, lu=lu, v=v, rtol=rtol, atol=atol)

    umu6 = [-1.0, -0.5, -0.1, 0.1, 0.5, 1.0]

    def six(top_up, bot_dn, rtol, zero=True, top_dn=None):
        e = []
        for iu, v in zip((3, 4, 5), top_up):
            e.append(I(iu, 0, 0, v, rtol=rtol))
        for iu, v in zip((0, 1, 2), bot_dn):
            e.append(I(iu, 1, 0, v, rtol=rtol))
        if zero:
            for iu in (0, 1, 2):
                e.append(I(iu, 0, 0, 0.0 if top_dn is None else top_dn, atol=1e-10 if top_dn is None else 1e-5))
            if top_dn is None:
                for iu in (3, 4, 5):
                    e.append(I(iu, 1, 0, 0.0, atol=1e-10))
        return e

    # --- C-DISORT intensity tables, no correction (IntensityValidation*) ---
    iso = dict(beam_mu=0.1, albedo=0.0)
    for name, src, tau, w, fb, itop, top_up, bot_dn, rtol, zero in [
        ("iv_1a", "tests/test_solver.cpp:2613-2670", 0.03125, 0.2, pi / 0.1, 0.0,
         [1.17770662e-01, 2.64170381e-02, 1.34041302e-02], [1.33826267e-02, 2.63323514e-02, 1.15897886e-01], 5e-3, True),
        ("iv_1b", "tests/test_solver.cpp:2672-2726", 0.03125, 1.0, pi / 0.1, 0.0,
         [6.22883782e-01, 1.39762939e-01, 7.09191639e-02], [7.08109342e-02, 1.39336700e-01, 6.13457861e-01], 5e-3, True),
        ("iv_1d", "tests/test_solver.cpp:2848-2868", 32.0, 0.2, pi / 0.1, 0.0,
         [2.62972304e-01, 9.06966927e-02, 5.02852820e-02], [], 5e-3, False),
        ("iv_1e", "tests/test_solver.cpp:2870-2884", 32.0, 1.0, pi / 0.1, 0.0,
         [1.93320904e+00, 1.02731911e+00, 7.97198966e-01], [2.71316407e-02, 1.87804598e-02, 1.16385133e-02], 5e-3, False),
    ]:
        add(name, src, 1, 16, 0, [tau], [w], [_iso_rows(17)], ["user_mu"], dict(iso, beam_flux=fb, iso_top=itop),
            six(top_up, bot_dn, rtol, zero), accuracy=1e-4, mu_user=umu6, phi_user=[0.0])
    add("iv_1c", "tests/test_solver.cpp:2826-2846", 1, 16, 0, [0.03125], [0.99], [_iso_rows(17)], ["user_mu"],
        dict(iso, beam_flux=0.0, iso_top=1.0),
        six([1.33177353e-01, 2.99878840e-02, 1.52233390e-02], [9.84446844e-01, 9.69362629e-01, 8.63945617e-01],
            1e-3, True, top_dn=1.0) + [I(iu, 1, 0, 0.0, atol=1e-10) for iu in (3, 4, 5)],
        accuracy=1e-4, mu_user=umu6, phi_user=[0.0])
    add("iv_1f", "tests/test_solver.cpp:2886-2903", 1, 16, 0, [32.0], [0.99], [_iso_rows(17)], ["user_mu"],
        dict(iso, beam_flux=0.0, iso_top=1.0),
        six([8.77510240e-01, 8.15135758e-01, 7.52714753e-01], [1.86840458e-03, 1.26492283e-03, 7.79280391e-04],
            5e-3, True, top_dn=1.0),
        accuracy=1e-4, mu_user=umu6, phi_user=[0.0])
    add("iv_2a", "tests/test_solver.cpp:2728-2789", 1, 16, 0, [0.2], [0.5], [_ray_rows(17)], ["user_mu"],
        dict(beam_flux=1.0, beam_mu=0.08),
        six([2.82015706e-02, 7.33147775e-03, 2.43710005e-03], [2.27937786e-03, 6.84949135e-03, 1.46808276e-02], 5e-3),
        accuracy=1e-4, mu_user=umu6, phi_user=[0.0])
    umu_r = [-0.981986, -0.538263, -0.018014, 0.018014, 0.538263, 0.981986]
    for name, src, tau, w, top_up, bot_dn, zero in [
        ("iv_2a_disotest", "tests/test_solver.cpp:2933-2951", 0.2, 0.5,
         [1.61795654e-01, 2.11501379e-02, 7.86712687e-03], [7.71896562e-03, 2.00777929e-02, 2.57684687e-02], True),
        ("iv_2b", "tests/test_solver.cpp:2953-2965", 0.2, 1.0,
         [3.47677916e-01, 4.87120142e-02, 1.89386842e-02], [1.86027204e-02, 4.64061388e-02, 6.77603294e-02], False),
        ("iv_2c", "tests/test_solver.cpp:2967-2980", 5.0, 0.5,
         [1.62565965e-01, 2.45785751e-02, 1.01498340e-02], [1.70004351e-04, 3.97168092e-05, 1.32472007e-05], False),
        ("iv_2d", "tests/test_solver.cpp:2982-2994", 5.0, 1.0,
         [3.64009637e-01, 8.26992712e-02, 4.92369811e-02], [1.05949663e-02, 7.69337480e-03, 3.79276359e-03], False),
    ]:
        add(name, src, 1, 16, 0, [tau], [w], [_ray_rows(17)], ["user_mu"], dict(beam_flux=pi, beam_mu=0.080442),
            six(top_up, bot_dn, 5e-3, zero), accuracy=1e-4, mu_user=umu_r, phi_user=[0.0])
    for name, src, tau, top_up, bot_dn in [
        ("iv_3a", "tests/test_solver.cpp:3024-3038", 1.0,
         [1.51211231e-01, 1.01866361e-01, 3.66538569e-02], [2.67923725e+00, 2.68485407e-01, 2.13997993e-01]),
        ("iv_3b", "tests/test_solver.cpp:3040-3055", 8.0,
         [3.79792564e-01, 5.20401891e-01, 4.89946294e-01], [6.66732911e-01, 4.22352708e-01, 2.36362270e-01]),
    ]:
        add(name, src, 1, 16, 0, [tau], [1.0], [_hg_rows(0.75, 17)], ["user_mu"], dict(beam_flux=pi, beam_mu=1.0),
            six(top_up, bot_dn, 1e-2, False), accuracy=1e-4, mu_user=umu6, phi_user=[0.0])

    # --- Nakajima-Tanaka corrected tables (IntensityCorrectionHG3a/3b) ---
    for name, src, tau, top_up, bot_dn in [
        ("ic_hg3a", "tests/test_solver.cpp:3895-3910", 1.0,
         [1.57317307e-01, 1.00094999e-01, 5.51969895e-02], [2.94697018e+00, 2.60364310e-01, 2.09966872e-01]),
        ("ic_hg3b", "tests/test_solver.cpp:3912-3927", 8.0,
         [3.85898755e-01, 5.18534731e-01, 5.11459930e-01], [6.68756301e-01, 4.22340072e-01, 2.36358326e-01]),
    ]:
        add(name, src, 1, 16, 16, [tau], [1.0], [_hg_rows(0.75, 17)],
            ["user_mu", "user_tau", "corr_buras", "corr_nakajima"], dict(beam_flux=pi, beam_mu=1.0),
            six(top_up, bot_dn, 5e-4, False), tau_user=[0.0, tau], mu_user=umu6, phi_user=[0.0])

    # --- Buras-Emde corrected tables (NewICHG3a/3b): the same HG g = 0.75
    # layer with its phase function tabulated at 100 equidistant cosines
    # (tests/test_solver.cpp:3940-3988) and intensity_corr_nakajima off ---
    ct = [-1.0 + 2.0 * it / 99 for it in range(100)]
    hg_tab = [(1.0 - 0.75 ** 2) / (1.0 + 0.75 ** 2 - 1.5 * x) ** 1.5 for x in ct]
    for name, src, tau, top_up, bot_dn in [
        ("ic_buras_hg3a", "tests/test_solver.cpp:3989-4008", 1.0,
         [1.51254859e-01, 1.01163783e-01, 3.92360939e-02], [3.06133539e+00, 2.66600162e-01, 2.13783659e-01]),
        ("ic_buras_hg3b", "tests/test_solver.cpp:4009-4026", 8.0,
         [3.79836194e-01, 5.19661316e-01, 4.92942191e-01], [6.69642948e-01, 4.22349774e-01, 2.36362060e-01]),
    ]:
        add(name, src, 1, 16, 16, [tau], [1.0], [_hg_rows(0.75, 17)],
            ["user_mu", "user_tau", "corr_buras"], dict(beam_flux=pi, beam_mu=1.0),
            six(top_up, bot_dn, 1e-3, False), tau_user=[0.0, tau], mu_user=umu6, phi_user=[0.0],
            mu_phase=ct, phase=[hg_tab])

    # --- Thermal intensities (disotest 7b, 7c, 9c, 10a) ---
    add("disotest_7b", "tests/test_solver.cpp:3415-3472", 1, 16, 0, [100.0], [0.95], [_hg_rows(0.75, 17)],
        ["thermal", "user_tau", "user_mu"], dict(beam_flux=0.0, beam_mu=0.5, wn_lo=2702.99, wn_hi=2703.01),
        [F("flux_up", 0, 1.10948697e-06, rtol=1e-2), F("flux_tau_divergence", 0, 8.23218837e-08, rtol=1e-2),
         F("flux_down", 1, 2.07786355e-05, rtol=1e-2), I(1, 0, 0, 4.65744239e-07, rtol=1e-2),
         I(0, 1, 0, 7.52310868e-06, rtol=1e-2)],
        tau_user=[0.0, 100.0], mu_user=[-1.0, 1.0], phi_user=[0.0], temper=[200.0, 300.0])
    add("disotest_7c", "tests/test_solver.cpp:3474-3544", 1, 12, 0, [1.0], [0.5], [_hg_rows(0.8, 13)],
        ["thermal", "user_tau", "user_mu"],
        dict(beam_flux=200.0, beam_mu=0.5, iso_top=100.0, temp_top=100.0, temp_bot=320.0, emis_top=1.0,
             wn_lo=0.0, wn_hi=50000.0),
        [F("flux_direct_beam", 0, 1.0e2, rtol=1e-2), F("flux_down", 0, 3.19829585e+02, rtol=1e-2),
         F("flux_up", 0, 4.29571753e+02, rtol=1e-2), F("flux_tau_divergence", 0, -8.04270275e+01, rtol=1e-2),
         F("flux_direct_beam", 1, 3.67879441e+01, rtol=1e-2), F("flux_down", 1, 3.54099239e+02, rtol=1e-2),
         F("flux_up", 1, 4.47017662e+02, rtol=1e-2), F("flux_direct_beam", 2, 1.35335283e+01, rtol=1e-2),
         F("flux_down", 2, 3.01333770e+02, rtol=1e-2), F("flux_up", 2, 5.94576146e+02, rtol=1e-2),
         I(0, 0, 0, 1.01804919e+02, rtol=1e-2), I(2, 0, 0, 1.41637169e+02, rtol=1e-2),
         I(3, 0, 0, 1.48415776e+02, rtol=1e-2), I(2, 2, 0, 1.89259465e+02, rtol=1e-2),
         I(3, 2, 0, 1.89259465e+02, rtol=1e-2), I(2, 0, 1, 1.27777845e+02, rtol=1e-2),
         I(3, 0, 1, 1.48415776e+02, rtol=1e-2)],
        tau_user=[0.0, 0.5, 1.0], mu_user=[-1.0, -0.1, 0.1, 1.0], phi_user=[0.0, 90.0], temper=[300.0, 200.0])

    def nine_c(name, src, nstr, utau, umu, phi, expect):
        add(name, src, 6, nstr, 0, [float(l + 1) for l in range(6)], [0.65 + l * 0.05 for l in range(6)],
            [_hg_rows((l + 1) / 7.0, nstr + 1) for l in range(6)], ["thermal", "user_tau", "user_mu"],
            dict(beam_flux=pi, beam_mu=0.5, iso_top=1.0, albedo=0.5, temp_bot=700.0, temp_top=550.0, emis_top=1.0,
                 wn_lo=999.0, wn_hi=1000.0),
            expect, tau_user=utau, mu_user=umu, phi_user=phi, temper=[600.0 + 10.0 * l for l in range(7)])

    nine_c("disotest_9c", "tests/test_solver.cpp:3309-3360,3546-3595", 8, [0.0, 1.05, 2.1, 6.0, 21.0],
           [-1.0, -0.2, 0.2, 1.0], [60.0, 120.0, 180.0],
           [F("flux_direct_beam", 0, 1.57079633e+00, rtol=1e-2), F("flux_down", 0, 6.09217377e+00, rtol=1e-2),
            F("flux_up", 0, 4.68413572e+00, rtol=1e-2), F("flux_direct_beam", 1, 1.92354108e-01, rtol=1e-2),
            F("flux_down", 1, 4.97279304e+00, rtol=1e-2), F("flux_up", 1, 4.24381392e+00, rtol=1e-2),
            F("flux_direct_beam", 2, 2.35549970e-02, rtol=1e-2), F("flux_down", 2, 4.46616368e+00, rtol=1e-2),
            F("flux_up", 2, 4.16941147e+00, rtol=1e-2), F("flux_down", 3, 4.22730788e+00, rtol=1e-2),
            F("flux_up", 3, 4.30666805e+00, rtol=1e-2), F("flux_down", 4, 4.73766598e+00, rtol=1e-2),
            F("flux_up", 4, 5.11524388e+00, rtol=1e-2),
            I(0, 0, 0, 1.93919914e+00, rtol=1e-2), I(2, 0, 0, 1.61854966e+00, rtol=1e-2),
            I(3, 0, 0, 1.43872527e+00, rtol=1e-2), I(0, 2, 0, 1.48510758e+00, rtol=1e-2),
            I(1, 2, 0, 1.35008578e+00, rtol=1e-2), I(2, 4, 0, 1.62823270e+00, rtol=1e-2),
            I(3, 4, 0, 1.62823270e+00, rtol=1e-2), I(2, 0, 1, 1.57894559e+00, rtol=1e-2),
            I(3, 0, 1, 1.43872527e+00, rtol=1e-2), I(0, 0, 2, 1.93919914e+00, rtol=1e-2),
            I(2, 0, 2, 1.56558604e+00, rtol=1e-2)])
    nine_c("disotest_10a", "tests/test_solver.cpp:3309-3360,3597-3637", 4, [0.0, 2.1, 21.0],
           [-0.788675129, -0.211324871, 0.211324871, 0.788675129], [60.0, 120.0],
           [F("flux_direct_beam", 0, 1.57079633e+00, rtol=1e-2), F("flux_down", 0, 6.09217377e+00, rtol=1e-2),
            F("flux_up", 0, 4.68841229e+00, rtol=1e-2), F("flux_direct_beam", 1, 2.35549970e-02, rtol=1e-2),
            F("flux_down", 1, 4.47000071e+00, rtol=1e-2), F("flux_up", 1, 4.17037311e+00, rtol=1e-2),
            F("flux_down", 2, 4.73856448e+00, rtol=1e-2), F("flux_up", 2, 5.11569312e+00, rtol=1e-2),
            I(0, 0, 0, 1.93919914e+00, rtol=1e-2), I(2, 0, 0, 1.61511864e+00, rtol=1e-2),
            I(3, 0, 0, 1.46918463e+00, rtol=1e-2), I(0, 1, 0, 1.44581337e+00, rtol=1e-2),
            I(2, 1, 0, 1.33080189e+00, rtol=1e-2), I(3, 1, 0, 1.32690167e+00, rtol=1e-2),
            I(2, 2, 0, 1.62837570e+00, rtol=1e-2), I(3, 2, 0, 1.62837570e+00, rtol=1e-2),
            I(2, 0, 1, 1.57636831e+00, rtol=1e-2), I(3, 0, 1, 1.45769864e+00, rtol=1e-2)])

    # --- RPV surface (RpvBRDF1, TransparentBrdfRpv, MultiLayerBrdfRpv) ---
    add("rpv_1", "tests/test_solver.cpp:4133-4189", 1, 16, 16, [0.3], [0.9], [_iso_rows(17)],
        ["user_tau", "user_mu", "rpv"], dict(beam_flux=1.0, beam_mu=0.6),
        [I(3, 0, 0, 9.33168983e-02, rtol=1e-3), I(4, 0, 0, 5.20589621e-02, rtol=1e-3),
         I(5, 0, 0, 4.53763077e-02, rtol=1e-3), I(0, 1, 0, 2.37650510e-02, rtol=1e-3),
         I(1, 1, 0, 4.10887262e-02, rtol=1e-3), I(2, 1, 0, 8.22310683e-02, rtol=1e-3)],
        tau_user=[0.0, 0.3], mu_user=umu6, phi_user=[0.0], rpv=[0.1, 0.9, -0.2, 0.0, 0.0, 0.0, 1.0])
    mu30 = float(np.cos(pi / 6.0))
    rpv_veg = [0.027, 0.647, -0.169, 0.0, 0.0, 0.0, 1.0]
    tr = [F("flux_direct_beam", 0, 8.66025404e-01, rtol=1e-3), F("flux_up", 0, 4.84172253e-02, rtol=1e-3)]
    for j, vals in enumerate([[1.49228151e-02, 1.24765511e-02, 1.07957108e-02, 1.56340156e-02],
                              [1.89187154e-02, 1.59930084e-02, 1.41192980e-02, 1.56340156e-02],
                              [2.50577685e-02, 2.15197688e-02, 2.00132506e-02, 1.56340156e-02]]):
        tr += [I(iu, 0, j, v, rtol=1e-3) for iu, v in enumerate(vals)]
    add("rpv_transparent", "tests/test_solver.cpp:4827-4862,4928-4958", 1, 16, 16, [1e-10], [0.0], [_iso_rows(17)],
        ["user_tau", "user_mu", "rpv"], dict(beam_flux=1.0, beam_mu=mu30), tr,
        tau_user=[0.0], mu_user=[0.1, 0.2, 0.5, 1.0], phi_user=[0.0, 90.0, 180.0], rpv=rpv_veg)
    ml = [F("flux_direct_beam", 0, 8.66025404e-01, rtol=1e-3), F("flux_up", 0, 1.94651642e-01, rtol=1e-3),
          F("flux_up", 1, 7.85885055e-02, rtol=1e-3), F("flux_down", 1, 1.51472115e-01, rtol=1e-3),
          F("flux_up", 2, 4.12358568e-02, rtol=1e-3), F("flux_down", 2, 2.99006899e-01, rtol=1e-3)]
    for (lu, j), vals in {(0, 0): [1.09772501e-01, 9.81705389e-02, 6.23665968e-02, 4.84158344e-02],
                          (2, 0): [1.51962027e-02, 1.25279793e-02, 1.03952134e-02, 1.26614890e-02],
                          (0, 2): [1.17340949e-01, 1.07569692e-01, 7.88376901e-02, 4.84158344e-02]}.items():
        ml += [I(iu, lu, j, v, rtol=1e-3) for iu, v in enumerate(vals)]
    add("rpv_multilayer", "tests/test_solver.cpp:5000-5049,5129-5162", 2, 16, 16, [0.32, 0.32], [1.0, 1.0],
        [_ray_rows(17), _hg_rows(0.7, 17)], ["user_tau", "user_mu", "rpv"], dict(beam_flux=1.0, beam_mu=mu30), ml,
        tau_user=[0.0, 0.32, 0.64], mu_user=[0.1, 0.2, 0.5, 1.0], phi_user=[0.0, 90.0, 180.0], rpv=rpv_veg)

    # --- Hapke, Cox-Munk and Ross-Li (Ambrals) surfaces: parameter blocks
    # [b0, hh, w], [u10, refractive index, do_shadow], [iso, vol, geo] ---
    hapke = [1.0, 0.06, 0.6, 0.0, 0.0, 0.0, 0.0]
    add("hapke_1", "tests/test_solver.cpp:4083-4102", 1, 16, 16, [1.0], [1.0], [_hg_rows(0.75, 17)],
        ["user_tau", "user_mu", "hapke"], dict(beam_flux=pi, beam_mu=0.5),
        six([1.03364126e+00, 4.09048726e-01, 1.29319940e-01], [1.42300629e-01, 2.55523611e+00, 1.00072067e+00],
            1e-3, False),
        tau_user=[0.0, 1.0], mu_user=umu6, phi_user=[0.0], rpv=hapke)
    add("hapke_2", "tests/test_solver.cpp:4104-4122", 1, 16, 16, [0.5], [0.5], [_iso_rows(17)],
        ["user_tau", "user_mu", "hapke"], dict(beam_flux=1.0, beam_mu=0.8),
        six([4.86338877e-02, 4.47684548e-02, 3.66203890e-02], [1.72356855e-02, 2.73309928e-02, 3.98292120e-02],
            1e-3, False),
        tau_user=[0.0, 0.5], mu_user=umu6, phi_user=[0.0], rpv=hapke)

    def transparent(name, src, kind, params, phis, flux_up, vals, zero_atol=None):
        e = [F("flux_direct_beam", 0, 8.66025404e-01, rtol=1e-3), F("flux_up", 0, flux_up, rtol=1e-3)]
        for j, row in enumerate(vals):
            for iu, v in enumerate(row):
                e.append(I(iu, 0, j, v, atol=zero_atol) if v == 0.0 else I(iu, 0, j, v, rtol=1e-3))
        add(name, src, 1, 16, 16, [1e-10], [0.0], [_iso_rows(17)], ["user_tau", "user_mu", kind],
            dict(beam_flux=1.0, beam_mu=mu30), e, tau_user=[0.0], mu_user=[0.1, 0.2, 0.5, 1.0], phi_user=phis,
            rpv=params)

    transparent("hapke_transparent", "tests/test_solver.cpp:4827-4893", "hapke", hapke, [0.0, 90.0, 180.0],
                1.80952123e-01, [[7.77476519e-02, 7.54425629e-02, 6.93954190e-02, 5.36729764e-02],
                                 [6.40460953e-02, 6.26773809e-02, 5.81122830e-02, 5.36729764e-02],
                                 [5.19251669e-02, 5.17173008e-02, 5.00653878e-02, 5.36729764e-02]])
    transparent("coxmunk_transparent", "tests/test_solver.cpp:4827-4862,4895-4926", "coxmunk",
                [12.0, 1.34, 0.0, 0.0, 0.0, 0.0, 0.0], [0.0, 45.0, 90.0], 2.26858686e-02,
                [[2.42487295e-03, 2.88842623e-03, 8.53078961e-03, 3.17964917e-02],
                 [3.12678310e-05, 9.02852476e-05, 1.49471537e-03, 3.17964917e-02],
                 [0.0, 0.0, 2.39371326e-05, 3.17964917e-02]], zero_atol=1e-5)
    transparent("rossli_transparent", "tests/test_solver.cpp:4827-4862,4960-4990", "rossli",
                [0.091, 0.02, 0.01, 0.0, 0.0, 0.0, 0.0], [0.0, 90.0, 180.0], 6.80276718e-02,
                [[2.06350335e-02, 2.22951259e-02, 2.37114733e-02, 2.31711490e-02],
                 [1.22654153e-02, 1.78864384e-02, 2.10323702e-02, 2.31711490e-02],
                 [4.53790437e-03, 1.40564310e-02, 1.94761137e-02, 2.31711490e-02]])

    def multilayer(name, src, kind, params, phis, fluxes, table):
        e = [F("flux_direct_beam", 0, 8.66025404e-01, rtol=1e-3)]
        e += [F(q, lu, v, rtol=1e-3) for (q, lu), v in fluxes.items()]
        for (lu, j), vals in table.items():
            e += [I(iu, lu, j, v, rtol=1e-3) for iu, v in enumerate(vals)]
        add(name, src, 2, 16, 16, [0.32, 0.32], [1.0, 1.0], [_ray_rows(17), _hg_rows(0.7, 17)],
            ["user_tau", "user_mu", kind], dict(beam_flux=1.0, beam_mu=mu30), e, tau_user=[0.0, 0.32, 0.64],
            mu_user=[0.1, 0.2, 0.5, 1.0], phi_user=phis, rpv=params)

    multilayer("hapke_multilayer", "tests/test_solver.cpp:5000-5088", "hapke", hapke, [0.0, 90.0, 180.0],
               {("flux_up", 0): 2.83184981e-01, ("flux_up", 1): 1.91350794e-01, ("flux_down", 1): 1.75701065e-01,
                ("flux_up", 2): 1.63014593e-01, ("flux_down", 2): 3.32252296e-01},
               {(0, 0): [1.25984467e-01, 1.19170604e-01, 9.42674171e-02, 7.79502273e-02],
                (1, 0): [1.23054866e-01, 1.05759903e-01, 7.46212352e-02, 5.09419742e-02],
                (2, 0): [7.41448782e-02, 6.96104265e-02, 6.06974324e-02, 4.66773354e-02],
                (0, 2): [1.33177919e-01, 1.26289727e-01, 1.02989027e-01, 7.79502273e-02]})
    multilayer("coxmunk_multilayer", "tests/test_solver.cpp:5000-5049,5090-5127", "coxmunk",
               [12.0, 1.34, 1.0, 0.0, 0.0, 0.0, 0.0], [0.0, 45.0, 90.0],
               {("flux_up", 0): 1.78716405e-01, ("flux_up", 1): 5.75091570e-02, ("flux_down", 1): 1.46328004e-01,
                ("flux_up", 2): 1.75230219e-02, ("flux_down", 2): 2.91229302e-01},
               {(0, 0): [1.06393677e-01, 9.40595383e-02, 5.80077466e-02, 5.36951653e-02],
                (2, 0): [1.85109472e-03, 2.29113269e-03, 5.36341643e-03, 2.53947801e-02],
                (0, 2): [9.60808047e-02, 8.62969047e-02, 5.75538280e-02, 5.36951653e-02]})
    multilayer("rossli_multilayer", "tests/test_solver.cpp:5000-5049,5164-5200", "rossli",
               [0.091, 0.02, 0.01, 0.0, 0.0, 0.0, 0.0], [0.0, 90.0, 180.0],
               {("flux_up", 0): 2.06450422e-01, ("flux_up", 1): 9.33373536e-02, ("flux_down", 1): 1.54422183e-01,
                ("flux_up", 2): 5.65249803e-02, ("flux_down", 2): 3.02497242e-01},
               {(0, 0): [1.11775573e-01, 1.00828251e-01, 6.73995934e-02, 5.31758411e-02],
                (2, 0): [1.94024875e-02, 1.93417692e-02, 1.93954026e-02, 1.86845785e-02],
                (0, 2): [1.19144501e-01, 1.09042063e-01, 8.01501157e-02, 5.31758411e-02],
                (1, 1): [6.47230847e-02, 5.12384525e-02, 3.05390598e-02, 2.26323978e-02]})

    # --- Pseudo-spherical beam through the general solver (SphericalGeometry) ---
    add("spherical_general", "tests/test_solver.cpp:4459-4538", 2, 16, 0, [0.5, 1.0], [0.9, 0.8],
        [_hg_rows(0.8, 17)] * 2, ["user_mu", "spherical"], dict(beam_flux=1.0, beam_mu=0.5, albedo=0.1, radius=6371.0),
        [F("flux_direct_beam", 0, 5.0e-01, rtol=1e-3), F("flux_up", 0, 8.80100177e-02, rtol=1e-3),
         F("flux_direct_beam", 2, 2.55423483e-02, rtol=1e-3), F("flux_down", 2, 1.99378710e-01, rtol=1e-3),
         F("flux_up", 2, 2.25146963e-02, rtol=1e-3)],
        mu_user=[-0.5, 0.1, 0.5, 1.0], phi_user=[0.0], alt=[60.0, 30.0, 0.0])
    return cases


if __name__ == "__main__":
    with open(os.path.join(HERE, "reference_cases.json"), "w") as f:
        json.dump(reference_cases(), f, indent=1)
    with open(os.path.join(HERE, "numerics_ref.json"), "w") as f:
        json.dump(numerics_ref(), f)
    with open(os.path.join(HERE, "intensity_cases.json"), "w") as f:
        json.dump(intensity_cases(), f, indent=1)
    print("wrote reference_cases.json, numerics_ref.json, intensity_cases.json")
