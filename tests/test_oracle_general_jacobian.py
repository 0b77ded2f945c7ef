"""Temperature Jacobians of the general oracle (DisortSolver::
computeThermalJacobian, src/DisortSolver.cpp:2130-2377): the Jacobians of
the m = 0 fluxes at the output levels (user optical depths included), with
a Lambertian or BRDF surface and the diffusion lower boundary.

Pins: (1) central differences of the general oracle's forward solve with
the Planck difference scaled out (the outputs are affine in the Planck
values; see test_oracle_jacobian.py), at 1e-8; (2) the reference's own
DisortSolver checks (tests/test_jacobian.cpp:19-111, :206-247) at their
tolerances; (3) the flux-only oracle's Jacobian on the same problems
(tests/test_jacobian.cpp:381-405 cross-checks the two solvers at 1e-6; the
two restatements agree to 1e-12); (4) index_from_bottom
(tests/test_jacobian.cpp:316-355); (5) a 1e-6 error planted in the
Jacobian stage fails pin (1)."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
from disort_amd import general as G
from disort_amd.batch import OUT_FIELDS, S_TEMP_BOT, scalars
from test_oracle_jacobian import JAC_FIELDS, thermal_batch

PS_JACOBIAN = 17
BRDF = {"rpv": (G.G_BRDF_RPV, [0.027, 0.647, -0.169, 0.0, 0.0, 0.0, 1.0]),
        "hapke": (G.G_BRDF_HAPKE, [1.0, 0.06, 0.6, 0, 0, 0, 0]),
        "coxmunk": (G.G_BRDF_COXMUNK, [7.0, 1.34, 1.0, 0, 0, 0, 0]),
        "rossli": (G.G_BRDF_ROSSLI, [0.12, 0.03, 0.02, 0, 0, 0, 0])}


def general_batch(tlev, tsurf, nstr, ssa, beam=0.0, flags=0, dtau=0.4, albedo=0.1, wn=(500.0, 800.0), g=None,
                  nmom=None, tau_user=None, brdf=None, temp_top=0.0, iso_top=0.0, user_mu=False, n=1):
    """The reference's buildThermalConfig (tests/test_jacobian.cpp:19-69):
    dtau 0.4, chi_1 = 0.3 ssa, albedo 0.1, 500-800 cm^-1, beam mu 0.5, one
    azimuth; g gives a Henyey-Greenstein phase function instead."""
    tlev = np.asarray(tlev, np.float64)
    L = tlev.size - 1
    P = max(nstr + 1, nmom or 0)
    pm = np.zeros((n, L, P))
    pm[..., 0] = 1.0
    if g is None:
        pm[..., 1] = 0.3 * ssa
    else:
        pm[:] = (np.asarray(g, np.float64)[:, None] ** np.arange(P)[None, :])[None]
    f = flags | 1  # thermal emission
    rpv = None
    if brdf:
        bf, par = BRDF[brdf]
        f |= bf
        rpv = np.tile(par, (n, 1))
    if tau_user is not None:
        f |= G.G_USER_TAU
    mu = None
    if user_mu:
        f |= G.G_USER_MU
        mu = np.array([-0.9, -0.4, 0.3, 0.8])
    sc = scalars(n, beam_flux=beam, beam_mu=0.5, albedo=albedo, temp_bot=tsurf, temp_top=temp_top, iso_top=iso_top,
                 wn_lo=wn[0], wn_hi=wn[1])
    dt = np.broadcast_to(np.asarray(dtau, np.float64), (L,))
    return G.GeneralBatch(flags=f, nstr=nstr, dtau=np.tile(dt, (n, 1)), ssa=np.full((n, L), ssa), pmom=pm, scal=sc,
                          temper=np.tile(tlev, (n, 1)), rpv=rpv, tau_user=tau_user, mu_user=mu, phi_user=[0.0],
                          accuracy=1e-4)


def perturbed(gb, dof, h):
    L = gb.nlyr
    temper = gb.temper.copy()
    scal = gb.scal.copy()
    if dof <= L:
        temper[0, dof] += h
    else:
        scal[0, S_TEMP_BOT] += h
    return G.GeneralBatch(flags=gb.flags, nstr=gb.nstr, dtau=gb.dtau, ssa=gb.ssa, pmom=gb.pmom, scal=scal,
                          temper=temper, rpv=gb.rpv, tau_user=gb.tau_user, mu_user=gb.mu_user,
                          phi_user=gb.phi_user, accuracy=gb.accuracy)


def fd_failures(gb, rtol, atol, linear=True):
    L = gb.nlyr
    _, jac = ol.general_jacobian_one(gb, 0)
    lo, hi = gb.scal[0, 8], gb.scal[0, 9]
    bad = []
    for dof in range(L + 2):
        t0 = gb.temper[0, dof] if dof <= L else gb.scal[0, S_TEMP_BOT]
        h = 1e-3 * max(abs(t0), 1.0)
        fp = ol.general_one(perturbed(gb, dof, h), 0)[0]
        fm = ol.general_one(perturbed(gb, dof, -h), 0)[0]
        scale = (ol.planck(lo, hi, t0 + h) - ol.planck(lo, hi, t0 - h)) / (2 * h) / ol.planck_deriv(lo, hi, t0)
        for f, name in enumerate(JAC_FIELDS):
            q = OUT_FIELDS.index(name)
            fd = (fp[q] - fm[q]) / (2 * h)
            ana = jac[f, :, dof] * (scale if linear else 1.0)
            floor = atol + 1e-13 * np.abs(fp[q]).max() / h
            err = np.abs(ana - fd) - (rtol * np.maximum(np.abs(fd), np.abs(ana)) + floor)
            if (err > 0).any():
                lu = int(np.argmax(err))
                bad.append((name, dof, lu, ana[lu], fd[lu]))
    return bad


TL5 = [210.0, 235.0, 255.0, 270.0, 280.0]
CASES = [
    ("single_layer_8", dict(tlev=[200.0, 260.0], tsurf=280.0, nstr=8, ssa=0.0), 1e-3),
    ("multilayer_profile_16", dict(tlev=[200.0 + 8.0 * i for i in range(11)], tsurf=285.0, nstr=16, ssa=0.0), 1e-3),
    ("scattering_half_16", dict(tlev=TL5, tsurf=290.0, nstr=16, ssa=0.5), 1.5e-3),
    ("scattering_high_16", dict(tlev=TL5, tsurf=290.0, nstr=16, ssa=0.9), 2e-3),
    ("thermal_plus_beam_16", dict(tlev=[200.0, 240.0, 270.0], tsurf=288.0, nstr=16, ssa=0.6, beam=1.0), 1.5e-3),
    ("user_tau_8", dict(tlev=TL5, tsurf=290.0, nstr=8, ssa=0.7, dtau=[0.3, 1.2, 0.5, 0.8],
                        tau_user=[[0.0, 0.1, 0.3, 1.0, 2.0, 2.8]], beam=0.5), None),
    ("rpv_12", dict(tlev=TL5, tsurf=295.0, nstr=12, ssa=0.4, brdf="rpv", beam=1.0, user_mu=True), None),
    ("hapke_8", dict(tlev=TL5, tsurf=300.0, nstr=8, ssa=0.3, brdf="hapke"), None),
    ("coxmunk_16", dict(tlev=TL5, tsurf=292.0, nstr=16, ssa=0.2, brdf="coxmunk",
                        tau_user=[[0.05, 0.9, 1.6]]), None),
    ("rossli_6", dict(tlev=TL5, tsurf=288.0, nstr=6, ssa=0.6, brdf="rossli", temp_top=170.0, iso_top=0.1), None),
    ("diffusion_delta_m_12", dict(tlev=TL5, tsurf=290.0, nstr=12, ssa=0.8, flags=8, g=[0.8, 0.6, 0.9, 0.7],
                                  nmom=30, dtau=[0.5, 2.0, 0.1, 1.0]), None),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_general_jacobian_matches_linearised_finite_difference(case):
    bad = fd_failures(general_batch(**case[1]), rtol=1e-8, atol=1e-14)
    assert not bad, bad[:5]


@pytest.mark.parametrize("case", [c for c in CASES if c[2]], ids=[c[0] for c in CASES if c[2]])
def test_general_jacobian_reference_finite_difference_tolerance(case):
    """tests/test_jacobian.cpp:73-111 at the reference's rtol, atol 1e-9."""
    bad = fd_failures(general_batch(**case[1]), rtol=case[2], atol=1e-9, linear=False)
    assert not bad, bad[:5]


@pytest.mark.parametrize("case", [c for c in CASES if c[2]], ids=[c[0] for c in CASES if c[2]])
def test_general_and_flux_oracle_jacobians_agree(case):
    kw = case[1]
    _, jg = ol.general_jacobian_one(general_batch(**kw), 0)
    _, jf = ol.flux_jacobian_one(thermal_batch(**kw), 0)
    np.testing.assert_allclose(jg, jf, rtol=1e-11, atol=1e-12 * np.abs(jf).max())


def test_general_jacobian_index_from_bottom_reversal():
    tlev = [210.0, 240.0, 265.0, 285.0]
    L = len(tlev) - 1
    tu = np.array([[0.0, 0.25, 0.6, 1.2]])
    _, j0 = ol.general_jacobian_one(general_batch(tlev, 295.0, 16, 0.3, tau_user=tu), 0)
    # Bottom-up: temperatures reversed, user depths measured from the bottom.
    gb = general_batch(tlev[::-1], 295.0, 16, 0.3, tau_user=1.2 - tu[:, ::-1])
    _, j1 = ol.general_jacobian_one(gb, 0, index_from_bottom=True)
    cols = list(range(L, -1, -1)) + [L + 1]
    np.testing.assert_allclose(j1, j0[:, ::-1][:, :, cols], rtol=1e-12, atol=1e-15)


def test_general_jacobian_leaves_forward_outputs_unchanged():
    gb = general_batch([220.0, 250.0, 275.0], 285.0, 8, 0.2, brdf="rpv", user_mu=True, beam=1.0)
    flux, _ = ol.general_jacobian_one(gb, 0)
    assert np.array_equal(flux, ol.general_one(gb, 0)[0])


def test_planted_general_jacobian_error_is_caught():
    l = ol.lib()
    l.oracle_set_perturbation.argtypes = [C.c_int, C.c_double]
    try:
        l.oracle_set_perturbation(PS_JACOBIAN, 1e-6)
        assert fd_failures(general_batch(**CASES[6][1]), rtol=1e-8, atol=1e-14)
    finally:
        l.oracle_set_perturbation(0, 0.0)
