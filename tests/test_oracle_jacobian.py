"""Temperature Jacobians of the flux-only oracle (the analytic linearisation
of DisortFluxSolver::computeThermalJacobian, src/FluxSolver.cpp:786-831 and
:1294-1497) pinned against central finite differences of the oracle's own
forward solve, on the reference's test problems (buildFluxThermalConfig,
tests/test_jacobian.cpp:117-180 and :361-376) and further ones (diffusion
lower boundary, Delta-M, top emission and isotropic sources, spherical beam).

Every output is affine in the level Planck values B(T_k) and in B(T_bottom),
so a central difference in T_k equals (d out / d B_k) times the central
difference of B(T_k) up to rounding alone.  The pin divides the analytic
column by dB/dT and multiplies by the Planck difference: that removes the
truncation error of the finite difference and lets the comparison run at
1e-8 instead of the reference's 1e-3.  The plain finite-difference check at
the reference's own tolerances is kept as well.  dB/dT itself is pinned in
test_oracle_numerics.py."""
import numpy as np
import pytest

import oracle_lib as ol
from disort_amd.batch import F_DIFFUSION_BC, F_SPHERICAL, F_THERMAL, OUT_FIELDS, scalars, uniform_batch

JAC_FIELDS = ("flux_up", "flux_down", "mean_intensity", "mean_intensity_down", "mean_intensity_up",
              "flux_tau_divergence")
FLUX_STAGES = 32
PS_JACOBIAN = 17


def thermal_batch(tlev, tsurf, nstr, ssa, beam=0.0, flags=0, dtau=0.4, g=None, nmom=None, temp_top=0.0,
                  iso_top=0.0, iso_bot=0.0, albedo=0.1, wn=(500.0, 800.0), alt=None, radius=0.0):
    """The reference's buildFluxThermalConfig (tests/test_jacobian.cpp:117-140):
    dtau 0.4, chi_1 = 0.3 ssa, albedo 0.1, 500-800 cm^-1, T_top 0, beam
    mu 0.5; g gives a Henyey-Greenstein phase function instead."""
    tlev = np.asarray(tlev, np.float64)
    L = tlev.size - 1
    P = max(nstr + 1, nmom or 0)
    pm = np.zeros((1, L, P))
    pm[..., 0] = 1.0
    if g is None:
        pm[..., 1] = 0.3 * ssa
    else:
        pm[0] = np.asarray(g, np.float64)[:, None] ** np.arange(P)[None, :]
    dt = np.broadcast_to(np.asarray(dtau, np.float64), (L,))[None].copy()
    sc = scalars(1, beam_flux=beam, beam_mu=0.5, albedo=albedo, temp_bot=tsurf, temp_top=temp_top,
                 iso_top=iso_top, iso_bot=iso_bot, wn_lo=wn[0], wn_hi=wn[1], radius=radius)
    return uniform_batch(dt, np.full((1, L), ssa), pm, nstr, sc, flags=F_THERMAL | flags, temper=tlev[None],
                         alt=None if alt is None else np.asarray(alt, np.float64)[None])


def perturbed(fb, dof, h):
    """Copy of the one-instance batch with temperature dof moved by h."""
    L = int(fb.nlyr[0])
    temper = fb.temper.reshape(1, -1).copy()
    scal = fb.scal.reshape(1, -1).copy()
    if dof <= L:
        temper[0, dof] += h
    else:
        from disort_amd.batch import S_TEMP_BOT
        scal[0, S_TEMP_BOT] += h
    return uniform_batch(fb.dtau.reshape(1, L), fb.ssa.reshape(1, L), fb.pmom.reshape(1, L, -1), int(fb.nstr[0]),
                         scal, flags=fb.flags, temper=temper,
                         alt=None if fb.alt is None else fb.alt.reshape(1, -1))


def dof_temperature(fb, dof):
    from disort_amd.batch import S_TEMP_BOT
    L = int(fb.nlyr[0])
    return fb.temper[dof] if dof <= L else fb.scal[S_TEMP_BOT]


def fd_failures(fb, rtol, atol, linear=True):
    """Compare the oracle's analytic Jacobian with central differences of its
    forward solve.  linear=True scales the Planck difference out (see the
    module docstring); False is the reference's plain check."""
    L = int(fb.nlyr[0])
    _, jac = ol.flux_jacobian_one(fb, 0)
    lo, hi = fb.scal[8], fb.scal[9]
    bad = []
    for dof in range(L + 2):
        t0 = dof_temperature(fb, dof)
        h = 1e-3 * max(abs(t0), 1.0)
        fp = ol.flux_one(perturbed(fb, dof, h), 0)
        fm = ol.flux_one(perturbed(fb, dof, -h), 0)
        dbdt = ol.planck_deriv(lo, hi, t0)
        db_fd = (ol.planck(lo, hi, t0 + h) - ol.planck(lo, hi, t0 - h)) / (2 * h)
        for f, name in enumerate(JAC_FIELDS):
            fd = (fp[OUT_FIELDS.index(name)] - fm[OUT_FIELDS.index(name)]) / (2 * h)
            ana = jac[f, :, dof] * (db_fd / dbdt if linear else 1.0)
            scale = np.maximum(np.abs(fd), np.abs(ana))
            # the finite difference carries the forward solve's rounding / (2h)
            floor = atol + 1e-13 * np.abs(fp[OUT_FIELDS.index(name)]).max() / h
            err = np.abs(ana - fd) - (rtol * scale + floor)
            if (err > 0).any():
                lu = int(np.argmax(err))
                bad.append((name, dof, lu, ana[lu], fd[lu]))
    return bad


TLEV_10 = [200.0 + 8.0 * i for i in range(11)]
TLEV_8 = [205.0 + 9.0 * i for i in range(9)]

# (id, batch kwargs, reference rtol or None)
CASES = [
    ("single_layer_8", dict(tlev=[200.0, 260.0], tsurf=280.0, nstr=8, ssa=0.0), 1e-3),
    ("multilayer_scattering_16", dict(tlev=TLEV_8, tsurf=290.0, nstr=16, ssa=0.7), 2e-3),
    ("thermal_plus_beam_16", dict(tlev=[210.0, 245.0, 275.0], tsurf=288.0, nstr=16, ssa=0.5, beam=1.0), 1.5e-3),
    ("multilayer_profile_16", dict(tlev=TLEV_10, tsurf=285.0, nstr=16, ssa=0.0), 1e-3),
    ("scattering_high_16", dict(tlev=[210.0, 235.0, 255.0, 270.0, 280.0], tsurf=290.0, nstr=16, ssa=0.9), 2e-3),
    ("diffusion_bc_8", dict(tlev=[210.0, 235.0, 255.0, 270.0, 280.0], tsurf=290.0, nstr=8, ssa=0.6,
                            flags=F_DIFFUSION_BC, dtau=[0.3, 0.5, 0.2, 0.9]), None),
    ("delta_m_12", dict(tlev=[220.0, 240.0, 262.0, 281.0], tsurf=292.0, nstr=12, ssa=0.8, g=[0.85, 0.7, 0.9],
                        nmom=24, dtau=[0.2, 1.5, 0.7], albedo=0.35), None),
    ("sources_top_emission_4", dict(tlev=[190.0, 230.0, 260.0], tsurf=300.0, nstr=4, ssa=0.3, temp_top=180.0,
                                    iso_top=0.2, iso_bot=0.05, beam=0.7, wn=(900.0, 1100.0)), None),
    ("thin_layer_6", dict(tlev=[230.0, 231.0, 260.0], tsurf=270.0, nstr=6, ssa=0.4, dtau=[5e-5, 0.8]), None),
    ("spherical_beam_16", dict(tlev=[210.0, 240.0, 265.0, 282.0], tsurf=290.0, nstr=16, ssa=0.4, beam=1.0,
                               flags=F_SPHERICAL, alt=[3e5, 2e5, 1e5, 0.0], radius=6.371e6), None),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_jacobian_matches_linearised_finite_difference(case):
    bad = fd_failures(thermal_batch(**case[1]), rtol=1e-8, atol=1e-14)
    assert not bad, bad[:5]


@pytest.mark.parametrize("case", [c for c in CASES if c[2]], ids=[c[0] for c in CASES if c[2]])
def test_jacobian_reference_finite_difference_tolerance(case):
    """tests/test_jacobian.cpp:145-176 at the reference's rtol, atol 1e-9."""
    bad = fd_failures(thermal_batch(**case[1]), rtol=case[2], atol=1e-9, linear=False)
    assert not bad, bad[:5]


def test_jacobian_leaves_forward_outputs_unchanged():
    """tests/test_jacobian.cpp:284-309: requesting the Jacobian changes no output."""
    fb = thermal_batch([220.0, 250.0, 275.0], 285.0, 8, 0.0)
    out, _ = ol.flux_jacobian_one(fb, 0)
    assert np.array_equal(out, ol.flux_one(fb, 0))


def test_jacobian_index_from_bottom_reversal():
    """tests/test_jacobian.cpp:316-355 for the flux solver: inputs supplied
    bottom-up give the top-down Jacobian with rows and level columns reversed
    and the surface column last."""
    tlev = [210.0, 240.0, 265.0, 285.0]
    L = len(tlev) - 1
    _, j0 = ol.flux_jacobian_one(thermal_batch(tlev, 295.0, 16, 0.3), 0)
    _, j1 = ol.flux_jacobian_one(thermal_batch(tlev[::-1], 295.0, 16, 0.3), 0, index_from_bottom=True)
    cols = list(range(L, -1, -1)) + [L + 1]
    np.testing.assert_allclose(j1, j0[:, ::-1][:, :, cols], rtol=1e-12, atol=1e-15)


def test_spherical_equals_plane_parallel_without_beam():
    """tests/test_jacobian.cpp:254-278: with no beam the pseudo-spherical
    option changes nothing in the thermal Jacobian."""
    kw = dict(tlev=[210.0, 240.0, 265.0, 282.0], tsurf=290.0, nstr=16, ssa=0.4)
    _, jp = ol.flux_jacobian_one(thermal_batch(**kw), 0)
    _, js = ol.flux_jacobian_one(thermal_batch(**kw, flags=F_SPHERICAL, alt=[3e5, 2e5, 1e5, 0.0],
                                               radius=6.371e6), 0)
    np.testing.assert_allclose(js, jp, rtol=0, atol=1e-10)


def test_surface_column_sums_with_levels_to_isothermal_derivative():
    """An isothermal column with a black surface and no scattering: raising
    every temperature together by dT changes every level's upward flux by
    pi dB/dT dT (the upwelling field stays B everywhere).  Pins the sum of
    all columns of flux_up, independently of the linearisation."""
    T = 250.0
    fb = thermal_batch([T] * 5, T, 8, 0.0, albedo=0.0, dtau=[0.3, 1.1, 0.05, 2.0])
    _, jac = ol.flux_jacobian_one(fb, 0)
    want = np.pi * ol.planck_deriv(500.0, 800.0, T)
    np.testing.assert_allclose(jac[0].sum(axis=1), want, rtol=1e-12)


def test_planted_jacobian_error_is_caught():
    """A 1e-6 relative error planted in the oracle's Jacobian output fails
    the linearised finite-difference pin."""
    import ctypes as C
    l = ol.lib()
    l.oracle_set_perturbation.argtypes = [C.c_int, C.c_double]
    try:
        l.oracle_set_perturbation(FLUX_STAGES + PS_JACOBIAN, 1e-6)
        assert fd_failures(thermal_batch(**CASES[1][1]), rtol=1e-8, atol=1e-14)
    finally:
        l.oracle_set_perturbation(0, 0.0)
