"""Temperature Jacobians through the DisORT++-compatible facade:
DisortFluxSolver<N> / DisortSolver with compute_temperature_jacobian, the
spectral entry points and index_from_bottom, against the oracle's analytic
linearisation (tests/test_jacobian.cpp is the reference's coverage).
Criterion: |gpu - oracle| <= 1e-9 |oracle| + 1e-12 max|oracle| per config."""
import numpy as np
import pytest

import oracle_lib as ol
from test_oracle_general_jacobian import general_batch
from test_oracle_jacobian import JAC_FIELDS, thermal_batch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def d():
    import disort_amd.disortpp as mod

    return mod


def jac_of(res):
    return np.array([np.array(getattr(res, f"{'flux_divergence' if f == 'flux_tau_divergence' else f}_temperature_jac"))
                     for f in JAC_FIELDS])


def check(got, ref, label):
    assert got.shape == ref.shape, (label, got.shape, ref.shape)
    floor = 1e-12 * np.abs(ref).max()
    bad = ~(np.abs(got - ref) <= 1e-9 * np.abs(ref) + floor)
    assert not bad.any(), f"{label}: {int(bad.sum())} Jacobian entries off"


def flux_config(d, tlev, tsurf, nstr, ssa, beam=0.0, rev=False):
    """tests/test_jacobian.cpp:117-140 (buildFluxThermalConfig)."""
    L = len(tlev) - 1
    c = d.DisortFluxConfig(L, nstr)
    c.use_thermal_emission = True
    c.direct_beam_flux = beam
    c.direct_beam_mu = 0.5
    c.surface_albedo = 0.1
    c.temperature_bottom = tsurf
    c.temperature_top = 0.0
    c.emissivity_top = 1.0
    c.wavenumber_low, c.wavenumber_high = 500.0, 800.0
    c.allocate()
    c.delta_tau = [0.4] * L
    c.single_scat_albedo = [ssa] * L
    pm = np.zeros((L, nstr + 1))
    pm[:, 0], pm[:, 1] = 1.0, 0.3 * ssa
    c.phase_function_moments = pm.tolist()
    c.temperature = list(tlev[::-1]) if rev else list(tlev)
    c.index_from_bottom = rev
    c.compute_temperature_jacobian = True
    return c


def general_config(d, tlev, tsurf, nstr, ssa, beam=0.0, rev=False, only_fluxes=False):
    """tests/test_jacobian.cpp:19-69 (buildThermalConfig)."""
    L = len(tlev) - 1
    c = d.DisortConfig(L, nstr, 0)
    c.flags.use_thermal_emission = True
    c.flags.comp_only_fluxes = only_fluxes
    c.flags.compute_temperature_jacobian = True
    c.flags.index_from_bottom = rev
    c.bc.direct_beam_flux = beam
    c.bc.direct_beam_mu = 0.5
    c.bc.surface_albedo = 0.1
    c.bc.temperature_bottom = tsurf
    c.bc.temperature_top = 0.0
    c.bc.emissivity_top = 1.0
    c.wavenumber_low, c.wavenumber_high = 500.0, 800.0
    c.accuracy_fourier_series = 1e-4
    c.num_phi = 1
    c.allocate()
    c.phi_user = [0.0]
    c.delta_tau = [0.4] * L
    c.single_scat_albedo = [ssa] * L
    pm = np.zeros((L, nstr + 1))
    pm[:, 0], pm[:, 1] = 1.0, 0.3 * ssa
    c.phase_function_moments = pm.tolist()
    c.temperature = list(tlev[::-1]) if rev else list(tlev)
    return c


TLEV = [210.0, 240.0, 265.0, 282.0]


@pytest.mark.parametrize("nstr,ssa,beam", [(8, 0.0, 0.0), (16, 0.7, 0.0), (16, 0.5, 1.0)])
def test_flux_solver_jacobian(d, nstr, ssa, beam):
    r = d.create_flux_solver(nstr).solve(flux_config(d, TLEV, 290.0, nstr, ssa, beam))
    _, ref = ol.flux_jacobian_one(thermal_batch(TLEV, 290.0, nstr, ssa, beam=beam), 0)
    check(jac_of(r), ref, f"DisortFluxSolver{nstr}")


def test_flux_solver_jacobian_index_from_bottom(d):
    r = d.DisortFluxSolver16().solve(flux_config(d, TLEV, 295.0, 16, 0.3, rev=True))
    _, ref = ol.flux_jacobian_one(thermal_batch(TLEV[::-1], 295.0, 16, 0.3), 0, index_from_bottom=True)
    check(jac_of(r), ref, "index_from_bottom")


def test_jacobian_off_or_without_thermal_is_empty(d):
    c = flux_config(d, TLEV, 290.0, 8, 0.2)
    c.compute_temperature_jacobian = False
    r = d.DisortFluxSolver8().solve(c)
    assert len(r.flux_up_temperature_jac) == 0
    c = flux_config(d, TLEV, 290.0, 8, 0.2)
    c.use_thermal_emission = False
    c.direct_beam_flux = 1.0
    r = d.DisortFluxSolver8().solve(c)
    assert len(r.flux_up_temperature_jac) == 0 and len(r.flux_divergence_temperature_jac) == 0


def test_flux_spectral_bands_jacobian(d):
    bands = [(300.0, 400.0), (500.0, 800.0), (1000.0, 1005.0)]
    rs = d.solve_flux_spectral_bands(flux_config(d, TLEV, 290.0, 12, 0.4), bands)
    for (lo, hi), r in zip(bands, rs):
        _, ref = ol.flux_jacobian_one(thermal_batch(TLEV, 290.0, 12, 0.4, wn=(lo, hi)), 0)
        check(jac_of(r), ref, f"band {lo}-{hi}")


def test_flux_batch_mixed_jacobian_requests(d):
    a = flux_config(d, TLEV, 290.0, 8, 0.2)
    b = flux_config(d, TLEV[:3], 280.0, 16, 0.6)
    b.compute_temperature_jacobian = False
    c = flux_config(d, TLEV, 285.0, 16, 0.1, rev=True)  # temperatures given bottom-up
    ra, rb, rc = d.solve_flux_batch([a, b, c])
    check(jac_of(ra), ol.flux_jacobian_one(thermal_batch(TLEV, 290.0, 8, 0.2), 0)[1], "a")
    assert len(rb.flux_up_temperature_jac) == 0
    check(jac_of(rc), ol.flux_jacobian_one(thermal_batch(TLEV[::-1], 285.0, 16, 0.1), 0, True)[1], "c")


@pytest.mark.parametrize("nstr,ssa,beam,only", [(8, 0.0, 0.0, False), (16, 0.5, 0.0, False), (16, 0.6, 1.0, True)])
def test_general_solver_jacobian(d, nstr, ssa, beam, only):
    r = d.DisortSolver().solve(general_config(d, TLEV, 290.0, nstr, ssa, beam, only_fluxes=only))
    _, ref = ol.general_jacobian_one(general_batch(TLEV, 290.0, nstr, ssa, beam=beam), 0)
    check(jac_of(r), ref, f"DisortSolver nstr={nstr}")


def test_general_solver_jacobian_user_tau_rpv_index_from_bottom(d):
    c = general_config(d, TLEV, 295.0, 12, 0.4, beam=1.0, rev=True)
    c.flags.use_user_tau = True
    c.num_user_tau = 4
    c.tau_user = [0.0, 0.3, 0.9, 1.2]  # measured from the bottom
    c.flags.use_lambertian_surface = False
    c.flags.brdf_type = d.BrdfType.RPV
    spec = d.RpvBrdfSpec()
    spec.rho0, spec.k, spec.theta, spec.sigma, spec.t1, spec.t2, spec.scale = 0.027, 0.647, -0.169, 0.0, 0.0, 0.0, 1.0
    c.brdf.set_rpv(spec)
    r = d.DisortSolver().solve(c)
    gb = general_batch(TLEV[::-1], 295.0, 12, 0.4, beam=1.0, brdf="rpv", tau_user=[[0.0, 0.3, 0.9, 1.2]])
    _, ref = ol.general_jacobian_one(gb, 0, index_from_bottom=True)
    check(jac_of(r), ref, "user tau + RPV + index_from_bottom")


def test_spectral_bands_general_jacobian(d):
    bands = [(400.0, 500.0), (500.0, 800.0)]
    rs = d.solve_spectral_bands(general_config(d, TLEV, 290.0, 8, 0.3), bands)
    for (lo, hi), r in zip(bands, rs):
        _, ref = ol.general_jacobian_one(general_batch(TLEV, 290.0, 8, 0.3, wn=(lo, hi)), 0)
        check(jac_of(r), ref, f"band {lo}-{hi}")
