"""The DisORT++-compatible facade (`disortpp`) on the GPU: same numbers as
the C ABI path and the oracle, reference semantics for the spectral
overrides, index_from_bottom and the general solver's flux-only mode.
Tolerance: |gpu - oracle| <= 1e-9 |oracle| + 1e-12 per output."""
import math

import numpy as np
import pytest

import oracle_lib as ol
from disort_amd import workloads as W
from disort_amd.batch import OUT_FIELDS

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-9, 1e-12


def as_array(res):
    return np.array([getattr(res, f) for f in OUT_FIELDS])


def check(got, ref, label):
    assert np.isfinite(got).all(), label
    bad = ~(np.abs(got - ref) <= RTOL * np.abs(ref) + ATOL)
    assert not bad.any(), f"{label}: {bad.sum()} outputs out of tolerance"


def c1_flux_config(d):
    c = d.DisortFluxConfig(4, 16)
    c.allocate()
    c.delta_tau = [0.1, 0.5, 1.0, 2.0]
    c.single_scat_albedo = [0.9] * 4
    c.set_henyey_greenstein(0.8)
    c.direct_beam_flux = math.pi
    c.direct_beam_mu = 0.5
    c.surface_albedo = 0.1
    return c


@pytest.fixture(scope="module")
def d():
    import disort_amd.disortpp as mod

    return mod


def test_c1_flux_solver(d):
    r = d.DisortFluxSolver16().solve(c1_flux_config(d))
    assert r.num_levels() == 5
    check(as_array(r), ol.flux_one(W.c1(), 0), "c1 via DisortFluxSolver16")
    assert r.net_flux(0) == pytest.approx(r.flux_direct_beam[0] + r.flux_down[0] - r.flux_up[0])


def test_general_solver_flux_mode_equals_flux_solver(d):
    cfg = d.DisortConfig(4, 16, 0)
    cfg.flags.comp_only_fluxes = True
    cfg.allocate()
    cfg.delta_tau = [0.1, 0.5, 1.0, 2.0]
    cfg.single_scat_albedo = [0.9] * 4
    cfg.set_henyey_greenstein(0.8)
    cfg.bc.direct_beam_flux = math.pi
    cfg.bc.direct_beam_mu = 0.5
    cfg.bc.surface_albedo = 0.1
    r = d.DisortSolver().solve(cfg)
    f = d.DisortFluxSolver16().solve(c1_flux_config(d))
    np.testing.assert_array_equal(as_array(r), as_array(f))
    # The reference fills these in the caller's config.
    assert cfg.num_user_tau == 5 and cfg.tau_user == pytest.approx((0.0, 0.1, 0.6, 1.6, 3.6), abs=1e-15)
    assert cfg.num_user_mu == 16 and len(r.mu_angles) == 16
    cmu, _ = ol.gauss(8)
    np.testing.assert_allclose(r.mu_angles[8:], cmu[:8], rtol=0, atol=0)
    np.testing.assert_allclose(r.mu_angles[:8], -cmu[:8][::-1], rtol=0, atol=0)


def test_index_from_bottom(d):
    top = c1_flux_config(d)
    top.use_thermal_emission = True
    top.temperature = [200.0, 220.0, 240.0, 260.0, 280.0]
    top.temperature_bottom = 290.0
    top.wavenumber_low, top.wavenumber_high = 500.0, 600.0
    bot = c1_flux_config(d)
    bot.use_thermal_emission = True
    bot.index_from_bottom = True
    bot.delta_tau = top.delta_tau[::-1]
    bot.single_scat_albedo = top.single_scat_albedo[::-1]
    bot.phase_function_moments = top.phase_function_moments[::-1]
    bot.temperature = top.temperature[::-1]
    bot.temperature_bottom = 290.0
    bot.wavenumber_low, bot.wavenumber_high = 500.0, 600.0
    a = as_array(d.DisortFluxSolver16().solve(top))
    b = as_array(d.DisortFluxSolver16().solve(bot))
    np.testing.assert_array_equal(b, a[:, ::-1])


def test_flux_spectral_bands_c2_subset(d):
    n, first = 48, 2500
    b = W.c2(n, first_band=first)
    L, N = 60, 16
    cfg = d.DisortFluxConfig(L, N)
    cfg.use_thermal_emission = True
    cfg.allocate()
    s = b.scal.reshape(n, -1)
    cfg.temperature = list(b.temper[: L + 1])
    cfg.temperature_top, cfg.temperature_bottom, cfg.emissivity_top = 0.0, 300.0, 1.0
    cfg.direct_beam_flux, cfg.direct_beam_mu = math.pi, float(s[0, 1])
    bands = [(float(s[i, 8]), float(s[i, 9])) for i in range(n)]
    res = d.solve_flux_spectral_bands(
        cfg, bands,
        delta_tau=b.dtau.reshape(n, L).tolist(),
        single_scat_albedo=b.ssa.reshape(n, L).tolist(),
        phase_function_moments=b.pmom.reshape(n, L, N + 1).tolist(),
        surface_albedo=s[:, 7].tolist(),
    )
    assert len(res) == n
    for i in (0, 17, n - 1):
        check(as_array(res[i]), ol.flux_one(b, i), f"band {first + i}")


def test_spectral_bands_general_api(d):
    cfg = d.DisortConfig(4, 16, 0)
    cfg.flags.comp_only_fluxes = True
    cfg.allocate()
    cfg.delta_tau = [0.1, 0.5, 1.0, 2.0]
    cfg.single_scat_albedo = [0.9] * 4
    cfg.set_henyey_greenstein(0.8)
    cfg.bc.direct_beam_flux = math.pi
    cfg.bc.direct_beam_mu = 0.5
    out = d.solve_spectral_bands(cfg, [(100.0, 101.0)] * 3, surface_albedo=[0.0, 0.1, 0.2])
    ref = as_array(d.DisortFluxSolver16().solve(c1_flux_config(d)))
    check(as_array(out[1]), ref, "band 1 of solve_spectral_bands vs DisortFluxSolver16")
    assert out[2].flux_up[0] > out[1].flux_up[0] > out[0].flux_up[0]


def test_ragged_flux_batch(d):
    cfgs = []
    for k, (L, N) in enumerate([(3, 4), (7, 8), (12, 16), (5, 6), (9, 14)]):
        c = d.DisortFluxConfig(L, N)
        c.allocate()
        c.delta_tau = [0.05 * (j + 1) * (k + 1) for j in range(L)]
        c.single_scat_albedo = [0.5 + 0.04 * j for j in range(L)]
        c.set_henyey_greenstein(0.3 + 0.1 * k)
        c.direct_beam_flux, c.direct_beam_mu, c.surface_albedo = 1.0, 0.3 + 0.1 * k, 0.2
        cfgs.append(c)
    batch = d.solve_flux_batch(cfgs)
    for c, r in zip(cfgs, batch):
        single = d.create_flux_solver(c.num_streams).solve(c)
        check(as_array(r), as_array(single), f"ragged batch nstr={c.num_streams} nlyr={c.num_layers}")
