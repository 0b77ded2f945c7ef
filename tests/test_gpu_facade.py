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


# --- DisortSolver with intensities (the general GPU path) -------------------
def iv_config(d, tau, w, beam, rpv=None, nmom=0):
    """IntensityValidation-style single layer (tests/test_solver.cpp:2613-2670)."""
    c = d.DisortConfig(1, 16, nmom)
    c.bc.direct_beam_flux = beam
    c.bc.direct_beam_mu = 0.1 if rpv is None else 0.6
    c.flags.use_user_mu = True
    c.accuracy_fourier_series = 1e-4
    c.num_user_mu = 6
    c.num_phi = 1
    if rpv is not None:
        c.flags.use_user_tau = True
        c.num_user_tau = 2
        c.flags.use_lambertian_surface = False
        c.flags.brdf_type = d.BrdfType.RPV
        c.accuracy_fourier_series = 0.0
    c.allocate()
    c.set_isotropic(0)
    c.delta_tau = [tau]
    c.single_scat_albedo = [w]
    c.mu_user = [-1.0, -0.5, -0.1, 0.1, 0.5, 1.0]
    c.phi_user = [0.0]
    if rpv is not None:
        c.tau_user = [0.0, tau]
        spec = d.RpvBrdfSpec()
        spec.rho0, spec.k, spec.theta = rpv
        c.brdf.set_rpv(spec)
    return c


def test_general_solver_intensity_table_1a(d):
    c = iv_config(d, 0.03125, 0.2, math.pi / 0.1)
    r = d.DisortSolver().solve(c)
    assert r.num_user_mu() == 6 and r.num_phi() == 1 and r.num_user_tau() == 2
    # C-DISORT values quoted by tests/test_solver.cpp:2654-2669 (5e-3).
    for (iu, lu), v in {(3, 0): 1.17770662e-01, (4, 0): 2.64170381e-02, (5, 0): 1.34041302e-02,
                        (0, 1): 1.33826267e-02, (1, 1): 2.63323514e-02, (2, 1): 1.15897886e-01}.items():
        assert r.intensities(iu, lu, 0) == pytest.approx(v, rel=5e-3)
    for iu in range(3):
        assert abs(r.intensities(iu, 0, 0)) < 1e-10 and abs(r.intensities(iu + 3, 1, 0)) < 1e-10


def test_general_solver_rpv_table(d):
    """RpvBRDF1 (tests/test_solver.cpp:4133-4189), 1e-3."""
    c = iv_config(d, 0.3, 0.9, 1.0, rpv=(0.1, 0.9, -0.2), nmom=16)
    r = d.DisortSolver().solve(c)
    for (iu, lu), v in {(3, 0): 9.33168983e-02, (4, 0): 5.20589621e-02, (5, 0): 4.53763077e-02,
                        (0, 1): 2.37650510e-02, (1, 1): 4.10887262e-02, (2, 1): 8.22310683e-02}.items():
        assert r.intensities(iu, lu, 0) == pytest.approx(v, rel=1e-3)


def general_cfg(d, L=5, seed=3, rev=False):
    r = np.random.default_rng(seed)
    c = d.DisortConfig(L, 8, 0)
    c.flags.use_user_mu = True
    c.flags.use_user_tau = True
    c.flags.use_thermal_emission = True
    c.flags.index_from_bottom = rev
    c.num_user_mu, c.num_user_tau, c.num_phi = 4, 3, 3
    c.allocate()
    dt = r.uniform(0.1, 1.0, L)
    c.delta_tau = list(dt[::-1] if rev else dt)
    w = r.uniform(0.3, 0.95, L)
    c.single_scat_albedo = list(w[::-1] if rev else w)
    g = r.uniform(0.1, 0.8, L)
    for lc in range(L):
        c.set_henyey_greenstein(float(g[L - 1 - lc] if rev else g[lc]), lc)
    t = np.linspace(220.0, 290.0, L + 1)
    c.temperature = list(t[::-1] if rev else t)
    c.wavenumber_low, c.wavenumber_high = 600.0, 700.0
    c.bc.temperature_bottom, c.bc.temperature_top = 295.0, 0.0
    c.bc.direct_beam_flux, c.bc.direct_beam_mu, c.bc.surface_albedo = math.pi, 0.55, 0.2
    tot = float(dt.sum())
    tu = [0.0, 0.4 * tot, tot]
    c.tau_user = [tot - x for x in tu[::-1]] if rev else tu
    c.mu_user = [-0.8, -0.3, 0.4, 0.9]
    c.phi_user = [0.0, 60.0, 180.0]
    return c


def test_general_solver_index_from_bottom(d):
    top = d.DisortSolver().solve(general_cfg(d))
    bot = d.DisortSolver().solve(general_cfg(d, rev=True))
    np.testing.assert_allclose(np.array(bot.intensity)[::-1], np.array(top.intensity), rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(bot.flux_up[::-1], top.flux_up, rtol=1e-10, atol=1e-13)


def test_spectral_bands_with_intensities(d):
    """solve_spectral_bands on the full solver: one GPU batch, each band
    equal to a single DisortSolver solve."""
    c = general_cfg(d)
    bands = [(600.0 + 10 * k, 610.0 + 10 * k) for k in range(4)]
    out = d.solve_spectral_bands(c, bands, surface_albedo=[0.0, 0.1, 0.2, 0.3])
    for k, (lo, hi) in enumerate(bands):
        one = general_cfg(d)
        one.wavenumber_low, one.wavenumber_high = lo, hi
        one.bc.surface_albedo = 0.1 * k
        ref = d.DisortSolver().solve(one)
        np.testing.assert_allclose(np.array(out[k].intensity), np.array(ref.intensity), rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(out[k].flux_up, ref.flux_up, rtol=1e-12, atol=1e-15)


def test_general_solver_buras_emde_table(d):
    """NewICHG3a (tests/test_solver.cpp:3940-4008): HG g = 0.75 tabulated at
    100 cosines, intensity_corr_buras without intensity_corr_nakajima; the
    C-DISORT values the reference quotes (at 1e-3; they hold to 1e-8)."""
    c = d.DisortConfig(1, 16, 16)
    c.flags.use_user_tau = True
    c.flags.use_user_mu = True
    c.flags.intensity_corr_buras = True
    c.flags.intensity_corr_nakajima = False
    c.num_phase_func_angles = 100
    c.num_user_tau = 2
    c.num_user_mu = 6
    c.num_phi = 1
    c.allocate()
    c.set_henyey_greenstein(0.75, 0)
    ct = [-1.0 + 2.0 * it / 99 for it in range(100)]
    c.mu_phase_function = ct
    c.phase_function = [[(1 - 0.75 ** 2) / (1 + 0.75 ** 2 - 1.5 * x) ** 1.5 for x in ct]]
    c.single_scat_albedo = [1.0]
    c.delta_tau = [1.0]
    c.tau_user = [0.0, 1.0]
    c.mu_user = [-1.0, -0.5, -0.1, 0.1, 0.5, 1.0]
    c.phi_user = [0.0]
    c.bc.direct_beam_flux = math.pi
    c.bc.direct_beam_mu = 1.0
    r = d.DisortSolver().solve(c)
    for (iu, lu), v in {(3, 0): 1.51254859e-01, (4, 0): 1.01163783e-01, (5, 0): 3.92360939e-02,
                        (0, 1): 3.06133539e+00, (1, 1): 2.66600162e-01, (2, 1): 2.13783659e-01}.items():
        assert r.intensities(iu, lu, 0) == pytest.approx(v, rel=1e-8)


def special_config(d, dtau, ssa, mu, albedo, g=0.8, nstr=16, user_mu=True):
    c = d.DisortConfig(len(dtau), nstr, 0)
    c.flags.ibcnd = d.BoundaryConditionType.Special
    c.flags.use_user_mu = user_mu
    c.num_user_mu = len(mu) if user_mu else 0
    c.allocate()
    c.delta_tau = list(dtau)
    c.single_scat_albedo = list(ssa)
    for lc in range(len(dtau)):
        c.set_henyey_greenstein(g, lc)
    if user_mu:
        c.mu_user = list(mu)
    c.bc.surface_albedo = albedo
    return c


def test_special_bc_tables(d):
    """SpecialBC13a / 13c (tests/test_solver.cpp:4206-4318): the C-DISORT
    values to 11 digits (asserted at 1e-10)."""
    for dtau, ssa, alb, trn in (([1.0], [0.99], 5.4525838793e-01, 8.4499871160e-01),
                                ([0.5, 0.5], [0.99, 0.50], 2.7620066856e-01, 5.0331887919e-01)):
        c = special_config(d, dtau, ssa, [0.5], 0.5)
        r = d.DisortSolver().solve(c)
        assert r.albedo_medium[0] == pytest.approx(alb, rel=1e-10)
        assert r.transmissivity_medium[0] == pytest.approx(trn, rel=1e-10)
        assert list(r.mu_angles) == [-0.5, 0.5] and c.num_user_mu == 1


@pytest.mark.parametrize("nstr,L,user_mu,albedo", [(8, 1, True, 0.0), (16, 3, True, 0.3), (32, 4, False, 0.2),
                                                   (64, 2, True, 0.0)])
def test_special_bc_against_oracle(d, nstr, L, user_mu, albedo):
    from disort_amd import general as G
    from disort_amd.batch import scalars

    r = np.random.default_rng(nstr + L)
    dtau = r.uniform(0.05, 2.0, L)
    ssa = r.uniform(0.3, 0.999, L)
    mu = [0.15, 0.4, 0.75, 1.0]
    c = special_config(d, dtau, ssa, mu, albedo, g=0.7, nstr=nstr, user_mu=user_mu)
    res = d.DisortSolver().solve(c)
    pm = np.repeat((0.7 ** np.arange(nstr + 1))[None, None, :], L, axis=1)
    gb = G.GeneralBatch(flags=G.G_USER_MU if user_mu else 0, nstr=nstr, dtau=dtau[None], ssa=ssa[None], pmom=pm,
                        scal=scalars(1, albedo=albedo), mu_user=mu if user_mu else None, phi_user=[0.0])
    alb, trn, sph = ol.general_albtrans(gb, 0)
    check(np.array(res.albedo_medium), alb, "albedo_medium")
    check(np.array(res.transmissivity_medium), trn, "transmissivity_medium")
    assert res.spherical_albedo == pytest.approx(sph[0], rel=1e-9, abs=1e-12)
    assert res.spherical_transmissivity == pytest.approx(sph[1], rel=1e-9, abs=1e-12)


def test_special_bc_spectral_bands(d):
    c = special_config(d, [0.4, 0.9], [0.95, 0.6], [0.3, 0.8], 0.1)
    bands = [(500.0 + k, 501.0 + k) for k in range(3)]
    res = d.solve_spectral_bands(c, bands, single_scat_albedo=[[0.95, 0.6], [0.9, 0.5], [0.99, 0.2]])
    one = [d.DisortSolver().solve(special_config(d, [0.4, 0.9], w, [0.3, 0.8], 0.1))
           for w in ([0.95, 0.6], [0.9, 0.5], [0.99, 0.2])]
    for a, b in zip(res, one):
        np.testing.assert_allclose(a.albedo_medium, b.albedo_medium, rtol=1e-12)
        np.testing.assert_allclose(a.transmissivity_medium, b.transmissivity_medium, rtol=1e-12)
