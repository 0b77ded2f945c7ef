"""NumPy arrays as the per-band overrides of the facade's spectral drivers
(solve_flux_spectral_bands and friends; the reference's bindings take nested
lists, src/python_bindings.cpp:633-763).  The array form must give exactly
the nested-list form's results (the same GPU batch, bit for bit), packed
from the arrays without per-band configurations."""
import math
import time

import numpy as np
import pytest

import oracle_lib as ol
from disort_amd import workloads as W
from disort_amd.batch import OUT_FIELDS

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def d():
    import disort_amd.disortpp as mod

    return mod


def arr(res):
    return np.array([[getattr(r, f) for f in OUT_FIELDS] for r in res])


def c2_setup(d, n, first=0, thermal=True):
    b = W.c2(n, first_band=first)
    L, N = 60, 16
    cfg = d.DisortFluxConfig(L, N)
    cfg.use_thermal_emission = thermal
    cfg.allocate()
    s = b.scal.reshape(n, -1)
    if thermal:
        cfg.temperature = list(b.temper[: L + 1])
    cfg.temperature_top, cfg.temperature_bottom, cfg.emissivity_top = 0.0, 300.0, 1.0
    cfg.direct_beam_flux, cfg.direct_beam_mu = math.pi, float(s[0, 1])
    bands = [(float(s[i, 8]), float(s[i, 9])) for i in range(n)]
    ov = dict(delta_tau=b.dtau.reshape(n, L), single_scat_albedo=b.ssa.reshape(n, L),
              phase_function_moments=b.pmom.reshape(n, L, N + 1), surface_albedo=s[:, 7].copy())
    return b, cfg, bands, ov


def as_lists(ov):
    return {k: v.tolist() for k, v in ov.items()}


def test_c2_arrays_equal_lists_and_oracle(d):
    n = 400
    b, cfg, bands, ov = c2_setup(d, n, first=1234)
    d.solve_flux_spectral_bands(cfg, bands[:2], **{k: v[:2] for k, v in ov.items()})  # context, kernels
    t0 = time.perf_counter()
    ra = d.solve_flux_spectral_bands(cfg, bands, **ov)
    t1 = time.perf_counter()
    rl = d.solve_flux_spectral_bands(cfg, bands, **as_lists(ov))
    t2 = time.perf_counter()
    print(f"[arrays] {n} bands: arrays {1e3 * (t1 - t0):.1f} ms, lists {1e3 * (t2 - t1):.1f} ms", flush=True)
    np.testing.assert_array_equal(arr(ra), arr(rl))
    for i in (0, 201, n - 1):
        ref = ol.flux_one(b, i)
        got = arr([ra[i]])[0]
        assert (np.abs(got - ref) <= 1e-9 * np.abs(ref) + 1e-12).all(), i


def test_c2_full_10000_bands_through_arrays(d):
    """The whole C2 workload through the facade with NumPy overrides."""
    n = 10000
    b, cfg, bands, ov = c2_setup(d, n)
    t0 = time.perf_counter()
    res = d.solve_flux_spectral_bands(cfg, bands, **ov)
    print(f"[arrays] C2 10000 bands via solve_flux_spectral_bands: {time.perf_counter() - t0:.3f} s", flush=True)
    assert len(res) == n
    for i in (0, 4999, 9999):
        ref = ol.flux_one(b, i)
        got = arr([res[i]])[0]
        assert (np.abs(got - ref) <= 1e-9 * np.abs(ref) + 1e-12).all(), i


def test_index_from_bottom_and_wider_rows(d):
    """Rows longer than num_layers / nmomNstr()+1 and bottom-up input use
    the same entries the nested lists do (the first num_layers after the
    reversal)."""
    n, L, N = 12, 8, 8
    r = np.random.default_rng(4)
    cfg = d.DisortFluxConfig(L, N, 12)
    cfg.use_thermal_emission = True
    cfg.index_from_bottom = True
    cfg.allocate()
    cfg.temperature = list(np.linspace(290, 210, L + 1))
    cfg.temperature_bottom = 295.0
    cfg.direct_beam_flux, cfg.direct_beam_mu = 1.0, 0.6
    bands = [(500.0 + 10 * i, 520.0 + 10 * i) for i in range(n)]
    g = r.uniform(0.1, 0.8, (n, L + 3))
    ov = dict(delta_tau=r.uniform(0.01, 2.0, (n, L + 3)), single_scat_albedo=r.uniform(0.1, 0.95, (n, L + 3)),
              phase_function_moments=g[..., None] ** np.arange(16)[None, None, :],
              emissivity_top=r.uniform(0.5, 1.0, n), isotropic_flux_top=r.uniform(0, 0.1, n))
    ra = d.solve_flux_spectral_bands(cfg, bands, **ov)
    rl = d.solve_flux_spectral_bands(cfg, bands, **as_lists(ov))
    np.testing.assert_array_equal(arr(ra), arr(rl))


def test_non_contiguous_and_integer_arrays(d):
    n = 16
    b, cfg, bands, ov = c2_setup(d, n, first=77)
    ref = arr(d.solve_flux_spectral_bands(cfg, bands, **ov))
    wide = np.zeros((n, 2 * 60))
    wide[:, ::2] = ov["delta_tau"]
    ov2 = dict(ov, delta_tau=wide[:, ::2], surface_albedo=np.asfortranarray(ov["surface_albedo"]))
    np.testing.assert_array_equal(arr(d.solve_flux_spectral_bands(cfg, bands, **ov2)), ref)
    cfg.direct_beam_flux = 0.0
    z = d.solve_flux_spectral_bands(cfg, bands, direct_beam_flux=np.zeros(n, dtype=np.int64), **{
        k: v for k, v in ov.items()})
    assert all(r.flux_direct_beam[-1] == 0.0 for r in z)


def test_ragged_lists_still_accepted(d):
    n, L = 3, 4
    cfg = d.DisortFluxConfig(L, 8)
    cfg.allocate()
    cfg.single_scat_albedo = [0.5] * L
    cfg.set_henyey_greenstein(0.5)
    cfg.direct_beam_flux, cfg.direct_beam_mu = 1.0, 0.5
    dt = [[0.1] * L, [0.2] * (L + 2), [0.3] * L]  # ragged: the list form
    res = d.solve_flux_spectral_bands(cfg, [(0.0, 0.0)] * n, delta_tau=dt)
    for i, r in enumerate(res):
        c = d.DisortFluxConfig(L, 8)
        c.allocate()
        c.single_scat_albedo = [0.5] * L
        c.set_henyey_greenstein(0.5)
        c.direct_beam_flux, c.direct_beam_mu = 1.0, 0.5
        c.delta_tau = dt[i]
        np.testing.assert_array_equal(arr([r]), arr([d.DisortFluxSolver8().solve(c)]))


def test_invalid_band_raises_reference_message(d):
    n = 10
    b, cfg, bands, ov = c2_setup(d, n, first=5)
    ssa = ov["single_scat_albedo"].copy()
    ssa[7, 3] = 1.5
    with pytest.raises(ValueError, match=r"^single_scat_albedo\[3\] must be in \[0, 1\]$"):
        d.solve_flux_spectral_bands(cfg, bands, **dict(ov, single_scat_albedo=ssa))


def test_jacobians_through_arrays(d):
    n = 6
    b, cfg, bands, ov = c2_setup(d, n, first=300)
    cfg.compute_temperature_jacobian = True
    ra = d.solve_flux_spectral_bands(cfg, bands, **ov)
    rl = d.solve_flux_spectral_bands(cfg, bands, **as_lists(ov))
    for x, y in zip(ra, rl):
        np.testing.assert_array_equal(np.array(x.flux_up_temperature_jac), np.array(y.flux_up_temperature_jac))
        np.testing.assert_array_equal(np.array(x.flux_divergence_temperature_jac),
                                      np.array(y.flux_divergence_temperature_jac))
    assert np.array(ra[0].flux_up_temperature_jac).shape == (61, 62)


def test_general_spectral_arrays_equal_lists(d):
    n, L = 5, 4
    r = np.random.default_rng(9)
    cfg = d.DisortConfig(L, 8, 0)
    cfg.num_phi = 2
    cfg.allocate()
    cfg.phi_user = [0.0, 90.0]
    cfg.set_henyey_greenstein(0.6)
    cfg.bc.direct_beam_flux, cfg.bc.direct_beam_mu = 1.0, 0.7
    ov = dict(delta_tau=r.uniform(0.1, 1.0, (n, L)), single_scat_albedo=r.uniform(0.2, 0.9, (n, L)),
              surface_albedo=r.uniform(0, 0.3, n))
    bands = [(100.0, 101.0)] * n
    ra = d.solve_spectral_bands(cfg, bands, **ov)
    rl = d.solve_spectral_bands(cfg, bands, **as_lists(ov))
    for x, y in zip(ra, rl):
        np.testing.assert_array_equal(np.array(x.flux_up), np.array(y.flux_up))
        np.testing.assert_array_equal(np.array(x.intensity), np.array(y.intensity))
