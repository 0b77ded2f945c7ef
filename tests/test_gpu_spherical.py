"""Pseudo-spherical beam on the GPU (workgroup kernels) against the oracle,
which is pinned on the C-DISORT spherical values the reference tests quote
(tests/test_solver.cpp:4459-4532).  Tolerance |gpu - oracle| <= 1e-9 |oracle|
+ 1e-12 per output."""
import numpy as np
import pytest

from disort_amd import workloads as W
from disort_amd.batch import F_DELTA_M_PLUS, F_SPHERICAL, F_THERMAL, FluxBatch, scalars, uniform_batch

from test_gpu_flux import solve_and_check
from test_oracle_flux import FIELD, spherical_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from disort_amd._capi import Context

    return Context(0)


def test_cdisort_spherical_case(ctx):
    b = spherical_case()
    solve_and_check(ctx, b, label="C-DISORT spherical case")
    assert b.out[FIELD["flux_down"], 2] == pytest.approx(1.99378710e-01, rel=1e-3)
    assert b.out[FIELD["flux_direct_beam"], 2] == pytest.approx(2.55423483e-02, rel=1e-3)


@pytest.mark.parametrize("nstr", [4, 16])
def test_reference_grazing_configs(ctx, nstr):
    # tests/test_flux_solver.cpp:308-359: 5 layers of 0.1, omega 0.9,
    # isotropic, mu0 0.3, altitudes 100..0 km, radius 6371 km, albedo 0.1.
    L = 5
    pm = np.zeros((1, L, nstr + 1))
    pm[..., 0] = 1.0
    b = uniform_batch(np.full((1, L), 0.1), np.full((1, L), 0.9), pm, nstr,
                      scalars(1, beam_flux=1.0, beam_mu=0.3, albedo=0.1, radius=6371.0), flags=F_SPHERICAL,
                      alt=np.array([[100.0, 80.0, 60.0, 40.0, 20.0, 0.0]]))
    solve_and_check(ctx, b, label=f"grazing spherical nstr={nstr}")


def spherical_batch(n, L, nstr, seed, thermal=False, extra=0, pstride=None, deep=False):
    b = W.random_flux_batch(n, L, nstr, seed=seed, thermal=thermal, beam=True, flags_extra=extra, pstride=pstride)
    if deep:
        b.dtau[:] = np.repeat(np.linspace(0.5, 8.0, L)[None, :], n, axis=0).reshape(-1)
        b.ssa[:] = 0.3
    alt = np.repeat(np.linspace(100.0, 0.0, L + 1)[None, :], n, axis=0)
    sc = b.scal.reshape(n, -1).copy()
    sc[:, 10] = 6371.0
    return FluxBatch(flags=b.flags | F_SPHERICAL, nstr=b.nstr, nlyr=b.nlyr, pstride=b.pstride, dtau=b.dtau,
                     ssa=b.ssa, pmom=b.pmom, scal=sc, temper=b.temper, alt=alt)


@pytest.mark.parametrize("nstr", [4, 6, 8, 12, 16, 32])
def test_random_spherical(ctx, nstr):
    solve_and_check(ctx, spherical_batch(6, 8, nstr, seed=700 + nstr), label=f"spherical nstr={nstr}")


def test_spherical_thermal_delta_m_plus(ctx):
    solve_and_check(ctx, spherical_batch(4, 7, 16, seed=801, thermal=True), label="spherical + thermal")
    solve_and_check(ctx, spherical_batch(4, 7, 16, seed=802, extra=F_DELTA_M_PLUS, pstride=18),
                    label="spherical + delta-M+")


def test_spherical_deep_thermal_column(ctx):
    # Absorption optical depth beyond the cut-off with thermal emission on:
    # layers below the cut keep no slant path (xba = 0 branch).
    solve_and_check(ctx, spherical_batch(3, 10, 16, seed=803, thermal=True, deep=True), label="deep spherical")


def test_spherical_needs_workgroup_variant(ctx):
    from disort_amd._capi import V_LANE, V_WORKGROUP, DeviceFluxBatch, DisortError

    dev = DeviceFluxBatch(ctx, spherical_batch(2, 5, 8, seed=5))
    assert dev.get_variant(8) == V_WORKGROUP
    with pytest.raises(DisortError):
        dev.set_variant(8, V_LANE)
    dev.close()


def test_graph_recaptures_on_new_altitudes(ctx):
    # The captured graph holds the device address of every batch array: a
    # replay with a new altitude buffer (all other pointers unchanged) must
    # solve with the new altitudes, as the eager execute does.
    from disort_amd._capi import DeviceFluxBatch

    b = spherical_batch(4, 6, 8, seed=811)
    dev = DeviceFluxBatch(ctx, b)
    first_alt = dev.cstruct.alt
    dev.execute_graph()
    g1 = dev.download().out.copy()
    dev.execute()
    e1 = dev.download().out.copy()
    np.testing.assert_array_equal(g1, e1)
    alt2 = np.ascontiguousarray(b.alt * 0.5)
    p2 = ctx.alloc(alt2.nbytes)
    dev._bufs.append(p2)
    ctx.h2d(p2, alt2)
    dev.cstruct.alt = p2
    dev.execute_graph()
    g2 = dev.download().out.copy()
    dev.execute()
    e2 = dev.download().out.copy()
    np.testing.assert_array_equal(g2, e2)
    assert not np.array_equal(g2, g1)
    b2 = spherical_batch(4, 6, 8, seed=811)
    b2.alt[:] = alt2
    solve_and_check(ctx, b2, label="spherical, halved altitudes")
    np.testing.assert_allclose(e2, b2.out, rtol=1e-12, atol=1e-15)
    dev.cstruct.alt = first_alt
    dev.close()
