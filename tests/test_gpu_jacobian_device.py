"""Device-resident temperature Jacobians (disort_jac_plan_create /
disort_jac_execute) against the host-array entry point disort_flux_jacobian
(itself checked against the oracle in test_gpu_jacobian.py) and against the
oracle directly on columns of the surrogate workload's shape.

Expansion form (JAC_EXPAND): the expanded solves are the same kernels on
the same inputs, so the raw responses agree bit for bit; dB/dT is evaluated
on the device here and on the host there, so the Jacobians may differ by a
few ulps (bound 1e-13 relative, with the absolute floor relative to the
column's largest entry).  Factored form (the default for nstr <= 16): the
unit-source right-hand sides eliminated on one factorisation per column;
checked against the oracle at the Jacobian bar (1e-9 relative, 1e-12 of the
column's largest entry, extended-precision arbitration)."""
import numpy as np
import pytest

import oracle_lib as ol
from disort_amd import workloads as W
from disort_amd.batch import F_DIFFUSION_BC, F_THERMAL

from disort_amd.batch import F_SPHERICAL
from test_gpu_jacobian import assert_jac_parity
from test_oracle_jacobian import CASES as REF_CASES, thermal_batch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from disort_amd._capi import Context

    return Context(0)


def device_jacobian(ctx, fb, max_expanded=0, streams=1, expand=False):
    from disort_amd._capi import JAC_EXPAND, DeviceFluxBatch, JacPlan

    dev = DeviceFluxBatch(ctx, fb)
    flags = int(fb.flags) | (JAC_EXPAND if expand else 0)
    jp = JacPlan(ctx, fb.n, int(fb.nstr[0]), int(fb.nlyr[0]), int(fb.pstride[0]), flags, max_expanded)
    if streams > 1:
        jp.set_streams(streams)
        jp.set_chunks(streams)
    jac = ctx.alloc(jp.jac_bytes)
    L = int(fb.nlyr[0])
    try:
        jp.execute(dev.cstruct, jac)
        ctx.sync()
        out = np.empty((fb.n, 6, L + 1, L + 2))
        ctx.d2h(out, jac)
        st = jp.status()
        chunk = jp.chunk
    finally:
        ctx.free(jac)
        jp.close()
        dev.close()
    return out, st, chunk


@pytest.mark.parametrize("nstr,nlyr,extra,max_expanded", [(8, 12, 0, 5 * 14), (16, 20, F_DIFFUSION_BC, 0),
                                                          (4, 7, 0, 9)])
def test_device_jacobian_matches_host_entry(ctx, nstr, nlyr, extra, max_expanded):
    """Chunked expansion (a last partial chunk included) gives the host entry
    point's Jacobians."""
    fb = W.random_flux_batch(37, nlyr, nstr, seed=11 + nstr, flags_extra=extra)
    print(f"[jac device] expansion nstr {nstr} nlyr {nlyr} flags {fb.flags} max_expanded {max_expanded}", flush=True)
    got, st, chunk = device_jacobian(ctx, fb, max_expanded, expand=True)
    assert not st.any()
    if max_expanded:
        assert chunk == max(1, max_expanded // (nlyr + 2))
    ref = np.stack(ctx.flux_jacobian(fb))
    floor = 1e-13 * np.abs(ref).max(axis=(1, 2, 3), keepdims=True)
    assert np.all(np.abs(got - ref) <= 1e-13 * np.abs(ref) + floor)


FACTORED = [("n8_l12_chunked", 8, 12, 0, 3 * 14, 1), ("n16_l20_diffusion", 16, 20, F_DIFFUSION_BC, 0, 1),
            ("n4_l7_chunked_streams", 4, 7, 0, 5 * 9, 2), ("n12_l9", 12, 9, 0, 0, 1),
            ("n6_l1", 6, 1, 0, 0, 1), ("n10_l15_diffusion_streams", 10, 15, F_DIFFUSION_BC, 0, 3)]


@pytest.mark.parametrize("case", FACTORED, ids=[c[0] for c in FACTORED])
def test_device_jacobian_factored_matches_oracle(ctx, case):
    """Factored form on random columns of several stream counts (chunks with
    a last partial chunk, side streams, diffusion lower boundary, a single
    layer) against the oracle, and against the expansion form."""
    _, nstr, nlyr, extra, max_expanded, streams = case
    fb = W.random_flux_batch(13, nlyr, nstr, seed=29 + nstr + nlyr, flags_extra=extra)
    print(f"[jac device] factored {case[0]}", flush=True)
    got, st, chunk = device_jacobian(ctx, fb, max_expanded, streams)
    assert not st.any()
    if max_expanded:
        assert chunk == max(1, max_expanded // (nlyr + 2))
    for i in range(fb.n):
        _, ref = ol.flux_jacobian_one(fb, i)
        assert_jac_parity(got[i], ref, f"{case[0]} column {i}", fb, i)
    xp, _, _ = device_jacobian(ctx, fb, 0, 1, expand=True)
    scale = np.abs(xp).max(axis=(1, 2, 3), keepdims=True)
    assert np.abs(got - xp).max() <= 1e-7 * scale.max()


DEVICE_REF = [c for c in REF_CASES if not (c[1].get("flags", 0) & F_SPHERICAL)]


@pytest.mark.parametrize("case", DEVICE_REF, ids=[c[0] for c in DEVICE_REF])
def test_device_jacobian_factored_reference_problems(ctx, case):
    """The reference's Jacobian test problems (tests/test_jacobian.cpp and the
    oracle's extensions: single layer, scattering profiles, thermal plus
    beam, diffusion lower boundary, Delta-M with albedo, top emission and
    isotropic sources, a layer thinner than the Planck-slope cut-off)
    through the factored device plan, against the oracle.  The spherical
    beam case is left out: the device plan does not take that flag."""
    fb = thermal_batch(**case[1])
    print(f"[jac device] factored reference problem {case[0]}", flush=True)
    got, st, _ = device_jacobian(ctx, fb)
    assert not st.any()
    _, ref = ol.flux_jacobian_one(fb, 0)
    assert_jac_parity(got[0], ref, case[0], fb, 0)


def test_device_jacobian_surrogate_columns_match_oracle(ctx):
    """Columns of the surrogate workload (nlyr 60, nstr 16, thermal and beam;
    the beam does not enter the Jacobian) on three side streams, against the
    oracle's linearisation."""
    from disort_amd.surrogate import c5_columns

    fb = c5_columns(5, seed=404)
    print("[jac device] surrogate columns, 3 streams", flush=True)
    got, st, _ = device_jacobian(ctx, fb, max_expanded=2 * 62, streams=3)
    assert not st.any()
    for i in range(fb.n):
        _, ref = ol.flux_jacobian_one(fb, i)
        assert_jac_parity(got[i], ref, f"column {i}", fb, i)


def test_device_jacobian_plan_rejects(ctx):
    from disort_amd._capi import DisortError, JacPlan
    from disort_amd.batch import F_SPHERICAL

    with pytest.raises(DisortError, match="thermal"):
        JacPlan(ctx, 4, 8, 10, 9, 0)
    with pytest.raises(DisortError):
        JacPlan(ctx, 4, 8, 10, 9, F_THERMAL | F_SPHERICAL)
    with pytest.raises(DisortError):
        JacPlan(ctx, 4, 12 + 1, 10, 14, F_THERMAL)
    with pytest.raises(DisortError):
        JacPlan(ctx, 4, 8, 10, 8, F_THERMAL)  # pstride below nstr + 1
