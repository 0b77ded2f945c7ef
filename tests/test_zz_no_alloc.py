"""disort_flux_execute_static (the flux solve that allocates nothing: every
intermediate lives in the library's static device array) against the
existing device-resident entry point disort_flux_execute on the same device
inputs, through the C ABI.  Two calls in a row, the second with a smaller
batch (fewer instances, fewer layers, other stream count), so that the
second call runs over what the first left in the shared array.

Bar: the flux path's parity bar |static - execute| <= 1e-9 |execute| + 1e-12
on every output (the two run the same kernels on the same inputs, so the
results are in fact identical)."""
import ctypes as C

import numpy as np
import pytest

from disort_amd import workloads as W
from disort_amd.batch import F_SPHERICAL, FluxBatch

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-9, 1e-12


@pytest.fixture(scope="module")
def ctx():
    from disort_amd._capi import Context

    return Context(0)


def run_both(ctx, batch, label):
    from disort_amd._capi import DeviceFluxBatch, _check, lib

    n, N, L, P = batch.n, int(batch.nstr[0]), int(batch.nlyr[0]), int(batch.pstride[0])
    print(f"[no-alloc] {label}: n={n} nstr={N} nlyr={L} pstride={P} flags={batch.flags}", flush=True)
    dev = DeviceFluxBatch(ctx, batch)
    dev.execute()
    ref = dev.download().out.copy()
    assert (dev.host.status == 0).all()
    need = lib().disort_flux_static_bytes(n, N, L)
    assert 0 < need <= lib().disort_static_scratch_bytes(), (need, label)
    out_p = ctx.alloc(batch.out.nbytes)
    st_p = ctx.alloc(4 * n)
    try:
        s = type(dev.cstruct).from_buffer_copy(dev.cstruct)
        s.out = out_p
        s.status = st_p
        # The static path ignores the offset and shape arrays of the batch.
        s.nstr = s.nlyr = s.pstride = s.lay_off = s.mom_off = None
        _check(ctx.ptr, lib().disort_flux_execute_static(ctx.ptr, C.byref(s), N, L, P))
        ctx.sync()
        got = np.empty_like(batch.out)
        ctx.d2h(got, out_p)
        st = np.full(n, -1, np.int32)
        ctx.d2h(st, st_p)
    finally:
        ctx.free(out_p)
        ctx.free(st_p)
        dev.close()
    assert (st == 0).all(), st
    assert np.isfinite(got).all()
    bad = np.abs(got - ref) > RTOL * np.abs(ref) + ATOL
    assert not bad.any(), f"{label}: {np.argwhere(bad)[:5].tolist()}"
    return got, ref


def test_static_matches_execute_two_calls(ctx):
    # First call: 600 C2 bands (N=16, L=60, thermal), second: a smaller
    # batch of 40 instances with 8 streams and 25 layers.
    g1, r1 = run_both(ctx, W.c2(600), "C2 600 bands")
    g2, r2 = run_both(ctx, W.random_flux_batch(40, 25, 8, seed=3, thermal=True, beam=True), "40 x (8, 25)")
    assert np.array_equal(g1, r1) and np.array_equal(g2, r2)


def test_static_workgroup_kernels(ctx):
    # 32 streams (workgroup kernels) and a pseudo-spherical beam.
    run_both(ctx, W.random_flux_batch(12, 10, 32, seed=5, beam=True), "12 x (32, 10)")
    b = W.random_flux_batch(6, 8, 8, seed=6, thermal=False, beam=True)
    alt = np.repeat(np.linspace(100.0, 0.0, 9)[None, :], 6, axis=0)
    sc = b.scal.reshape(6, -1).copy()
    sc[:, 10] = 6371.0
    sb = FluxBatch(flags=b.flags | F_SPHERICAL, nstr=b.nstr, nlyr=b.nlyr, pstride=b.pstride, dtau=b.dtau,
                   ssa=b.ssa, pmom=b.pmom, scal=sc, alt=alt)
    run_both(ctx, sb, "spherical 6 x (8, 8)")


def test_static_rejects_oversized_batch(ctx):
    from disort_amd._capi import DisortInvalid, _check, lib
    from disort_amd.batch import CFluxBatch

    n = 100000
    assert lib().disort_flux_static_bytes(n, 16, 60) > lib().disort_static_scratch_bytes()
    s = CFluxBatch()
    s.n = n
    s.flags = 1
    dummy = ctx.alloc(8)
    try:
        s.dtau = s.ssa = s.pmom = s.temper = s.scal = s.out = dummy
        with pytest.raises(DisortInvalid, match="bytes of static scratch; the maximum is"):
            _check(ctx.ptr, lib().disort_flux_execute_static(ctx.ptr, C.byref(s), 16, 60, 17))
    finally:
        ctx.free(dummy)
