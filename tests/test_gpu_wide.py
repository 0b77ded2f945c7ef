"""Workgroup kernels (N = 32, 64, and the alternative schedule at N = 8, 16)
against the oracle, and the ragged C4-style batch of mixed stream counts.
Tolerance: |gpu - oracle| <= 1e-9 |oracle| + 1e-12 per output."""
import numpy as np
import pytest

import oracle_lib as ol
from disort_amd import workloads as W
from disort_amd.batch import F_DELTA_M_PLUS, F_DIFFUSION_BC, F_THERMAL

from test_gpu_flux import assert_parity, oracle_out, solve_and_check

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from disort_amd._capi import Context

    return Context(0)


@pytest.mark.parametrize("nstr", [32, 64])
@pytest.mark.parametrize("thermal,beam", [(True, True), (False, True), (True, False)])
def test_wide_random_batches(ctx, nstr, thermal, beam):
    b = W.random_flux_batch(6, 7, nstr, seed=300 + nstr, thermal=thermal, beam=beam)
    solve_and_check(ctx, b, label=f"nstr={nstr} thermal={thermal} beam={beam}")


@pytest.mark.parametrize("nstr", [32, 64])
def test_wide_delta_m_plus_and_diffusion(ctx, nstr):
    b = W.random_flux_batch(3, 5, nstr, seed=400 + nstr, flags_extra=F_DELTA_M_PLUS, pstride=nstr + 2)
    solve_and_check(ctx, b, label=f"delta-M+ nstr={nstr}")
    b = W.random_flux_batch(3, 5, nstr, seed=500 + nstr, flags_extra=F_DIFFUSION_BC)
    solve_and_check(ctx, b, label=f"diffusion nstr={nstr}")


@pytest.mark.parametrize("nstr", [32, 64])
def test_wide_layer_cutoff(ctx, nstr):
    # Beam only, thick absorbing layers: the absorption cut-off fires after
    # layer 3 and the layer tasks below it do not run.
    L = 8
    b = W.random_flux_batch(4, L, nstr, seed=600 + nstr, thermal=False)
    b.dtau[:] = 4.0
    b.ssa[:] = 0.2
    ref = solve_and_check(ctx, b, label=f"lyrcut nstr={nstr}")
    assert (ref[2].reshape(4, L + 1)[:, -1] == 0).all()  # levels below the cut are zero


@pytest.mark.parametrize("nstr", [8, 16])
def test_workgroup_variant_at_small_n(ctx, nstr):
    from disort_amd._capi import V_LANE, V_LANE_PAIR, V_WORKGROUP, DeviceFluxBatch

    b = W.random_flux_batch(10, 12, nstr, seed=600 + nstr)
    ref = oracle_out(b)
    dev = DeviceFluxBatch(ctx, b)
    assert dev.get_variant(nstr) == (V_LANE_PAIR if nstr <= 12 else V_LANE)
    dev.set_variant(nstr, V_WORKGROUP)
    assert dev.get_variant(nstr) == V_WORKGROUP
    dev.execute()
    out = dev.download()
    assert (out.status == 0).all()
    assert_parity(out.out, ref, label=f"workgroup kernels nstr={nstr}")
    dev.close()


def test_variant_rejected_where_no_kernels(ctx):
    from disort_amd._capi import V_LANE, V_WORKGROUP, DeviceFluxBatch, DisortError

    b = W.random_flux_batch(2, 4, 32, seed=1)
    dev = DeviceFluxBatch(ctx, b)
    assert dev.get_variant(32) == V_WORKGROUP
    with pytest.raises(DisortError):
        dev.set_variant(32, V_LANE)  # no register kernels at 32 streams
    with pytest.raises(DisortError):
        dev.set_variant(8, V_LANE)  # no bucket with 8 streams
    dev.close()


def test_c4_ragged_subset(ctx):
    # 40 instances of the ragged workload: stream counts 4..64, 10..120 layers,
    # beam only (absorption cut-off active in deep columns).
    b = W.c4(40, first=1000)
    assert set(int(x) for x in b.nstr) == {4, 8, 16, 32, 64}
    solve_and_check(ctx, b, label="c4 subset")


@pytest.mark.parametrize("nstr", [4, 6, 8, 10, 12, 14, 16])
@pytest.mark.parametrize("flags", [0, F_DELTA_M_PLUS, F_DIFFUSION_BC])
def test_register_variants_agree_with_oracle(ctx, nstr, flags):
    """Both register eigen kernels (one lane per layer, two lanes per layer)
    against the oracle, with beam and thermal sources."""
    from disort_amd._capi import V_LANE, V_LANE_PAIR, DeviceFluxBatch

    P = nstr + 2 if flags & F_DELTA_M_PLUS else None
    b = W.random_flux_batch(9, 11, nstr, seed=700 + nstr + flags, flags_extra=flags, pstride=P)
    if flags & F_DELTA_M_PLUS:
        b.pmom.reshape(-1, nstr + 2)[:, -1] *= 0.9
    ref = oracle_out(b)
    for v in (V_LANE, V_LANE_PAIR):
        dev = DeviceFluxBatch(ctx, b)
        dev.set_variant(nstr, v)
        dev.execute()
        out = dev.download()
        assert (out.status == 0).all()
        assert_parity(out.out, ref, label=f"variant {v} nstr={nstr} flags={flags}")
        dev.close()
