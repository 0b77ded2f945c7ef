"""Bucket schedules of a ragged batch: concurrent buckets, hipGraph replay
and the autotuner give the same numbers as the serial eager solve (which is
itself checked against the oracle)."""
import numpy as np
import pytest

from disort_amd import workloads as W

from test_gpu_flux import assert_parity, oracle_out

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from disort_amd._capi import Context

    return Context(0)


@pytest.fixture(scope="module")
def ragged():
    return W.c4(60, first=2000)


def run(dev, how):
    getattr(dev, how)()
    dev.ctx.sync()
    return dev.download().out.copy()


def test_serial_concurrent_graph_agree(ctx, ragged):
    from disort_amd._capi import DeviceFluxBatch

    ref = oracle_out(ragged)
    dev = DeviceFluxBatch(ctx, ragged)
    serial = run(dev, "execute")
    assert_parity(serial, ref, label="serial eager")
    dev.set_concurrent(True)
    conc = run(dev, "execute")
    np.testing.assert_array_equal(conc, serial)
    # Graph replays, with eager solves of another batch in between; every
    # replay must reproduce the eager result.
    other = DeviceFluxBatch(ctx, W.random_flux_batch(5, 9, 16, seed=9))
    for rep in range(3):
        g = run(dev, "execute_graph")
        np.testing.assert_array_equal(g, serial, err_msg=f"graph replay {rep}")
        other.execute()
        ctx.sync()
    dev.set_concurrent(False)
    np.testing.assert_array_equal(run(dev, "execute_graph"), serial)
    other.close()
    dev.close()


def test_autotune_keeps_results(ctx, ragged):
    from disort_amd._capi import V_LANE, V_LANE_PAIR, V_WORKGROUP, DeviceFluxBatch

    ref = oracle_out(ragged)
    dev = DeviceFluxBatch(ctx, ragged)
    rep = dev.autotune(reps=2)
    got = {b["nstr"]: b for b in rep["buckets"]}
    assert sorted(got) == [4, 8, 16, 32, 64]
    assert got[4]["chosen"] in ("lane", "lane_pair", "workgroup") and got[64]["chosen"] == "workgroup"
    codes = {"lane": V_LANE, "lane_pair": V_LANE_PAIR, "workgroup": V_WORKGROUP}
    for n in (8, 16):
        assert got[n]["lane_ms"] > 0 and got[n]["lane_pair_ms"] > 0 and got[n]["workgroup_ms"] > 0
        assert dev.get_variant(n) == codes[got[n]["chosen"]]
    assert rep["schedule"] in ("serial", "concurrent")
    out = run(dev, "execute_graph")
    assert_parity(out, ref, label="autotuned graph")
    dev.close()
