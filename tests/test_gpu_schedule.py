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
        # the bucket with the most work takes the variant of the joint
        # (variant x schedule) search, the others their isolated choice
        want = rep["dominant_variant"] if n == rep["dominant_nstr"] else got[n]["chosen"]
        assert dev.get_variant(n) == codes[want]
    assert rep["schedule"] in ("serial", "concurrent")
    assert rep["dominant_nstr"] == max(got, key=lambda k: min(v for v in (got[k]["lane_ms"], got[k]["lane_pair_ms"],
                                                                        got[k]["workgroup_ms"]) if v > 0))
    assert len(rep["schedules"]) == 6 * sum(got[rep["dominant_nstr"]][f"{v}_ms"] > 0
                                            for v in ("lane", "lane_pair", "workgroup"))
    out = run(dev, "execute_graph")
    assert_parity(out, ref, label="autotuned graph")
    dev.close()


def test_chunked_stream_pool_agrees(ctx, ragged):
    """Buckets split into chunks and spread over a pool of side streams give
    the serial results bit for bit (every instance runs the same kernels),
    eagerly and as graph replays; schedule arguments are range-checked."""
    from disort_amd._capi import DeviceFluxBatch, DisortError

    dev = DeviceFluxBatch(ctx, ragged)
    serial = run(dev, "execute")
    for streams, chunks in [(2, 2), (3, 3), (4, 2), (1, 3), (3, 8)]:
        dev.set_streams(streams)
        for n in dev.buckets():
            dev.set_chunks(n, chunks)
        np.testing.assert_array_equal(run(dev, "execute"), serial, err_msg=f"streams {streams} chunks {chunks}")
        for rep in range(2):
            np.testing.assert_array_equal(run(dev, "execute_graph"), serial,
                                          err_msg=f"graph streams {streams} chunks {chunks} replay {rep}")
    for bad in (0, 5):
        with pytest.raises(DisortError):
            dev.set_streams(bad)
    with pytest.raises(DisortError):
        dev.set_chunks(16, 9)
    with pytest.raises(DisortError):
        dev.set_chunks(12, 2)  # no bucket with 12 streams
    dev.close()


def test_chunked_uniform_batch_matches_oracle(ctx):
    """A uniform 16-stream batch (the C2 shape at reduced size) split into
    three chunks on three streams agrees with the oracle."""
    from disort_amd._capi import DeviceFluxBatch

    b = W.c2(64, nlyr=60, nstr=16)
    ref = oracle_out(b)
    dev = DeviceFluxBatch(ctx, b)
    dev.set_streams(3)
    dev.set_chunks(16, 3)
    assert_parity(run(dev, "execute"), ref, label="c2 subset, 3 chunks on 3 streams")
    dev.close()
