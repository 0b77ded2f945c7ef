"""Surrogate network kernels on the GPU against the CPU oracle
(oracle/mlp_oracle.py, which emulates the BF16 storage points), the column
kernels of the surrogate, and a small end-to-end train / gate / serve run.

Tolerances: BF16 storage (8 significant bits) with FP32 accumulation against
an oracle that rounds at the same points and accumulates in float64: outputs
and gradients agree to a relative Frobenius error of 2e-2 (a flipped BF16
rounding of an intermediate moves a result by ~2^-8 of its size)."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as mo  # noqa: E402

sys.path.insert(0, os.path.join(ROOT, "tests"))
import oracle_lib as ol  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from disort_amd._capi import Context

    return Context(0)


def bf16_values(bits):
    """float32 values of BF16 bit patterns."""
    return (np.asarray(bits, np.uint32) << 16).view(np.float32)


def assert_features_close(got_bits, want_bits):
    """GPU features (BF16 bits) against the oracle's: the feature kernels take
    logarithms and exponentials in single precision, so a standardised value
    may land one BF16 step away, or change sign right at zero."""
    g, w = bf16_values(got_bits), bf16_values(want_bits)
    assert np.all(np.abs(g - w) <= 2.0 ** -7 * np.abs(w) + 1e-5), float(np.max(np.abs(g - w)))


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def make_net(ctx, widths, B, seed):
    from disort_amd import surrogate as S

    net = S.MLP(ctx, widths, max_batch=B)
    r = np.random.default_rng(seed)
    p = np.zeros(net.n_param, np.float32)
    for l, (wo, bo) in enumerate(net.layout()):
        p[wo:bo] = r.normal(0, 1 / np.sqrt(widths[l]), bo - wo)
        p[bo:bo + widths[l + 1]] = r.normal(0, 0.1, widths[l + 1])
    net.set_params(p)
    x = r.normal(0, 1, (B, widths[0])).astype(np.float32)
    return net, p, x


@pytest.mark.parametrize("widths", [(64, 128, 64), (256, 512, 512, 192)])
def test_forward_backward_vs_oracle(ctx, widths):
    from disort_amd import surrogate as S

    B = 128
    net, p, x = make_net(ctx, widths, B, seed=len(widths))
    dx = S.DevArray.of(ctx, mo.bf16_bits(x))
    dyv = S.DevArray(ctx, B * widths[-1] * 4)
    net.forward(dx.ptr, B, dyv.ptr)
    y = dyv.get(np.float32, (B, widths[-1]))
    yo, cache = mo.forward(p, widths, x)
    assert rel(y, yo) < 2e-2
    g_out = np.random.default_rng(1).normal(0, 1, (B, widths[-1])).astype(np.float32)
    ddy = S.DevArray.of(ctx, mo.bf16_bits(g_out))
    net.backward(ddy.ptr)
    g = net.get_grads()
    go = mo.backward(p, widths, cache, g_out)
    for l, (wo, bo) in enumerate(net.layout()):
        assert rel(g[wo:bo], go[wo:bo]) < 2e-2, f"dW layer {l}"
        assert rel(g[bo:bo + widths[l + 1]], go[bo:bo + widths[l + 1]]) < 2e-2, f"db layer {l}"
    net.close()


def test_train_step_is_loss_gradient_and_adam(ctx):
    from disort_amd import surrogate as S

    widths, B = (64, 128, 64), 64
    net, p, x = make_net(ctx, widths, B, seed=5)
    t = np.random.default_rng(2).normal(0, 1, (B, widths[-1])).astype(np.float32)
    mask = np.ones(widths[-1], np.float32)
    mask[50:] = 0.0
    dx, dt, dm = S.DevArray.of(ctx, mo.bf16_bits(x)), S.DevArray.of(ctx, t), S.DevArray.of(ctx, mask)
    loss = S.DevArray.of(ctx, np.zeros(1, np.float32))
    net.train_step(dx.ptr, dt.ptr, dm.ptr, B, 50, 1e-3, 1, loss.ptr)
    p1 = net.get_params()
    yo, cache = mo.forward(p, widths, x)
    lo, dyo = mo.mse(yo, t, mask, 50)
    assert float(loss.get(np.float32, 1)[0]) == pytest.approx(lo, rel=2e-2)
    go = mo.backward(p, widths, cache, dyo)
    po, _, _ = mo.adam(p, go, np.zeros_like(p), np.zeros_like(p), 1e-3, 0.9, 0.999, 1e-8, 1)
    # First Adam step moves every parameter with a clear gradient by lr*sign(g).
    clear = np.abs(go) > 1e-3 * np.abs(go).max()
    np.testing.assert_allclose(p1[clear], po[clear], atol=2e-5)
    net.close()


def test_training_reduces_loss(ctx):
    from disort_amd import surrogate as S

    widths, n = (64, 128, 128, 64), 4096
    r = np.random.default_rng(7)
    x = r.normal(0, 1, (n, widths[0])).astype(np.float32)
    A = r.normal(0, 1 / 8, (widths[0], widths[-1]))
    t = (np.tanh(x @ A) + 0.1 * (x[:, :64] ** 2) @ A).astype(np.float32)
    net = S.MLP(ctx, widths, max_batch=256)
    net.set_params(net.init_params(3, widths[0], widths[-1]))
    hist = S.train(ctx, net, S.DevArray.of(ctx, mo.bf16_bits(x)), S.DevArray.of(ctx, t), n, epochs=8, batch=256,
                   lr=2e-3)
    assert hist[-1] < 0.5 * hist[0], hist
    assert sum(b < a for a, b in zip(hist, hist[1:])) >= 0.8 * (len(hist) - 1), hist
    net.close()


def test_column_kernels(ctx):
    from disort_amd import surrogate as S
    from disort_amd._capi import Context  # noqa: F401

    l = S._mlp_lib()
    n = 100
    hb = S.c5_columns(n, seed=11)
    dc = S.DeviceColumns(ctx, hb)
    dc.solve()
    out = dc.dev.download()
    feats = S.raw_features(hb)
    tg = S.host_targets(out)
    norm = S.Normalizer.fit(feats, tg, margin=1e-5)  # the kernel's logs and exps are single precision
    dn = norm.device(ctx)
    n_pad = 128
    x = S.DevArray(ctx, n_pad * 256 * 2)
    ood = S.DevArray(ctx, n * 4)
    assert l.disort_surrogate_features(ctx.ptr, n, n_pad, 60, 17, C.byref(dc.cstruct), dn["f_shift"].ptr,
                                       dn["f_scale"].ptr, dn["f_lo"].ptr, dn["f_hi"].ptr, 256, x.ptr, ood.ptr, None) == 0
    xs = x.get(np.uint16, (n_pad, 256))
    want = mo.bf16_bits(((feats[:, :242].astype(np.float32) - norm.f_shift[:242]) * norm.f_scale[:242]).astype(np.float32))
    # Same float32 standardisation; the logarithms in single precision.
    assert_features_close(xs[:n, :242], want)
    assert (xs[n:] == 0).all() and (xs[:, 242:] == 0).all()
    assert (ood.get(np.int32, n) == 0).all()  # the columns define the range
    # With the network's input width the 3 x 61 level features follow (oracle restatement).
    w0, nf = S.WIDTHS[0], S.N_FEAT
    dtau, ssa = hb.dtau.reshape(n, 60), hb.ssa.reshape(n, 60)
    mu0 = hb.scal.reshape(n, 12)[:, 1]
    sc12 = hb.scal.reshape(n, 12)
    temper = hb.temper.reshape(n, 61)
    b_lev = np.array([[ol.planck(sc12[i, 8], sc12[i, 9], tv) for tv in temper[i]] for i in range(n)])
    b_bot = np.array([ol.planck(sc12[i, 8], sc12[i, 9], sc12[i, 5]) for i in range(n)])
    full = np.concatenate([mo.features(dtau, ssa, hb.pmom.reshape(n, 60, -1), temper, mu0, sc12[:, 7]),
                           mo.level_features(dtau, ssa, mu0),
                           mo.thermal_features(b_lev, b_bot, sc12[:, 7], mu0 * sc12[:, 0])], axis=1)
    np.testing.assert_allclose(S.raw_features(hb), full, rtol=1e-12, atol=1e-15)
    x2 = S.DevArray(ctx, n_pad * w0 * 2)
    assert l.disort_surrogate_features(ctx.ptr, n, n_pad, 60, 17, C.byref(dc.cstruct), dn["f_shift"].ptr,
                                       dn["f_scale"].ptr, dn["f_lo"].ptr, dn["f_hi"].ptr, w0, x2.ptr, ood.ptr, None) == 0
    xs2 = x2.get(np.uint16, (n_pad, w0))
    want2 = mo.bf16_bits(((full.astype(np.float32) - norm.f_shift[:nf]) * norm.f_scale[:nf]).astype(np.float32))
    assert_features_close(xs2[:n, :nf], want2)
    assert (xs2[n:] == 0).all() and (xs2[:, nf:] == 0).all()
    assert (ood.get(np.int32, n) == 0).all()
    t = S.DevArray(ctx, n_pad * 192 * 4)
    assert l.disort_surrogate_targets(ctx.ptr, n, n_pad, 61, C.byref(dc.cstruct), dn["t_shift"].ptr,
                                      dn["t_scale"].ptr, 192, t.ptr, None) == 0
    ts = t.get(np.float32, (n_pad, 192))
    np.testing.assert_allclose(ts[:n, :183], (tg - norm.t_shift[:183]) * norm.t_scale[:183], rtol=1e-5, atol=1e-5)
    # Decoding the standardised targets returns the solver's fluxes.
    flux, spread = S.DevArray(ctx, n * 3 * 61 * 8), S.DevArray(ctx, n * 4)
    assert l.disort_surrogate_decode(ctx.ptr, n, 60, 192, (C.c_void_p * 1)(t.ptr), 1, dn["t_shift"].ptr,
                                     dn["t_scale"].ptr,
                                     C.byref(dc.cstruct), 1.0, 1e-3, flux.ptr, spread.ptr, None) == 0
    f = flux.get(np.float64, (n, 3, 61))
    ref = S.host_fluxes(out)
    np.testing.assert_allclose(f[:, :2], ref[:, :2], rtol=1e-5, atol=1e-6 * np.abs(ref).max())
    np.testing.assert_allclose(f[:, 2], ref[:, 2], rtol=1e-12, atol=1e-15)  # closed-form direct beam
    # Gather + plain solve + scatter of a subset reproduces those columns.
    idx = np.array([3, 17, 42, 99], np.int32)
    sub = S._SubBatch(ctx, idx.size)
    didx = S.DevArray.of(ctx, idx)
    assert l.disort_surrogate_gather(ctx.ptr, idx.size, didx.ptr, 60, 17, C.byref(dc.cstruct), C.byref(sub.cstruct)) == 0
    sub.solve()
    f2 = S.DevArray.of(ctx, np.zeros((n, 3, 61)))
    assert l.disort_surrogate_scatter(ctx.ptr, idx.size, didx.ptr, 61, sub.cstruct.out, sub.cstruct.total_levels,
                                      f2.ptr) == 0
    g = f2.get(np.float64, (n, 3, 61))
    np.testing.assert_allclose(g[idx], ref[idx], rtol=1e-9, atol=1e-12)
    sub.close()
    dc.close()


def test_end_to_end_gate_small(ctx):
    """Train two members on 64k solver columns, calibrate the gate, serve 4k
    fresh columns: gated-out and sampled columns carry plain-DISORT fluxes,
    accepted columns respect the calibrated violation rate, and the trained
    network beats the method without it (predicting the training mean) by a
    clear margin."""
    from disort_amd import surrogate as S

    x, t, n, norm = S.build_dataset(ctx, 65536, seed=1)
    members = []
    for k in range(2):
        m = S.MLP(ctx, S.WIDTHS, max_batch=4096)
        m.set_params(m.init_params(100 + k, S.N_FEAT, S.N_OUT))
        members.append(m)
    cal = S.DeviceColumns(ctx, S.c5_columns(2048, seed=77))
    cal.solve()
    cal_ref = S.host_fluxes(cal.dev.download())
    # The method without the network: every column gets the training mean.
    mean_net = S.MLP(ctx, S.WIDTHS, max_batch=4096)
    mean_net.set_params(np.zeros(mean_net.n_param, np.float32))
    baseline = S.GatedSurrogate(ctx, [mean_net, mean_net], norm).calibrate(cal, cal_ref)
    for k, m in enumerate(members):
        S.train(ctx, m, x, t, n, epochs=10, batch=1024, lr=2e-3, seed=k)
    gs = S.GatedSurrogate(ctx, members, norm, alpha=0.05, tol=5e-2, metric="beam")
    after = gs.calibrate(cal, cal_ref)
    assert after["err_beam_median"] < 0.6 * baseline["err_beam_median"], (baseline, after)
    fresh = S.c5_columns(4096, seed=78)
    dc = S.DeviceColumns(ctx, fresh)
    dc.solve()
    ref = S.host_fluxes(dc.dev.download())
    flux, info = gs.serve(dc, sample_frac=0.02, seed=3)
    f = flux.get(np.float64, (dc.n, 3, 61))
    fl = info["flags"].get(np.int32, dc.n)
    summ = S.flags_summary(fl)
    routed = np.setdiff1d(np.arange(dc.n), summ["accepted_idx"])
    assert routed.size == info["n_plain"]
    # flags: routed = sampled or failing the gate; about 2% sampled
    assert ((fl & 1) != 0).tolist() == (((fl & 2) != 0) | ((fl & 4) == 0)).tolist()
    assert 0.01 < summ["n_sampled"] / dc.n < 0.03
    np.testing.assert_allclose(f[routed], ref[routed], rtol=1e-9, atol=1e-12)
    sc = fresh.scal.reshape(-1, 12)
    err = S.column_error(f, ref, sc[:, 1] * sc[:, 0], gs.metric)
    assert (err[routed] <= 1e-8).all()
    if summ["accepted_idx"].size > 100:
        assert (err[summ["accepted_idx"]] > gs.tol).mean() < 0.15
    # The monitor's count: sampled columns that pass the gate.
    assert info["monitor_sampled"] == int(((fl & 6) == 6).sum())

    # Online monitor: a gate that passes everything with a bound no column
    # meets trips after the first call; the next call routes every column to
    # plain DISORT, which still measures the surrogate on the sampled ones.
    strict = S.GatedSurrogate(ctx, members, norm, tau=np.inf, alpha=0.01, tol=1e-9, metric="beam")
    strict.monitor.min_samples = 50
    f1, i1 = strict.serve(dc, sample_frac=0.05, seed=4)
    assert i1["monitor_sampled"] >= 50 and i1["monitor_violations"] == i1["monitor_sampled"]
    assert not i1["tripped"] and strict.monitor.tripped
    f2, i2 = strict.serve(dc, sample_frac=0.05, seed=5)
    assert i2["tripped"] and i2["n_plain"] == dc.n
    np.testing.assert_allclose(f2.get(np.float64, (dc.n, 3, 61)), ref, rtol=1e-9, atol=1e-12)
    for a in (f1, f2, i1["flags"], i2["flags"], flux, info["flags"]):
        a.free()
    strict.close()
    gs.close()
    for m in members + [mean_net]:
        m.close()


@pytest.mark.parametrize("thermal", [True, False])
def test_twostream_kernel_matches_oracle(ctx, thermal):
    """The two-stream baseline kernel against the oracle's restatement
    (mlp_oracle.twostream, pinned by closed forms, exact layer splitting and
    its accuracy class against DISORT): 1e-10 of the incident beam."""
    from disort_amd import surrogate as S
    from disort_amd import _capi

    n, L = 300, 60
    hb = S.c5_columns(n, seed=41)
    if not thermal:
        hb.flags = 0
    print(f"[twostream] {n} columns thermal={thermal}", flush=True)
    dc = S.DeviceColumns(ctx, hb)
    work = S.DevArray(ctx, 12 * n * L * 8)
    ts = S.DevArray(ctx, n * S.TS_ROW * 8)
    l = S._mlp_lib()
    assert l.disort_surrogate_twostream(ctx.ptr, n, L, 17, C.byref(dc.cstruct), work.ptr, ts.ptr) == 0
    got = S.ts_fluxes(ts, n)
    sc = hb.scal.reshape(n, 12)
    bl = bs = None
    if thermal:
        t = hb.temper.reshape(n, L + 1)
        bl = np.stack([_capi.planck_band(1000.0, 1001.0, t[i]) for i in range(n)])
        bs = _capi.planck_band(1000.0, 1001.0, sc[:, 5])
    up, dn = mo.twostream(hb.dtau.reshape(n, L), hb.ssa.reshape(n, L), hb.pmom.reshape(n, L, -1)[:, :, 1],
                          sc[:, 1], sc[:, 0], sc[:, 7], bl, bs)
    u = (sc[:, 1] * sc[:, 0])[:, None]
    assert np.abs(got[:, 0] - up).max() / 1.0 <= 1e-10 * u.max()
    assert (np.abs(got[:, 0] - up) / u).max() <= 1e-10
    assert (np.abs(got[:, 1] - dn) / u).max() <= 1e-10
    if thermal:  # the level and surface Planck values shared with the feature kernel
        pl = ts.get(np.float64, (n, S.TS_ROW))[:, 2 * (L + 1):3 * (L + 1) + 1]
        np.testing.assert_allclose(pl, np.concatenate([bl, bs[:, None]], axis=1), rtol=1e-13)
    np.testing.assert_array_equal(ts.get(np.float64, (n, S.TS_ROW))[:, 3 * (L + 1) + 1:],
                                  hb.pmom.reshape(n, L, -1)[:, :, 1])
    for a in (work, ts):
        a.free()
    dc.close()


def test_residual_column_kernels(ctx):
    """With the two-stream baseline: the appended features are its net and
    sum in beam units, the targets are the solver's residuals over it, and
    decoding the standardised targets adds it back (the solver's fluxes)."""
    from disort_amd import surrogate as S

    l = S._mlp_lib()
    n, n_pad = 100, 128
    hb = S.c5_columns(n, seed=12)
    dc = S.DeviceColumns(ctx, hb)
    dc.solve()
    out = dc.dev.download()
    tsd = dc.twostream()
    tsh = S.ts_fluxes(tsd, n)
    feats = S.raw_features(hb, S.N_FEAT_TS, tsh)
    tg = S.host_targets(out, tsh)
    w0 = S.WIDTHS_TS[0]
    norm = S.Normalizer.fit(feats, tg, margin=1e-5, w0=w0)  # single-precision logs and exps
    dn = norm.device(ctx)
    x = S.DevArray(ctx, n_pad * w0 * 2)
    ood = S.DevArray(ctx, n * 4)
    assert l.disort_surrogate_features(ctx.ptr, n, n_pad, 60, 17, C.byref(dc.cstruct), dn["f_shift"].ptr,
                                       dn["f_scale"].ptr, dn["f_lo"].ptr, dn["f_hi"].ptr, w0, x.ptr, ood.ptr,
                                       tsd.ptr) == 0
    xs = x.get(np.uint16, (n_pad, w0))
    nf = S.N_FEAT_TS
    want = mo.bf16_bits(((feats.astype(np.float32) - norm.f_shift[:nf]) * norm.f_scale[:nf]).astype(np.float32))
    assert_features_close(xs[:n, :nf], want)
    assert (xs[n:] == 0).all() and (xs[:, nf:] == 0).all()
    assert (ood.get(np.int32, n) == 0).all()
    t = S.DevArray(ctx, n_pad * 192 * 4)
    assert l.disort_surrogate_targets(ctx.ptr, n, n_pad, 61, C.byref(dc.cstruct), dn["t_shift"].ptr,
                                      dn["t_scale"].ptr, 192, t.ptr, tsd.ptr) == 0
    ts_ = t.get(np.float32, (n_pad, 192))
    np.testing.assert_allclose(ts_[:n, :183], (tg - norm.t_shift[:183]) * norm.t_scale[:183], rtol=1e-5, atol=1e-5)
    flux, spread = S.DevArray(ctx, n * 3 * 61 * 8), S.DevArray(ctx, n * 4)
    assert l.disort_surrogate_decode(ctx.ptr, n, 60, 192, (C.c_void_p * 2)(t.ptr, t.ptr), 2, dn["t_shift"].ptr,
                                     dn["t_scale"].ptr, C.byref(dc.cstruct), 0.0, 1.0, flux.ptr, spread.ptr,
                                     tsd.ptr) == 0
    f = flux.get(np.float64, (n, 3, 61))
    ref = S.host_fluxes(out)
    u = (hb.scal.reshape(n, 12)[:, 1] * hb.scal.reshape(n, 12)[:, 0])[:, None]
    # float32 targets of residuals 10-40x smaller than the fluxes: decoded to ~1e-7 of the beam
    assert (np.abs(f[:, :2] - ref[:, :2]) / u[:, :, None]).max() < 2e-6
    assert (spread.get(np.float32, n) == 0).all()
    # Three members t, t + d, t - d on the net output of level 5: the mean is
    # the single member's, the spread twice their standard deviation.
    d = 0.3
    tp, tm = ts_.copy(), ts_.copy()
    tp[:n, 5] += np.float32(d)
    tm[:n, 5] -= np.float32(d)
    dp, dm = S.DevArray.of(ctx, tp), S.DevArray.of(ctx, tm)
    flux3 = S.DevArray(ctx, n * 3 * 61 * 8)
    assert l.disort_surrogate_decode(ctx.ptr, n, 60, 192, (C.c_void_p * 3)(t.ptr, dp.ptr, dm.ptr), 3,
                                     dn["t_shift"].ptr, dn["t_scale"].ptr, C.byref(dc.cstruct), 0.0, 1.0, flux3.ptr,
                                     spread.ptr, tsd.ptr) == 0
    # (t + d and t - d are rounded to float32, so their mean is t to float32 rounding)
    np.testing.assert_allclose(flux3.get(np.float64, (n, 3, 61)), f, rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(spread.get(np.float32, n), 2.0 * np.sqrt(2.0 / 3.0) * d / norm.t_scale[5], rtol=1e-4)
    for a in (x, ood, t, flux, spread, dp, dm, flux3):
        a.free()
    dc.close()
