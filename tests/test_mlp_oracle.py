"""Pins of the CPU oracle of the surrogate network (oracle/mlp_oracle.py)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -3.14159265, 0.0, 65504.0], np.float32)
    r = mo.bf16_round(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties to even -> 1.
    assert r[1] == 1.0
    # 1 + 1.5 * 2^-8 rounds up to 1 + 2^-7.
    assert r[2] == np.float32(1.0078125)
    assert r[3] == np.float32(1.0078125)
    assert r[4] == np.float32(-3.140625)
    assert r[5] == 0.0
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)
    b = mo.bf16_bits(np.array([1.0, -2.0], np.float32))
    assert list(b) == [0x3F80, 0xC000]


def test_gelu_closed_forms():
    assert mo.gelu(0.0) == 0.0
    z = np.linspace(-6, 6, 241)
    # GELU(z) - GELU(-z) = z for the tanh form.
    np.testing.assert_allclose(mo.gelu(z) - mo.gelu(-z), z, atol=1e-12)
    h = 1e-6
    fd = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(mo.gelu_grad(z), fd, atol=1e-8)
    assert mo.gelu_grad(0.0) == 0.5


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    widths = [6, 5, 7, 4]
    _, n = mo.layout(widths)
    p = rng.normal(0, 0.5, n)
    x = rng.normal(0, 1, (3, 6))
    t = rng.normal(0, 1, (3, 4))
    mask = np.array([1.0, 1.0, 0.0, 1.0])

    def loss(q):
        y, _ = mo.forward(q, widths, x, bf16=False)
        return mo.mse(y, t, mask, 3)[0]

    y, cache = mo.forward(p, widths, x, bf16=False)
    _, dy = mo.mse(y, t, mask, 3)
    g = mo.backward(p, widths, cache, dy, bf16=False)
    h = 1e-6
    fd = np.array([(loss(p + h * e) - loss(p - h * e)) / (2 * h) for e in np.eye(n)])
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-9)


def test_adam_recurrence():
    p, m, v = np.float32([1.0]), np.float32([0.0]), np.float32([0.0])
    g = np.float32([0.5])
    p1, m1, v1 = mo.adam(p, g, m, v, 1e-2, 0.9, 0.999, 1e-8, 1)
    # First step: m_hat = g, v_hat = g^2, so the step is lr * sign(g).
    assert p1[0] == pytest.approx(1.0 - 1e-2, abs=1e-7)
    assert m1[0] == pytest.approx(0.05) and v1[0] == pytest.approx(0.00025)
    p2, _, _ = mo.adam(p1, g, m1, v1, 1e-2, 0.9, 0.999, 1e-8, 2)
    assert p2[0] == pytest.approx(1.0 - 2e-2, abs=2e-7)


def test_features_layout():
    n, L = 2, 3
    dtau = np.array([[1e-4, 1.0, 10.0], [0.1, 0.2, 0.3]])
    ssa = np.full((n, L), 0.5)
    pmom = np.zeros((n, L, 5))
    pmom[:, :, 1] = 0.7
    temper = np.tile(np.array([200.0, 210.0, 230.0, 260.0]), (n, 1))
    f = mo.features(dtau, ssa, pmom, temper, np.array([0.3, 0.4]), np.array([0.1, 0.2]))
    assert f.shape == (2, 14)
    np.testing.assert_allclose(f[0, :3], [-4.0, 0.0, 1.0])
    np.testing.assert_allclose(f[0, 9:12], [205.0, 220.0, 245.0])
    assert f[1, 12] == 0.4 and f[1, 13] == 0.2


def test_level_features_closed_forms():
    dtau = np.array([[1e-4, 1.0, 10.0], [0.1, 0.2, 0.3]])
    ssa = np.array([[0.5, 0.5, 0.5], [0.0, 1.0, 0.25]])
    mu0 = np.array([0.5, 1.0])
    f = mo.level_features(dtau, ssa, mu0)
    assert f.shape == (2, 12)
    # optical depth of the four levels
    np.testing.assert_allclose(10.0 ** f[0, :4] - 1e-3, [0.0, 1e-4, 1.0001, 11.0001], rtol=1e-12, atol=1e-15)
    # Beer's law of the direct beam at mu0 = 0.5 and 1
    np.testing.assert_allclose(f[0, 4:8], np.exp(-2.0 * np.array([0.0, 1e-4, 1.0001, 11.0001])), rtol=1e-12)
    np.testing.assert_allclose(f[1, 4:8], np.exp(-np.array([0.0, 0.1, 0.3, 0.6])), rtol=1e-12)
    # absorption depth: half the optical depth at omega = 0.5; 0.1, 0.1, 0.325 in the second column
    np.testing.assert_allclose(10.0 ** f[0, 8:12] - 1e-3, [0.0, 5e-5, 0.50005, 5.50005], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(10.0 ** f[1, 8:12] - 1e-3, [0.0, 0.1, 0.1, 0.325], rtol=1e-12, atol=1e-15)


def test_emission_features_closed_forms():
    dtau = np.array([[1e-4, 1.0, 10.0], [0.1, 0.2, 0.3]])
    ssa = np.array([[0.5, 0.5, 0.5], [0.0, 1.0, 0.25]])
    g = np.full((2, 3), 0.3)
    f = mo.emission_features(dtau, ssa, g, np.array([0.1, 0.2]))
    assert f.shape == (2, 3 * 3 + 1 + 6 * 4)
    np.testing.assert_allclose(f[0, :3], [-4.0, 0.0, 1.0])
    assert f[1, 9] == 0.2 and (f[:, 6:9] == 0.3).all()
    b = 10
    tau0 = np.array([0.0, 1e-4, 1.0001, 11.0001])
    np.testing.assert_allclose(10.0 ** f[0, b:b + 4] - 1e-3, tau0, rtol=1e-12, atol=1e-15)
    # absorption depth: half the optical depth at omega = 0.5; 0.1, 0.1, 0.325 in the second column
    np.testing.assert_allclose(10.0 ** f[0, b + 4:b + 8] - 1e-3, tau0 / 2, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(10.0 ** f[1, b + 4:b + 8] - 1e-3, [0.0, 0.1, 0.1, 0.325], rtol=1e-12, atol=1e-15)
    # diffuse transmissions: exp(-x / 0.6) from the top, to the surface
    np.testing.assert_allclose(f[0, b + 8:b + 12], np.exp(-tau0 / 0.6), rtol=1e-12)
    np.testing.assert_allclose(f[1, b + 12:b + 16], np.exp(-np.array([0.0, 0.1, 0.1, 0.325]) / 0.6), rtol=1e-12)
    # top and surface transmissions multiply to the whole column's, at every level
    for q in (0, 1):
        top, bot = f[:, b + (8 + 4 * q):b + (12 + 4 * q)], f[:, b + (16 + 4 * q):b + (20 + 4 * q)]
        np.testing.assert_allclose(top * bot, top[:, -1:] * np.ones((1, 4)), rtol=1e-12)
    assert (f[:, b + 19] == 1.0).all() and (f[:, b + 23] == 1.0).all() and (f[:, b + 8] == 1.0).all()


def _ts_inputs(n, L, seed):
    r = np.random.default_rng(seed)
    return (10.0 ** r.uniform(-3, 1, (n, L)), r.uniform(0, 0.99, (n, L)), r.uniform(0, 0.9, (n, L)),
            r.uniform(0.1, 1, n), np.full(n, np.pi), r.uniform(0, 0.3, n))


def test_twostream_closed_forms():
    # Isothermal absorbing column over a black surface at the same temperature:
    # U = pi B at every level, no diffuse flux enters at the top.
    B = np.full((1, 9), 2.0)
    up, dn = mo.twostream(np.full((1, 8), 0.7), np.zeros((1, 8)), np.zeros((1, 8)), np.array([0.6]),
                          np.array([0.0]), np.array([0.0]), B, np.array([2.0]))
    np.testing.assert_allclose(up, np.pi * 2.0, rtol=1e-13)
    assert dn[0, 0] == 0.0 and np.all(np.diff(dn[0]) > 0) and dn[0, -1] < np.pi * 2.0
    # Pure absorber lit by the beam over a black surface: no diffuse light at all.
    dtau, _, g, mu0, f0, _ = _ts_inputs(3, 6, 1)
    up, dn = mo.twostream(dtau, np.zeros_like(dtau), g, mu0, f0, np.zeros(3))
    assert np.all(up == 0.0) and np.all(np.abs(dn) < 1e-15)
    # Conservative scattering (omega clamped at 1 - 1e-6) over a white surface:
    # the column reflects the incident beam and the net flux vanishes.
    up, dn = mo.twostream(np.full((1, 8), 2.0), np.ones((1, 8)), np.full((1, 8), 0.5), np.array([0.6]),
                          np.array([1.0]), np.array([1.0]))
    direct = 0.6 * np.exp(-np.concatenate([[0.0], np.cumsum(np.full(8, 2.0))]) / 0.6)
    assert abs(up[0, 0] / 0.6 - 1.0) < 1e-4
    assert np.all(np.abs(dn[0] + direct - up[0]) < 1e-4)


def test_twostream_layer_splitting_is_exact():
    """The per-layer solution is exact: splitting every layer into two equal
    halves (the Planck value at the split the linear interpolation) leaves
    the fluxes at the original levels unchanged, with the beam and the
    thermal source on."""
    n, L = 4, 7
    dtau, ssa, g, mu0, f0, alb = _ts_inputs(n, L, 2)
    r = np.random.default_rng(3)
    b = r.uniform(1.0, 3.0, (n, L + 1))
    bs = r.uniform(1.0, 3.0, n)
    up, dn = mo.twostream(dtau, ssa, g, mu0, f0, alb, b, bs)
    rep = lambda a: np.repeat(a, 2, axis=1)  # noqa: E731
    b2 = np.empty((n, 2 * L + 1))
    b2[:, 0::2] = b
    b2[:, 1::2] = 0.5 * (b[:, :-1] + b[:, 1:])
    up2, dn2 = mo.twostream(rep(dtau) / 2, rep(ssa), rep(g), mu0, f0, alb, b2, bs)
    np.testing.assert_allclose(up2[:, 0::2], up, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(dn2[:, 0::2], dn, rtol=1e-9, atol=1e-12)


def test_twostream_accuracy_class_against_disort():
    """On the surrogate workload's columns the two-stream net flux is within a
    few percent of the incident beam of the 16-stream DISORT oracle (a sign
    or coupling mistake in a source term moves it by tens of percent)."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests"))
    import oracle_lib as ol
    from disort_amd import _capi
    from disort_amd import surrogate as S

    n, L = 200, 60
    hb = S.c5_columns(n, seed=3)
    ol.flux_batch(hb, nthreads=4)
    ref = S.host_fluxes(hb)
    sc = hb.scal.reshape(n, 12)
    t = hb.temper.reshape(n, L + 1)
    bl = np.stack([_capi.planck_band(1000.0, 1001.0, t[i]) for i in range(n)])
    bs = _capi.planck_band(1000.0, 1001.0, sc[:, 5])
    up, dn = mo.twostream(hb.dtau.reshape(n, L), hb.ssa.reshape(n, L), hb.pmom.reshape(n, L, -1)[:, :, 1],
                          sc[:, 1], sc[:, 0], sc[:, 7], bl, bs)
    u = sc[:, 1] * sc[:, 0]
    err = np.abs((dn - up) - (ref[:, 1] - ref[:, 0])).max(axis=1) / u
    assert np.median(err) < 0.02 and np.quantile(err, 0.9) < 0.04, np.quantile(err, [0.5, 0.9])
    errs = np.abs((dn + up) - (ref[:, 1] + ref[:, 0])).max(axis=1) / u
    assert np.median(errs) < 0.03 and np.quantile(errs, 0.9) < 0.06, np.quantile(errs, [0.5, 0.9])
