"""CPU oracle of the flux-surrogate network (test infrastructure only: the
tests and bench.py's checks import it; the library never does).

The surrogate has no counterpart in the reference, so this oracle restates
the network the GPU kernels compute (disort_amd/csrc/mlp_kernels.hip):
hidden layers z = h W + b, h' = GELU(z) with the tanh form of GELU, a linear
last layer, the masked mean-squared-error loss and Adam.  With `bf16=True`
it rounds to bfloat16 (round to nearest even) at the points the GPU stores
BF16 -- the input, the weights, every pre-activation and activation and the
back-propagated gradients -- and accumulates in float64, so the GPU's FP32
accumulation differs from it only by accumulation rounding.

Pins (tests/test_mlp_oracle.py): the float64 backward pass against central
finite differences of the float64 forward pass; GELU and its derivative
against closed-form values; Adam against its textbook recurrence on scalars;
the column features against hand-computed layouts and the level features
against closed forms (optical depth sums, Beer's law, half the optical depth
at single-scattering albedo 0.5); the emission features of the Jacobian
surrogate against the same sums and the invariant that the transmissions
from the top and to the surface multiply to the whole column's.
"""
from __future__ import annotations

import numpy as np

K0 = 0.7978845608028654  # sqrt(2 / pi)
K1 = 0.044715


def bf16_round(x):
    """Round float32 values to the nearest bfloat16 (ties to even), as float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out).reshape(a.shape)


def bf16_bits(x):
    """bfloat16 bit patterns (uint16) of float32 values, ties to even."""
    return (bf16_round(x).view(np.uint32) >> 16).astype(np.uint16)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(K0 * (z + K1 * z ** 3)))


def gelu_grad(z):
    th = np.tanh(K0 * (z + K1 * z ** 3))
    return 0.5 * (1.0 + th) + 0.5 * z * (1.0 - th * th) * K0 * (1.0 + 3.0 * K1 * z * z)


def layout(widths):
    """Offsets (W_l, b_l) in the packed parameter vector, as the library packs them."""
    offs, o = [], 0
    for l in range(len(widths) - 1):
        wo = o
        o += widths[l] * widths[l + 1]
        offs.append((wo, o))
        o += widths[l + 1]
    return offs, o


def unpack(params, widths):
    offs, _ = layout(widths)
    Ws, bs = [], []
    for l, (wo, bo) in enumerate(offs):
        Ws.append(params[wo:wo + widths[l] * widths[l + 1]].reshape(widths[l], widths[l + 1]))
        bs.append(params[bo:bo + widths[l + 1]])
    return Ws, bs


def forward(params, widths, x, bf16=True):
    """Returns y and the cache of (h_in, z) per layer."""
    Ws, bs = unpack(np.asarray(params, np.float64), widths)
    rnd = (lambda a: bf16_round(a).astype(np.float64)) if bf16 else (lambda a: np.asarray(a, np.float64))
    h = rnd(x)
    cache = []
    L = len(Ws)
    for l in range(L):
        W = rnd(Ws[l]) if bf16 else Ws[l]
        z = h @ W + bs[l]
        if l + 1 < L:
            cache.append((h, rnd(z)))
            h = rnd(gelu(z))
        else:
            cache.append((h, None))
            return z, cache
    raise AssertionError


def backward(params, widths, cache, dy, bf16=True):
    """Packed parameter gradients for output gradient dy."""
    Ws, _ = unpack(np.asarray(params, np.float64), widths)
    rnd = (lambda a: bf16_round(a).astype(np.float64)) if bf16 else (lambda a: np.asarray(a, np.float64))
    offs, n = layout(widths)
    g = np.zeros(n)
    dz = rnd(dy)
    for l in range(len(Ws) - 1, -1, -1):
        h, _ = cache[l]
        wo, bo = offs[l]
        g[wo:wo + Ws[l].size] = (h.T @ dz).reshape(-1)
        g[bo:bo + Ws[l].shape[1]] = dz.sum(axis=0)
        if l > 0:
            W = rnd(Ws[l]) if bf16 else Ws[l]
            z_prev = cache[l - 1][1]
            dz = rnd((dz @ W.T) * gelu_grad(z_prev))
    return g


def mse(y, t, mask, n_valid=None):
    """Loss and dL/dy of sum mask (y - t)^2 / (B * n_valid)."""
    B, N = y.shape
    nv = n_valid or N
    d = (y - t) * mask[None, :]
    return float((d * d).sum() / (B * nv)), 2.0 * d / (B * nv)


def adam(p, g, m, v, lr, b1, b2, eps, t):
    """One Adam step (bias-corrected), float32 like the GPU kernel."""
    p, g, m, v = (np.asarray(a, np.float32).copy() for a in (p, g, m, v))
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1 = np.float32(1 - b1 ** t)
    c2 = np.float32(1 - b2 ** t)
    p = p - lr * (m / c1) / (np.sqrt(v / c2) + eps)
    return p, m, v


def features(dtau, ssa, pmom, temper, mu0, albedo):
    """Raw surrogate features of uniform columns: [n, 4L+2] (see surrogate_kernels.hip)."""
    n, L = dtau.shape
    f = np.empty((n, 4 * L + 2))
    f[:, :L] = np.log10(dtau)
    f[:, L:2 * L] = ssa
    f[:, 2 * L:3 * L] = pmom.reshape(n, L, -1)[:, :, 1]
    f[:, 3 * L:4 * L] = 0.5 * (temper[:, :-1] + temper[:, 1:])
    f[:, 4 * L] = mu0
    f[:, 4 * L + 1] = albedo
    return f


def level_features(dtau, ssa, mu0):
    """Level features of the surrogate [n, 3(L+1)]: log10(tau_v + 1e-3),
    exp(-tau_v / mu0) and log10(sum_{l<v} (1 - omega_l) dtau_l + 1e-3) at the
    L+1 levels, tau_v the optical depth from the top to level v."""
    n, L = dtau.shape
    f = np.empty((n, 3 * (L + 1)))
    tau = np.zeros(n)
    tab = np.zeros(n)
    for v in range(L + 1):
        if v > 0:
            tau = tau + dtau[:, v - 1]
            tab = tab + (1.0 - ssa[:, v - 1]) * dtau[:, v - 1]
        f[:, v] = np.log10(tau + 1e-3)
        f[:, L + 1 + v] = np.exp(-tau / mu0)
        f[:, 2 * (L + 1) + v] = np.log10(tab + 1e-3)
    return f


def thermal_features(b_levels, b_bottom, albedo, mu0f0):
    """Thermal-source features of the surrogate [n, L+2]: the level Planck
    values and the surface emission (1 - albedo) B(T_bottom), in units of the
    beam mu0 F0 (b_levels [n, L+1], b_bottom [n]: band Planck values)."""
    u = np.where(mu0f0 > 1e-300, mu0f0, 1.0)
    return np.concatenate([b_levels / u[:, None], ((1.0 - albedo) * b_bottom / u)[:, None]], axis=1)


def emission_features(dtau, ssa, g, albedo, mubar=0.6):
    """Emission features of the Jacobian surrogate [n, 3L + 1 + 6(L+1)]:
    log10(dtau), omega, g per layer, the surface albedo, then per level v
    log10(tau_v + 1e-3), log10(a_v + 1e-3), exp(-tau_v / mubar),
    exp(-a_v / mubar), exp(-(tau_L - tau_v) / mubar), exp(-(a_L - a_v) / mubar),
    a_v = sum_{l<v} (1 - omega_l) dtau_l the absorption depth."""
    n, L = dtau.shape
    tau = np.zeros((n, L + 1))
    ab = np.zeros((n, L + 1))
    for v in range(1, L + 1):
        tau[:, v] = tau[:, v - 1] + dtau[:, v - 1]
        ab[:, v] = ab[:, v - 1] + (1.0 - ssa[:, v - 1]) * dtau[:, v - 1]
    lev = [np.log10(tau + 1e-3), np.log10(ab + 1e-3), np.exp(-tau / mubar), np.exp(-ab / mubar),
           np.exp(-(tau[:, L:] - tau) / mubar), np.exp(-(ab[:, L:] - ab) / mubar)]
    return np.concatenate([np.log10(dtau), ssa, g, np.asarray(albedo).reshape(n, 1)] + lev, axis=1)


SQ3 = 1.7320508075688772  # sqrt(3)


def twostream(dtau, ssa, g, mu0, f0, albedo, b_lev=None, b_sfc=None):
    """Two-stream diffuse fluxes of the surrogate's baseline, in DisORT's
    output convention: returns (up, down) [n, L+1], down excluding the
    unscaled direct beam.

    Delta-Eddington scaling (f = g^2), quadrature coefficients
    gamma1 = sqrt(3)(2 - w(1+g))/2, gamma2 = sqrt(3) w (1-g)/2,
    gamma3 = (1 - sqrt(3) g mu0)/2, gamma4 = 1 - gamma3, and per layer the
    exact solution of
      dU/dt = gamma1 U - gamma2 D - w F0 gamma3 e^{-tau/mu0} - (gamma1-gamma2) pi B(t)
      dD/dt = gamma2 U - gamma1 D + w F0 gamma4 e^{-tau/mu0} + (gamma1-gamma2) pi B(t)
    (t the scaled depth in the layer, B linear in t): homogeneous modes
    e^{lambda (t - dtau)} (1, Gamma) and e^{-lambda t} (Gamma, 1), the
    exponential beam particular solution (mu0 moved by 0.1% near
    resonance) and the linear thermal one, pi B(t) +- pi B'/(gamma1+gamma2).
    Boundary conditions: no diffuse flux from the top; Lambertian surface
    U = a (D + mu0 F0 e^{-tau_L/mu0}) + (1 - a) pi B_sfc.  The 2L constants
    are solved with continuity at the interfaces by Gaussian elimination
    with partial pivoting on the 3 x 2 panel of each layer (a plain dense
    solve here)."""
    n, L = dtau.shape
    w0 = np.minimum(ssa, 1.0 - 1e-6)
    f = g * g
    dt = (1.0 - w0 * f) * dtau
    w = (1.0 - f) * w0 / (1.0 - w0 * f)
    gs = g / (1.0 + g)
    g1 = SQ3 * (2.0 - w * (1.0 + gs)) / 2.0
    g2 = SQ3 * w * (1.0 - gs) / 2.0
    g3 = (1.0 - SQ3 * gs * mu0[:, None]) / 2.0
    g4 = 1.0 - g3
    lam = np.sqrt((g1 - g2) * (g1 + g2))
    Gam = g2 / (g1 + lam)
    E = np.exp(-lam * dt)
    T = np.concatenate([np.zeros((n, 1)), np.cumsum(dt, axis=1)], axis=1)  # scaled level depths
    Tu = np.concatenate([np.zeros((n, 1)), np.cumsum(dtau, axis=1)], axis=1)  # unscaled
    thermal = b_lev is not None
    up_p0, up_p1, dn_p0, dn_p1 = (np.zeros((n, L)) for _ in range(4))
    for l in range(L):
        m = mu0.copy()
        near = np.abs(1.0 - (m * lam[:, l]) ** 2) < 1e-4
        m[near] *= 0.999
        det = 1.0 / m ** 2 - lam[:, l] ** 2
        a = f0 * w[:, l] * (g3[:, l] * (1.0 / m - g1[:, l]) - g2[:, l] * g4[:, l]) / det
        b = f0 * w[:, l] * (-g4[:, l] * (g1[:, l] + 1.0 / m) - g2[:, l] * g3[:, l]) / det
        e0, e1 = np.exp(-T[:, l] / m), np.exp(-T[:, l + 1] / m)
        up_p0[:, l], up_p1[:, l] = a * e0, a * e1
        dn_p0[:, l], dn_p1[:, l] = b * e0, b * e1
        if thermal:
            bt, bb = np.pi * b_lev[:, l], np.pi * b_lev[:, l + 1]
            s = (bb - bt) / dt[:, l] / (g1[:, l] + g2[:, l])
            up_p0[:, l] += bt + s
            up_p1[:, l] += bb + s
            dn_p0[:, l] += bt - s
            dn_p1[:, l] += bb - s
    up = np.empty((n, L + 1))
    dn = np.empty((n, L + 1))
    for i in range(n):
        M = np.zeros((2 * L, 2 * L))
        r = np.zeros(2 * L)
        M[0, 0], M[0, 1] = Gam[i, 0] * E[i, 0], 1.0
        r[0] = -dn_p0[i, 0]
        for l in range(L - 1):
            k = 1 + 2 * l
            M[k, 2 * l:2 * l + 2] = [1.0, Gam[i, l] * E[i, l]]
            M[k, 2 * l + 2:2 * l + 4] = [-E[i, l + 1], -Gam[i, l + 1]]
            r[k] = up_p0[i, l + 1] - up_p1[i, l]
            M[k + 1, 2 * l:2 * l + 2] = [Gam[i, l], E[i, l]]
            M[k + 1, 2 * l + 2:2 * l + 4] = [-Gam[i, l + 1] * E[i, l + 1], -1.0]
            r[k + 1] = dn_p0[i, l + 1] - dn_p1[i, l]
        al = albedo[i]
        M[-1, -2:] = [1.0 - al * Gam[i, -1], (Gam[i, -1] - al) * E[i, -1]]
        r[-1] = al * mu0[i] * f0[i] * np.exp(-T[i, L] / mu0[i]) - up_p1[i, -1] + al * dn_p1[i, -1]
        if thermal:
            r[-1] += (1.0 - al) * np.pi * b_sfc[i]
        x = np.linalg.solve(M, r)
        A, B = x[0::2], x[1::2]
        up[i, :L] = A * E[i] + B * Gam[i] + up_p0[i]
        dn[i, :L] = A * Gam[i] * E[i] + B + dn_p0[i]
        up[i, L] = A[-1] + B[-1] * Gam[i, -1] * E[i, -1] + up_p1[i, -1]
        dn[i, L] = A[-1] * Gam[i, -1] + B[-1] * E[i, -1] + dn_p1[i, -1]
    dn = dn + mu0[:, None] * f0[:, None] * (np.exp(-T / mu0[:, None]) - np.exp(-Tu / mu0[:, None]))
    return up, dn
