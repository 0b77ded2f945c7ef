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
