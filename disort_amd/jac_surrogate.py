"""Learned temperature Jacobians with an error gate (no counterpart in the
reference; the plain path is DisortFluxSolver's compute_temperature_jacobian,
src/FluxSolver.cpp:1294-1497, run on the GPU by disort_jac_execute).

What is learned: the Jacobians of the top-of-atmosphere upward flux and of
the surface downward flux with respect to the L+2 temperatures of a column
(the L+1 level temperatures and the surface temperature): the rows the
radiative kernels of retrievals and feedback analyses use.  With the optical
properties fixed, dF/dT_k = R_k dB/dT(T_k), where the response R_k = dF/dB_k
to a unit Planck value at degree of freedom k depends on the optical
properties only.  The network (547 emission features and the 124 responses
of the two-stream model -> 512 x 4 GELU -> 124) learns R / pi from the
optical properties as the residual over those two-stream responses; dB/dT
is evaluated exactly on the device, so the temperature profile never enters
the network.

The gate is the flux surrogate's (disort_amd/surrogate.py): two members and
their spread, the out-of-distribution check on the features, a sampled
re-solve routed on the device, and the online error monitor, with the plain
GPU Jacobian as fallback and judge.  Error of a column: max over both rows
and every degree of freedom of |dF/dT_k - ref_k| / (pi dB/dT(T_k)), the
Jacobian error as a fraction of the black-body flux derivative at that
temperature (equivalently the absolute error of R_k / pi).

Workload: the flux surrogate's C5 columns (nlyr=60, nstr=16, thermal, one
band 1000-1001 cm^-1); the beam does not enter the Jacobians.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from .batch import F_THERMAL, NSCAL, S_ALBEDO, S_TEMP_BOT, S_WN_HI, S_WN_LO, CFluxBatch, FluxBatch
from .surrogate import (NLEV, NLYR, NSTR, DevArray, GateMonitor, MLP, Normalizer, _ck, _mlp_lib, _pad64, c5_columns,
                        train)

D = NLYR + 2                          # degrees of freedom: L+1 levels and the surface
N_EFEAT = 3 * NLYR + 1 + 6 * NLEV      # 547 emission features
N_JOUT = 2 * D                         # 124 outputs: top flux_up and surface flux_down responses
N_EFEAT_TS = N_EFEAT + N_JOUT          # 671: + the two-stream responses (residual form)
JWIDTHS = (704, 512, 512, 512, 512, 128)
TS_ROW = 3 * NLEV + 1 + NLYR           # rows of the two-stream input array (disort_surrogate_twostream)
WORK_COLS = 65536                      # columns per two-stream response pass (work: 18 L doubles each)
MU_BAR = 0.6
JTOL = 2e-3
# Target scale floor (R / pi units, residual form): the outputs' residuals
# all have standard deviations of order 1e-2 or less, so the loss weighs
# their absolute errors nearly alike, as the gate's metric does.
T_FLOOR = 0.01


def _jlib():
    l = _mlp_lib()
    if not getattr(l, "_jac_bound", False):
        vp = C.c_void_p
        l.disort_jac_surrogate_features.argtypes = [vp, C.c_int64, C.c_int64, C.c_int, C.c_int, vp, vp, vp, vp, vp,
                                                    C.c_int, vp, vp, vp]
        l.disort_jac_surrogate_targets.argtypes = [vp, C.c_int64, C.c_int64, C.c_int, vp, vp, vp, vp, C.c_int, vp, vp]
        l.disort_jac_surrogate_decode.argtypes = [vp, C.c_int64, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp]
        l.disort_jac_twostream.argtypes = [vp, C.c_int64, C.c_int, C.c_int, vp, vp, C.c_int64, vp, vp]
        l.disort_jac_surrogate_monitor.argtypes = [vp, C.c_int64, vp, vp, C.c_int, vp, vp, vp, C.c_double, vp]
        l.disort_jac_surrogate_scatter.argtypes = [vp, C.c_int64, vp, C.c_int, vp, vp]
        l._jac_bound = True
    return l


# ---------------------------------------------------------------------------
# Host statistics
# ---------------------------------------------------------------------------
def emission_features(b: FluxBatch, resp: np.ndarray | None = None) -> np.ndarray:
    """Host copy of the device emission-feature map (statistics only)
    [n, 547], with the two-stream responses [n, 2, D] appended / pi when
    given ([n, 671])."""
    n, L = b.n, NLYR
    dtau = b.dtau.reshape(n, L)
    ssa = b.ssa.reshape(n, L)
    f = np.empty((n, N_EFEAT if resp is None else N_EFEAT_TS))
    if resp is not None:
        f[:, N_EFEAT:] = resp.reshape(n, -1) / np.pi
    f[:, :L] = np.log10(dtau)
    f[:, L:2 * L] = ssa
    f[:, 2 * L:3 * L] = b.pmom.reshape(n, L, -1)[:, :, 1]
    f[:, 3 * L] = b.scal.reshape(n, NSCAL)[:, S_ALBEDO]
    z = np.zeros((n, 1))
    tau = np.concatenate([z, np.cumsum(dtau, axis=1)], axis=1)
    tab = np.concatenate([z, np.cumsum((1.0 - ssa) * dtau, axis=1)], axis=1)
    blocks = [np.log10(tau + 1e-3), np.log10(tab + 1e-3), np.exp(-tau / MU_BAR), np.exp(-tab / MU_BAR),
              np.exp(-(tau[:, -1:] - tau) / MU_BAR), np.exp(-(tab[:, -1:] - tab) / MU_BAR)]
    base = 3 * L + 1
    for q, blk in enumerate(blocks):
        f[:, base + q * NLEV:base + (q + 1) * NLEV] = blk
    return f


def dbdt_dofs(b: FluxBatch) -> np.ndarray:
    """dB/dT at the D degrees of freedom of every column [n, D] (level temperatures, then the surface)."""
    n = b.n
    s = b.scal.reshape(n, NSCAL)
    t = np.concatenate([b.temper.reshape(n, NLEV), s[:, S_TEMP_BOT][:, None]], axis=1)
    out = np.empty_like(t)
    bands = np.stack([s[:, S_WN_LO], s[:, S_WN_HI]], axis=1)
    for lo, hi in np.unique(bands, axis=0):
        sel = (bands[:, 0] == lo) & (bands[:, 1] == hi)
        out[sel] = _capi.planck_band_deriv(lo, hi, t[sel]).reshape(-1, D)
    return out


def jac_rows(jac: np.ndarray) -> np.ndarray:
    """[n, 2, D] rows (top flux_up, surface flux_down) of full Jacobians [n, 6, L+1, D]."""
    return np.stack([jac[:, 0, 0, :], jac[:, 1, NLYR, :]], axis=1)


def jacobian_error(rows: np.ndarray, ref_rows: np.ndarray, dbdt: np.ndarray) -> np.ndarray:
    """Per column max over both rows and every k of |dF/dT_k - ref_k| / (pi dB/dT(T_k))."""
    return np.max(np.abs(rows - ref_rows) / (np.pi * dbdt[:, None, :]), axis=(1, 2))


# ---------------------------------------------------------------------------
# Device columns and the plain Jacobian
# ---------------------------------------------------------------------------
class JacColumns:
    """n uniform surrogate-workload columns resident on the device, with
    their plain Jacobian plan (built on first use) and Jacobian buffer."""

    def __init__(self, ctx, b: FluxBatch):
        self.ctx, self.n, self.host = ctx, b.n, b
        L, P = NLYR, NSTR + 1
        self.arrs = [DevArray.of(ctx, a) for a in (b.dtau, b.ssa, b.pmom, b.temper, b.scal)]
        s = CFluxBatch()
        s.n, s.flags = b.n, F_THERMAL
        s.dtau, s.ssa, s.pmom, s.temper, s.scal = (a.ptr for a in self.arrs)
        s.total_levels = b.n * (L + 1)
        self.cstruct = s
        self.P = P
        self.plan = None
        self.jac = None

    def plain(self):
        """Enqueue the plain GPU Jacobian of every column into self.jac [n, 6, L+1, D]."""
        if self.plan is None:
            self.plan = _capi.JacPlan(self.ctx, self.n, NSTR, NLYR, self.P, F_THERMAL)
            self.jac = DevArray(self.ctx, self.plan.jac_bytes)
        self.plan.execute(self.cstruct, self.jac.ptr)

    def responses(self) -> DevArray:
        """Enqueue the two-stream responses [n, 2, D] (top U and surface D for a
        unit Planck value at each degree of freedom); the buffers are kept."""
        if getattr(self, "_resp", None) is None:
            wc = min(self.n, WORK_COLS)
            self._rwork = DevArray(self.ctx, 18 * NLYR * wc * 8)
            self._rts = DevArray(self.ctx, self.n * TS_ROW * 8)
            self._resp = DevArray(self.ctx, self.n * 2 * D * 8)
        _ck(self.ctx, _jlib().disort_jac_twostream(self.ctx.ptr, self.n, NLYR, self.P, C.byref(self.cstruct),
                                                   self._rwork.ptr, min(self.n, WORK_COLS), self._rts.ptr,
                                                   self._resp.ptr))
        return self._resp

    def plain_rows_host(self) -> np.ndarray:
        self.ctx.sync()
        return jac_rows(self.jac.get(np.float64, (self.n, 6, NLEV, D)))

    def close(self):
        if self.plan is not None:
            self.plan.close()
            self.plan = None
        for a in self.arrs + ([self.jac] if self.jac is not None else []):
            a.free()
        for name in ("_rwork", "_rts", "_resp"):
            a = getattr(self, name, None)
            if a is not None:
                a.free()
            setattr(self, name, None)
        self.arrs, self.jac = [], None


# ---------------------------------------------------------------------------
# Dataset and training
# ---------------------------------------------------------------------------
def _identity(ctx, w0, wl):
    z = lambda w, v: DevArray.of(ctx, np.full(w, v, np.float32))  # noqa: E731
    return z(w0, 0.0), z(w0, 1.0), z(w0, -np.inf), z(w0, np.inf), z(wl, 0.0), z(wl, 1.0)


def raw_targets(ctx, cols: JacColumns, resp: DevArray | None = None) -> np.ndarray:
    """[n, N_JOUT] responses R / pi of the solved columns (cols.plain() done),
    less the two-stream responses / pi when resp is given, via the device
    target kernel."""
    l = _jlib()
    wl = JWIDTHS[-1]
    ids = _identity(ctx, JWIDTHS[0], wl)
    t = DevArray(ctx, cols.n * wl * 4)
    _ck(ctx, l.disort_jac_surrogate_targets(ctx.ptr, cols.n, cols.n, NLYR, cols.jac.ptr, C.byref(cols.cstruct),
                                            ids[4].ptr, ids[5].ptr, wl, t.ptr, resp.ptr if resp else None))
    out = t.get(np.float32, (cols.n, wl))[:, :N_JOUT].astype(np.float64)
    for a in ids + (t,):
        a.free()
    return out


def build_jac_dataset(ctx, n: int, seed: int, norm: Normalizer | None = None, chunk: int = 8192):
    """Solve the plain Jacobians of n fresh columns on the GPU and keep
    standardised BF16 emission features and FP32 response targets on the
    device.  Returns (x, t, n, norm)."""
    l = _jlib()
    w0, wl = JWIDTHS[0], JWIDTHS[-1]
    n_pad = _pad64(n)
    x = DevArray(ctx, n_pad * w0 * 2)
    t = DevArray(ctx, n_pad * wl * 4)
    dn = None
    fit = norm is None
    lo = hi = None
    done = k = 0
    while done < n:
        m = min(chunk, n - done)
        hb = c5_columns(m, seed=seed * 1000 + k)
        cols = JacColumns(ctx, hb)
        cols.plain()
        resp = cols.responses()
        if fit:
            # Standardisation from the first chunk; the feature range (the
            # gate's out-of-distribution box) from every training column.
            f = emission_features(hb, resp.get(np.float64, (m, 2, D)))
            if norm is None:
                norm = Normalizer.fit(f, raw_targets(ctx, cols, resp), margin=0.0, w0=w0, wl=wl, t_floor=T_FLOOR)
            lo = f.min(0) if lo is None else np.minimum(lo, f.min(0))
            hi = f.max(0) if hi is None else np.maximum(hi, f.max(0))
        if dn is None:
            dn = norm.device(ctx)
        m_pad = m if done + m < n else n_pad - done
        _ck(ctx, l.disort_jac_surrogate_features(ctx.ptr, m, m_pad, NLYR, NSTR + 1, C.byref(cols.cstruct),
                                                 dn["f_shift"].ptr, dn["f_scale"].ptr, dn["f_lo"].ptr, dn["f_hi"].ptr,
                                                 w0, x.ptr + done * w0 * 2, None, resp.ptr))
        _ck(ctx, l.disort_jac_surrogate_targets(ctx.ptr, m, m_pad, NLYR, cols.jac.ptr, C.byref(cols.cstruct),
                                                dn["t_shift"].ptr, dn["t_scale"].ptr, wl, t.ptr + done * wl * 4,
                                                resp.ptr))
        ctx.sync()
        st = cols.plan.status()
        if st.any():
            raise RuntimeError(f"plain Jacobian failed on {int((st != 0).sum())} training columns")
        cols.close()
        done += m
        k += 1
    if fit:
        pad = 1e-5 * np.maximum(1.0, np.maximum(np.abs(lo), np.abs(hi)))  # single-precision logs and exponentials
        norm.f_lo[:N_EFEAT_TS] = lo - pad
        norm.f_hi[:N_EFEAT_TS] = hi + pad
    return x, t, n, norm


def train_member(ctx, x, t, n, epochs, seed, batch=2048, lr=2e-3, max_batch=65536, log=None):
    """One ensemble member trained on the dataset (returns the MLP and its loss history)."""
    mlp = MLP(ctx, JWIDTHS, max_batch=max_batch)
    mlp.set_params(mlp.init_params(seed, N_EFEAT_TS, N_JOUT))
    hist = train(ctx, mlp, x, t, n, epochs, batch=batch, lr=lr, seed=seed, n_out=N_JOUT, log=log)
    return mlp, hist


# ---------------------------------------------------------------------------
# Gated serving
# ---------------------------------------------------------------------------
class _JacSub:
    """Device arrays, Jacobian plan and Jacobian buffer for m gathered columns."""

    def __init__(self, ctx, m: int):
        self.ctx, self.m = ctx, m
        L, P = NLYR, NSTR + 1
        self.arrs = [DevArray(ctx, m * L * 8), DevArray(ctx, m * L * 8), DevArray(ctx, m * L * P * 8),
                     DevArray(ctx, m * (L + 1) * 8), DevArray(ctx, m * NSCAL * 8)]
        s = CFluxBatch()
        s.n, s.flags = m, F_THERMAL
        s.dtau, s.ssa, s.pmom, s.temper, s.scal = (a.ptr for a in self.arrs)
        s.total_levels = m * (L + 1)
        self.cstruct = s
        self.plan = _capi.JacPlan(ctx, m, NSTR, L, P, F_THERMAL)
        self.jac = DevArray(ctx, self.plan.jac_bytes)

    def solve(self, cstruct=None):
        self.plan.execute(cstruct if cstruct is not None else self.cstruct, self.jac.ptr)

    def close(self):
        self.plan.close()
        for a in self.arrs + [self.jac]:
            a.free()


class JacGatedSurrogate:
    """Two members behind the OOD + spread + sampled re-solve gate, serving
    the top flux_up and surface flux_down temperature Jacobians [n, 2, D]."""

    SUB_GRANULE = 1024

    def __init__(self, ctx, members, norm: Normalizer, tau: float = 0.0, alpha: float = 0.01, tol: float = JTOL):
        self.ctx, self.members, self.norm = ctx, list(members), norm
        self.dn = norm.device(ctx)
        self.tau, self.alpha, self.tol = tau, alpha, tol
        self.monitor = GateMonitor(alpha)
        self._subs = {}
        self._bufs = {}

    def predict(self, cols: JacColumns):
        """Device Jacobian rows [n, 2, D] (FP64), spread [n] and OOD flags [n]."""
        ctx, l, n = self.ctx, _jlib(), cols.n
        n_pad = _pad64(n)
        w0, wl = JWIDTHS[0], JWIDTHS[-1]
        x = self._scratch("x", n_pad * w0 * 2)
        ood = DevArray(ctx, n * 4)
        ys = [self._scratch(f"y{k}", n_pad * wl * 4) for k in range(len(self.members))]
        rows = DevArray(ctx, n * 2 * D * 8)
        spread = DevArray(ctx, n * 4)
        resp = cols.responses()
        _ck(ctx, l.disort_jac_surrogate_features(ctx.ptr, n, n_pad, NLYR, NSTR + 1, C.byref(cols.cstruct),
                                                 self.dn["f_shift"].ptr, self.dn["f_scale"].ptr, self.dn["f_lo"].ptr,
                                                 self.dn["f_hi"].ptr, w0, x.ptr, ood.ptr, resp.ptr))
        for m, y in zip(self.members, ys):
            for b0 in range(0, n_pad, m.max_batch):
                bb = min(m.max_batch, n_pad - b0)
                m.infer(x.ptr + b0 * w0 * 2, bb, y.ptr + b0 * wl * 4)
        _ck(ctx, l.disort_jac_surrogate_decode(ctx.ptr, n, NLYR, wl, ys[0].ptr, ys[1].ptr if len(ys) > 1 else None,
                                               self.dn["t_shift"].ptr, self.dn["t_scale"].ptr, C.byref(cols.cstruct),
                                               rows.ptr, spread.ptr, resp.ptr))
        return rows, spread, ood

    def _scratch(self, name: str, nbytes: int) -> DevArray:
        """Grow-only device scratch kept between calls."""
        a = self._bufs.get(name)
        if a is None or a.nbytes < nbytes:
            if a is not None:
                a.free()
            a = self._bufs[name] = DevArray(self.ctx, nbytes)
        return a

    def calibrate(self, cols: JacColumns, ref_rows: np.ndarray):
        """Spread threshold such that at most alpha of the accepted
        calibration columns exceed tol.  ref_rows: [n, 2, D] plain Jacobian rows."""
        rows, spread, ood = self.predict(cols)
        self.ctx.sync()
        r = rows.get(np.float64, (cols.n, 2, D))
        sp = spread.get(np.float32, cols.n)
        od = ood.get(np.int32, cols.n).astype(bool)
        for a in (rows, spread, ood):
            a.free()
        err = jacobian_error(r, ref_rows, dbdt_dofs(cols.host))
        ok = ~od
        order = np.argsort(sp[ok])
        bad = (err[ok][order] > self.tol).astype(np.float64)
        frac_bad = np.cumsum(bad) / np.arange(1, bad.size + 1)
        good = np.nonzero(frac_bad <= self.alpha)[0]
        self.tau = float(sp[ok][order][good[-1]]) if good.size else 0.0
        acc = ok & (sp <= self.tau)
        return {"tau": self.tau, "accept_frac": float(acc.mean()), "ood_frac": float(od.mean()),
                "err_median_all": float(np.median(err)), "err_p90_all": float(np.quantile(err, 0.9)),
                "err_p99_all": float(np.quantile(err, 0.99)),
                "viol_frac_accepted": float((err[acc] > self.tol).mean()) if acc.any() else 0.0,
                "viol_frac_all": float((err > self.tol).mean())}, err

    def sub_size(self, m: int, cap: int) -> int:
        """Padded sub-batch size: a multiple of the granule on a grid of eight sizes per octave."""
        g = self.SUB_GRANULE
        q = -(-m // g)
        if q > 8:
            step = 1 << (q.bit_length() - 4)
            q = -(-q // step) * step
        return min(q * g, cap)

    def _sub(self, size: int) -> _JacSub:
        sub = self._subs.get(size)
        if sub is None:
            sub = self._subs[size] = _JacSub(self.ctx, size)
        return sub

    def serve(self, cols: JacColumns, sample_frac: float, seed: int = 0):
        """Gated Jacobian rows for a resident batch: the surrogate where the
        gate accepts, the plain GPU Jacobian for OOD / high-spread / sampled
        columns, or for all while the monitor is tripped.  Returns (device
        rows [n, 2, D], info) with info["flags"] the device flags (bit 0
        routed, bit 1 sampled, bit 2 passes the gate)."""
        ctx, l, n = self.ctx, _jlib(), cols.n
        rows, spread, ood = self.predict(cols)
        g = self.SUB_GRANULE
        cap = -(-max(n, 1) // g) * g
        flags, idx, cnt = DevArray(ctx, n * 4), DevArray(ctx, cap * 4), DevArray(ctx, 8)
        tripped = self.monitor.tripped
        _ck(ctx, l.disort_surrogate_route(ctx.ptr, n, spread.ptr, ood.ptr, self.tau, sample_frac, seed, int(tripped),
                                          flags.ptr, idx.ptr, cnt.ptr))
        m = int(cnt.get(np.int32, 1)[0])
        sampled = bad = 0
        if m:
            if m == n:
                # Every column routed: the plain Jacobian of the batch in place.
                idx.free()
                idx = DevArray.of(ctx, np.arange(n, dtype=np.int32))
                cols.plain()
                jac, plan = cols.jac, cols.plan
            else:
                mpad = self.sub_size(m, cap)
                _ck(ctx, l.disort_surrogate_pad_index(ctx.ptr, m, mpad, idx.ptr))
                sub = self._sub(mpad)
                _ck(ctx, l.disort_surrogate_gather(ctx.ptr, mpad, idx.ptr, NLYR, NSTR + 1, C.byref(cols.cstruct),
                                                   C.byref(sub.cstruct)))
                sub.solve()
                jac, plan = sub.jac, sub.plan
            _ck(ctx, l.disort_jac_surrogate_monitor(ctx.ptr, m, idx.ptr, flags.ptr, NLYR, jac.ptr, rows.ptr,
                                                    C.byref(cols.cstruct), self.tol, cnt.ptr))
            _ck(ctx, l.disort_jac_surrogate_scatter(ctx.ptr, m, idx.ptr, NLYR, jac.ptr, rows.ptr))
            st = plan.status()
            if st[:m].any():
                raise RuntimeError(f"plain Jacobian failed on {int((st[:m] != 0).sum())} routed columns")
            sampled, bad = (int(v) for v in cnt.get(np.int32, 2))
        self.monitor.update(sampled, bad)
        for a in (idx, cnt, spread, ood):
            a.free()
        return rows, {"n": n, "n_plain": m, "tripped": tripped, "monitor_sampled": sampled,
                      "monitor_violations": bad, "flags": flags}

    def close(self):
        for sub in self._subs.values():
            sub.close()
        for a in self._bufs.values():
            a.free()
        self._subs, self._bufs = {}, {}
