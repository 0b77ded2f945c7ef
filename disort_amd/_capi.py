"""ctypes binding of libdisort_amd.so (the C ABI in include/disort_amd.h).

The library is built in-tree by ``__graft_entry__.build()`` into
``disort_amd/lib/libdisort_amd.so``.  There is no CPU fallback: if the
library is missing or no GPU is present, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .batch import NOUT, CFluxBatch, FluxBatch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("DISORT_AMD_LIB") or os.path.join(_HERE, "lib", "libdisort_amd.so")

DISORT_OK = 0
DISORT_E_INVALID = 1
DISORT_E_UNSUPPORTED = 2
DISORT_E_NUMERIC = 3
MEM_HOST = 0
MEM_DEVICE = 1
V_DEFAULT, V_LANE, V_WORKGROUP, V_LANE_PAIR = 0, 1, 2, 3

_lib = None


class DisortError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[disort code {code}] {msg}")
        self.code = code


class DisortInvalid(DisortError, ValueError):
    pass


def lib():
    """Load libdisort_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP library first (__graft_entry__.build())"
            )
        l = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        l.disort_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
        l.disort_ctx_destroy.argtypes = [vp]
        l.disort_last_error.argtypes = [vp]
        l.disort_last_error.restype = C.c_char_p
        l.disort_ctx_stream.argtypes = [vp]
        l.disort_ctx_stream.restype = vp
        l.disort_sync.argtypes = [vp]
        l.disort_flux_solve.argtypes = [vp, vp, C.c_int]
        l.disort_flux_jacobian.argtypes = [vp, vp, vp]
        l.disort_flux_plan_create.argtypes = [vp, C.c_int64, vp, vp, vp, C.c_int32, C.POINTER(vp)]
        l.disort_flux_plan_destroy.argtypes = [vp, vp]
        l.disort_flux_execute.argtypes = [vp, vp, vp]
        l.disort_flux_status.argtypes = [vp, vp, vp]
        l.disort_device_alloc.argtypes = [vp, C.c_size_t, C.POINTER(vp)]
        l.disort_device_free.argtypes = [vp, vp]
        l.disort_memcpy_h2d.argtypes = [vp, vp, vp, C.c_size_t]
        l.disort_memcpy_d2h.argtypes = [vp, vp, vp, C.c_size_t]
        l.disort_set_timing.argtypes = [vp, C.c_int]
        l.disort_stage_times.argtypes = [vp, C.POINTER(C.c_double), C.c_int]
        l.disort_version.restype = C.c_char_p
        l.disort_flux_plan_set_variant.argtypes = [vp, vp, C.c_int, C.c_int]
        l.disort_flux_plan_get_variant.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_int)]
        l.disort_quadrature.argtypes = [C.c_int, vp, vp]
        l.disort_planck_band.argtypes = [C.c_double, C.c_double, C.c_int64, vp, vp]
        l.disort_planck_band_deriv.argtypes = [C.c_double, C.c_double, C.c_int64, vp, vp]
        l.disort_flux_plan_set_concurrent.argtypes = [vp, vp, C.c_int]
        l.disort_flux_plan_set_streams.argtypes = [vp, vp, C.c_int]
        l.disort_flux_plan_set_chunks.argtypes = [vp, vp, C.c_int, C.c_int]
        l.disort_flux_execute_graph.argtypes = [vp, vp, vp]
        l.disort_flux_autotune.argtypes = [vp, vp, vp, C.c_int]
        l.disort_flux_plan_report.argtypes = [vp]
        l.disort_flux_plan_report.restype = C.c_char_p
        l.disort_general_solve.argtypes = [vp, vp, C.c_int]
        l.disort_general_jacobian.argtypes = [vp, vp, vp]
        l.disort_flux_execute_static.argtypes = [vp, vp, C.c_int32, C.c_int32, C.c_int32]
        l.disort_flux_static_bytes.argtypes = [C.c_int64, C.c_int32, C.c_int32]
        l.disort_flux_static_bytes.restype = C.c_size_t
        l.disort_static_scratch_bytes.restype = C.c_size_t
        l.disort_jac_plan_create.argtypes = [vp, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int64,
                                             C.POINTER(vp)]
        l.disort_jac_plan_destroy.argtypes = [vp, vp]
        l.disort_jac_plan_chunk.argtypes = [vp]
        l.disort_jac_plan_chunk.restype = C.c_int64
        l.disort_jac_plan_flux.argtypes = [vp]
        l.disort_jac_plan_flux.restype = vp
        l.disort_jac_execute.argtypes = [vp, vp, vp, vp]
        l.disort_jac_status.argtypes = [vp, vp, vp]
        _lib = l
    return _lib


def _check(ctx_ptr, rc):
    if rc != DISORT_OK:
        msg = lib().disort_last_error(ctx_ptr).decode(errors="replace")
        if rc == DISORT_E_INVALID:
            raise DisortInvalid(rc, msg)
        raise DisortError(rc, msg)


class Context:
    """One device, one HIP stream (disort_ctx)."""

    def __init__(self, device: int = 0):
        self._ptr = C.c_void_p()
        rc = lib().disort_ctx_create(device, C.byref(self._ptr))
        if rc != DISORT_OK:
            raise DisortError(rc, "disort_ctx_create failed (is a GPU visible?)")

    @property
    def ptr(self):
        return self._ptr

    def close(self):
        if self._ptr:
            lib().disort_ctx_destroy(self._ptr)
            self._ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_timing(self, on: bool):
        lib().disort_set_timing(self._ptr, int(on))

    def stage_times(self):
        ms = (C.c_double * 8)()
        n = lib().disort_stage_times(self._ptr, ms, 8)
        return list(ms[:n])

    def sync(self):
        _check(self._ptr, lib().disort_sync(self._ptr))

    def solve(self, batch: FluxBatch) -> FluxBatch:
        """Solve a host batch in place (batch.out, batch.status)."""
        st = batch.c_struct()
        rc = lib().disort_flux_solve(self._ptr, C.byref(st), MEM_HOST)
        _check(self._ptr, rc)
        return batch

    def flux_jacobian(self, batch: FluxBatch) -> list:
        """Temperature Jacobians of a thermal host batch: one array
        [6, L+1, L+2] per instance (fields flux_up, flux_down,
        mean_intensity, mean_intensity_down, mean_intensity_up,
        flux_tau_divergence; dof L+1 is the surface temperature)."""
        per = (batch.nlyr.astype(np.int64) + 1) * (batch.nlyr + 2)
        jac = np.zeros(int(6 * per.sum()), np.float64)
        st = batch.c_struct()
        _check(self._ptr, lib().disort_flux_jacobian(self._ptr, C.byref(st), jac.ctypes.data))
        res, off = [], 0
        for i, L in enumerate(batch.nlyr.tolist()):
            res.append(jac[off:off + 6 * int(per[i])].reshape(6, L + 1, L + 2))
            off += 6 * int(per[i])
        return res

    def general_solve(self, gb):
        """Solve a host disort_amd.general.GeneralBatch in place (gb.flux,
        gb.inten, gb.inten_avg, gb.status)."""
        st = gb.c_struct()
        rc = lib().disort_general_solve(self._ptr, C.byref(st), MEM_HOST)
        _check(self._ptr, rc)
        return gb

    def general_jacobian(self, gb) -> np.ndarray:
        """Temperature Jacobians of a thermal GeneralBatch: [n, 6, ntau,
        nlyr+2] (fields as flux_jacobian; rows the output levels)."""
        jac = np.zeros((gb.n, 6, gb.ntau, gb.nlyr + 2), np.float64)
        st = gb.c_struct()
        _check(self._ptr, lib().disort_general_jacobian(self._ptr, C.byref(st), jac.ctypes.data))
        return jac

    # -- device helpers -----------------------------------------------------
    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        _check(self._ptr, lib().disort_device_alloc(self._ptr, max(int(nbytes), 1), C.byref(p)))
        return p.value

    def free(self, p: int):
        lib().disort_device_free(self._ptr, C.c_void_p(p))

    def h2d(self, dst: int, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        _check(self._ptr, lib().disort_memcpy_h2d(self._ptr, C.c_void_p(dst), arr.ctypes.data, arr.nbytes))

    def d2h(self, arr: np.ndarray, src: int):
        _check(self._ptr, lib().disort_memcpy_d2h(self._ptr, arr.ctypes.data, C.c_void_p(src), arr.nbytes))


class DeviceFluxBatch:
    """A FluxBatch resident in device memory plus its plan, for repeated
    asynchronous solves (disort_flux_plan_create / disort_flux_execute)."""

    def __init__(self, ctx: Context, batch: FluxBatch):
        self.ctx = ctx
        self.host = batch
        self._bufs = []
        l = lib()
        self.plan = C.c_void_p()
        rc = l.disort_flux_plan_create(
            ctx.ptr,
            batch.n,
            batch.nstr.ctypes.data,
            batch.nlyr.ctypes.data,
            batch.pstride.ctypes.data,
            int(batch.flags),
            C.byref(self.plan),
        )
        _check(ctx.ptr, rc)

        def up(a):
            if a is None:
                return None
            p = ctx.alloc(a.nbytes)
            self._bufs.append(p)
            ctx.h2d(p, a)
            return p

        s = CFluxBatch()
        s.n = batch.n
        s.flags = int(batch.flags)
        s.nstr = up(batch.nstr)
        s.nlyr = up(batch.nlyr)
        s.pstride = up(batch.pstride)
        s.lay_off = up(batch.lay_off)
        s.mom_off = up(batch.mom_off)
        s.dtau = up(batch.dtau)
        s.ssa = up(batch.ssa)
        s.pmom = up(batch.pmom)
        s.temper = up(batch.temper)
        s.alt = up(batch.alt)
        s.scal = up(batch.scal)
        self.out_ptr = ctx.alloc(batch.out.nbytes)
        self._bufs.append(self.out_ptr)
        s.out = self.out_ptr
        s.total_levels = int(batch.total_levels)
        s.status = None
        self.cstruct = s

    def set_variant(self, nstr: int, variant: int):
        """Kernel variant of the bucket with `nstr` streams (V_LANE / V_LANE_PAIR / V_WORKGROUP)."""
        _check(self.ctx.ptr, lib().disort_flux_plan_set_variant(self.ctx.ptr, self.plan, nstr, variant))

    def get_variant(self, nstr: int) -> int:
        v = C.c_int()
        _check(self.ctx.ptr, lib().disort_flux_plan_get_variant(self.ctx.ptr, self.plan, nstr, C.byref(v)))
        return v.value

    def set_concurrent(self, on: bool):
        """Run stream-count buckets concurrently on side streams (or serially)."""
        _check(self.ctx.ptr, lib().disort_flux_plan_set_concurrent(self.ctx.ptr, self.plan, int(on)))

    def set_streams(self, streams: int):
        """1: buckets serially on the context stream; 2..4: concurrently on that many side streams."""
        _check(self.ctx.ptr, lib().disort_flux_plan_set_streams(self.ctx.ptr, self.plan, int(streams)))

    def set_chunks(self, nstr: int, chunks: int):
        """Split the bucket of stream count `nstr` into `chunks` buckets (1..8)."""
        _check(self.ctx.ptr, lib().disort_flux_plan_set_chunks(self.ctx.ptr, self.plan, int(nstr), int(chunks)))

    def autotune(self, reps: int = 3) -> dict:
        """Time every kernel variant per bucket and both schedules; keep the fastest."""
        import json

        _check(self.ctx.ptr, lib().disort_flux_autotune(self.ctx.ptr, self.plan, C.byref(self.cstruct), reps))
        return json.loads(lib().disort_flux_plan_report(self.plan).decode())

    def execute_graph(self):
        """Enqueue one solve as a single hipGraph launch (captured on first use)."""
        _check(self.ctx.ptr, lib().disort_flux_execute_graph(self.ctx.ptr, self.plan, C.byref(self.cstruct)))

    def buckets(self):
        """Stream counts present in the batch."""
        return sorted(set(int(x) for x in self.host.nstr))

    def execute(self):
        """Enqueue one solve on the context stream (asynchronous)."""
        _check(self.ctx.ptr, lib().disort_flux_execute(self.ctx.ptr, self.plan, C.byref(self.cstruct)))

    def download(self) -> FluxBatch:
        self.ctx.sync()
        self.ctx.d2h(self.host.out, self.out_ptr)
        st = np.zeros(self.host.n, np.int32)
        _check(self.ctx.ptr, lib().disort_flux_status(self.ctx.ptr, self.plan, st.ctypes.data))
        self.host.status[:] = st
        return self.host

    def close(self):
        if self.plan:
            lib().disort_flux_plan_destroy(self.ctx.ptr, self.plan)
            self.plan = C.c_void_p()
        for p in self._bufs:
            self.ctx.free(p)
        self._bufs = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


JAC_EXPAND = 0x100  # DISORT_JAC_EXPAND: one forward solve per column x degree of freedom


class JacPlan:
    """Device-resident temperature Jacobians of n uniform columns
    (disort_jac_plan_create / disort_jac_execute): execute(batch, jac_ptr)
    writes jac [n, 6, nlyr+1, nlyr+2] (FP64, device) for a device batch
    descriptor, asynchronously on the context stream.  Factored form (the
    unit-source right-hand sides of the boundary system) for nstr <= 16
    unless flags has JAC_EXPAND."""

    def __init__(self, ctx: Context, n: int, nstr: int, nlyr: int, pstride: int, flags: int, max_expanded: int = 0):
        self.ctx, self.n, self.nstr, self.nlyr = ctx, int(n), int(nstr), int(nlyr)
        self.plan = C.c_void_p()
        _check(ctx.ptr, lib().disort_jac_plan_create(ctx.ptr, self.n, self.nstr, self.nlyr, int(pstride), int(flags),
                                                      int(max_expanded), C.byref(self.plan)))
        self.chunk = int(lib().disort_jac_plan_chunk(self.plan))

    @property
    def jac_bytes(self) -> int:
        return self.n * 6 * (self.nlyr + 1) * (self.nlyr + 2) * 8

    def set_streams(self, streams: int):
        """Schedule of the inner flux plan (see DeviceFluxBatch.set_streams)."""
        _check(self.ctx.ptr, lib().disort_flux_plan_set_streams(self.ctx.ptr, lib().disort_jac_plan_flux(self.plan),
                                                                int(streams)))

    def set_chunks(self, chunks: int):
        _check(self.ctx.ptr, lib().disort_flux_plan_set_chunks(self.ctx.ptr, lib().disort_jac_plan_flux(self.plan),
                                                               self.nstr, int(chunks)))

    def execute(self, cstruct, jac_ptr: int):
        _check(self.ctx.ptr, lib().disort_jac_execute(self.ctx.ptr, self.plan, C.byref(cstruct), C.c_void_p(jac_ptr)))

    def status(self) -> np.ndarray:
        st = np.zeros(self.n, np.int32)
        _check(self.ctx.ptr, lib().disort_jac_status(self.ctx.ptr, self.plan, st.ctypes.data))
        return st

    def close(self):
        if self.plan:
            lib().disort_jac_plan_destroy(self.ctx.ptr, self.plan)
            self.plan = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def version() -> str:
    return lib().disort_version().decode()


def planck_band(lo: float, hi: float, t) -> np.ndarray:
    """Band-integrated Planck function at host temperatures t (the kernels' arithmetic)."""
    t = np.ascontiguousarray(t, dtype=np.float64).reshape(-1)
    b = np.empty_like(t)
    rc = lib().disort_planck_band(float(lo), float(hi), t.size, t.ctypes.data, b.ctypes.data)
    if rc:
        raise DisortInvalid(rc, "disort_planck_band: invalid band or temperature")
    return b


def planck_band_deriv(lo: float, hi: float, t) -> np.ndarray:
    """dB/dT of the band-integrated Planck function at host temperatures t."""
    t = np.ascontiguousarray(t, dtype=np.float64).reshape(-1)
    d = np.empty_like(t)
    rc = lib().disort_planck_band_deriv(float(lo), float(hi), t.size, t.ctypes.data, d.ctypes.data)
    if rc:
        raise DisortInvalid(rc, "disort_planck_band_deriv: invalid band or temperature")
    return d
