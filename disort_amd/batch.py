"""Structure-of-arrays batch of flux-only DISORT instances.

The layout is the one `disort_flux_batch` in include/disort_amd.h describes:
per-instance offsets into flat FP64 arrays, so stream and layer counts may
differ between instances.  One instance is one (wavenumber band x column)
solve of ``DisortFluxSolver<N>::solve`` (src/FluxSolver.cpp:23-171).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

# Flags and scalar slots -- keep in sync with include/disort_amd.h.
F_THERMAL = 1
F_SPHERICAL = 2
F_DELTA_M_PLUS = 4
F_DIFFUSION_BC = 8

S_BEAM_FLUX, S_BEAM_MU, S_ISO_TOP, S_ISO_BOT = 0, 1, 2, 3
S_TEMP_TOP, S_TEMP_BOT, S_EMIS_TOP, S_ALBEDO = 4, 5, 6, 7
S_WN_LO, S_WN_HI, S_RADIUS = 8, 9, 10
NSCAL = 12

OUT_FIELDS = (
    "flux_direct_beam",
    "flux_down",
    "flux_up",
    "flux_tau_divergence",
    "mean_intensity",
    "mean_intensity_down",
    "mean_intensity_up",
    "mean_intensity_direct_beam",
)
NOUT = len(OUT_FIELDS)


class CFluxBatch(C.Structure):
    """ctypes mirror of `disort_flux_batch`."""

    _fields_ = [
        ("n", C.c_int64),
        ("flags", C.c_int32),
        ("reserved", C.c_int32),
        ("nstr", C.c_void_p),
        ("nlyr", C.c_void_p),
        ("pstride", C.c_void_p),
        ("lay_off", C.c_void_p),
        ("mom_off", C.c_void_p),
        ("dtau", C.c_void_p),
        ("ssa", C.c_void_p),
        ("pmom", C.c_void_p),
        ("temper", C.c_void_p),
        ("alt", C.c_void_p),
        ("scal", C.c_void_p),
        ("out", C.c_void_p),
        ("total_levels", C.c_int64),
        ("status", C.c_void_p),
    ]


def _ptr(a):
    return None if a is None else a.ctypes.data


@dataclass
class FluxBatch:
    """Host-side batch.  Arrays are C-contiguous numpy arrays."""

    flags: int
    nstr: np.ndarray
    nlyr: np.ndarray
    pstride: np.ndarray
    dtau: np.ndarray
    ssa: np.ndarray
    pmom: np.ndarray
    scal: np.ndarray
    temper: np.ndarray | None = None
    alt: np.ndarray | None = None
    lay_off: np.ndarray = field(init=False)
    mom_off: np.ndarray = field(init=False)
    out: np.ndarray = field(init=False)
    status: np.ndarray = field(init=False)

    def __post_init__(self):
        self.nstr = np.ascontiguousarray(self.nstr, dtype=np.int32)
        self.nlyr = np.ascontiguousarray(self.nlyr, dtype=np.int32)
        self.pstride = np.ascontiguousarray(self.pstride, dtype=np.int32)
        n = self.nstr.shape[0]
        self.lay_off = np.zeros(n, dtype=np.int64)
        if n > 1:
            self.lay_off[1:] = np.cumsum(self.nlyr[:-1], dtype=np.int64)
        self.mom_off = np.zeros(n, dtype=np.int64)
        if n > 1:
            self.mom_off[1:] = np.cumsum(
                self.nlyr[:-1].astype(np.int64) * self.pstride[:-1], dtype=np.int64
            )
        for name in ("dtau", "ssa", "pmom", "scal", "temper", "alt"):
            a = getattr(self, name)
            if a is not None:
                setattr(self, name, np.ascontiguousarray(a, dtype=np.float64).reshape(-1))
        total_layers = int(self.nlyr.sum())
        if self.dtau.shape[0] != total_layers or self.ssa.shape[0] != total_layers:
            raise ValueError("dtau/ssa length must equal sum(nlyr)")
        if self.pmom.shape[0] != int((self.nlyr.astype(np.int64) * self.pstride).sum()):
            raise ValueError("pmom length must equal sum(nlyr*pstride)")
        if self.scal.shape[0] != n * NSCAL:
            raise ValueError("scal must hold NSCAL entries per instance")
        self.total_levels = total_layers + n
        if self.flags & F_THERMAL and (self.temper is None or self.temper.shape[0] != self.total_levels):
            raise ValueError("thermal batches need temper of length sum(nlyr+1)")
        self.out = np.zeros((NOUT, self.total_levels), dtype=np.float64)
        self.status = np.zeros(n, dtype=np.int32)

    @property
    def n(self) -> int:
        return int(self.nstr.shape[0])

    def c_struct(self) -> CFluxBatch:
        s = CFluxBatch()
        s.n = self.n
        s.flags = int(self.flags)
        s.reserved = 0
        s.nstr = _ptr(self.nstr)
        s.nlyr = _ptr(self.nlyr)
        s.pstride = _ptr(self.pstride)
        s.lay_off = _ptr(self.lay_off)
        s.mom_off = _ptr(self.mom_off)
        s.dtau = _ptr(self.dtau)
        s.ssa = _ptr(self.ssa)
        s.pmom = _ptr(self.pmom)
        s.temper = _ptr(self.temper)
        s.alt = _ptr(self.alt)
        s.scal = _ptr(self.scal)
        s.out = _ptr(self.out)
        s.total_levels = int(self.total_levels)
        s.status = _ptr(self.status)
        return s

    def level_slice(self, i: int) -> slice:
        start = int(self.lay_off[i]) + i
        return slice(start, start + int(self.nlyr[i]) + 1)

    def instance_out(self, i: int) -> dict:
        sl = self.level_slice(i)
        return {name: self.out[f, sl].copy() for f, name in enumerate(OUT_FIELDS)}


def uniform_batch(
    dtau: np.ndarray,
    ssa: np.ndarray,
    pmom: np.ndarray,
    nstr: int,
    scal: np.ndarray,
    flags: int = 0,
    temper: np.ndarray | None = None,
    alt: np.ndarray | None = None,
) -> FluxBatch:
    """Batch of n instances sharing nlyr and moment count.

    dtau, ssa: [n, L]; pmom: [n, L, P] with P >= nstr+1; scal: [n, NSCAL];
    temper: [n, L+1] when thermal; alt: [n, L+1] level altitudes (top
    first) with the spherical beam.
    """
    dtau = np.asarray(dtau, dtype=np.float64)
    n, L = dtau.shape
    pmom = np.asarray(pmom, dtype=np.float64)
    P = pmom.shape[2]
    return FluxBatch(
        flags=flags,
        nstr=np.full(n, nstr, np.int32),
        nlyr=np.full(n, L, np.int32),
        pstride=np.full(n, P, np.int32),
        dtau=dtau,
        ssa=np.asarray(ssa, dtype=np.float64),
        pmom=pmom,
        scal=np.asarray(scal, dtype=np.float64),
        temper=None if temper is None else np.asarray(temper, dtype=np.float64),
        alt=None if alt is None else np.asarray(alt, dtype=np.float64),
    )


def scalars(
    n: int,
    beam_flux=0.0,
    beam_mu=0.5,
    iso_top=0.0,
    iso_bot=0.0,
    temp_top=0.0,
    temp_bot=0.0,
    emis_top=1.0,
    albedo=0.0,
    wn_lo=0.0,
    wn_hi=0.0,
    radius=0.0,
) -> np.ndarray:
    """[n, NSCAL] scalar block; each argument is a scalar or a length-n array."""
    s = np.zeros((n, NSCAL), dtype=np.float64)
    for slot, v in (
        (S_BEAM_FLUX, beam_flux),
        (S_BEAM_MU, beam_mu),
        (S_ISO_TOP, iso_top),
        (S_ISO_BOT, iso_bot),
        (S_TEMP_TOP, temp_top),
        (S_TEMP_BOT, temp_bot),
        (S_EMIS_TOP, emis_top),
        (S_ALBEDO, albedo),
        (S_WN_LO, wn_lo),
        (S_WN_HI, wn_hi),
        (S_RADIUS, radius),
    ):
        s[:, slot] = v
    return s
