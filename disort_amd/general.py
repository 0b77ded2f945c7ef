"""Host-side batch for the general solver (`disort_general_batch` in
include/disort_amd.h): n columns of one DisortConfig shape with per-column
optical properties, the way solveSpectralBands varies them
(src/python_bindings.cpp:96-128).  Outputs follow DisortResult
(src/DisortResult.hpp:16-125): fluxes [8][n][ntau], intensity
[n][ntau][numu][nphi], azimuthal average [n][ntau][numu].
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from .batch import NOUT, NSCAL, OUT_FIELDS

# Keep in sync with include/disort_amd.h.
G_USER_TAU = 16
G_USER_MU = 32
G_ONLY_FLUXES = 64
G_BRDF_RPV = 128
G_CORR_BURAS = 256
G_CORR_NAKAJIMA = 512
G_BRDF_HAPKE = 1024
G_BRDF_COXMUNK = 2048
G_BRDF_ROSSLI = 4096
G_BRDF_ANY = G_BRDF_RPV | G_BRDF_HAPKE | G_BRDF_COXMUNK | G_BRDF_ROSSLI
G_NO_CUTOFF = 8192
S_BEAM_PHI = 11
NRPV = 7


class CGeneralBatch(C.Structure):
    """ctypes mirror of `disort_general_batch`."""

    _fields_ = [
        ("n", C.c_int64),
        ("flags", C.c_int32),
        ("nstr", C.c_int32),
        ("nlyr", C.c_int32),
        ("pstride", C.c_int32),
        ("ntau", C.c_int32),
        ("numu", C.c_int32),
        ("nphi", C.c_int32),
        ("reserved", C.c_int32),
        ("accuracy", C.c_double),
        ("dtau", C.c_void_p),
        ("ssa", C.c_void_p),
        ("pmom", C.c_void_p),
        ("temper", C.c_void_p),
        ("alt", C.c_void_p),
        ("scal", C.c_void_p),
        ("rpv", C.c_void_p),
        ("tau_user", C.c_void_p),
        ("mu_user", C.c_void_p),
        ("phi_user", C.c_void_p),
        ("flux", C.c_void_p),
        ("inten", C.c_void_p),
        ("inten_avg", C.c_void_p),
        ("status", C.c_void_p),
        ("nphase", C.c_int32),
        ("reserved2", C.c_int32),
        ("mu_phase", C.c_void_p),
        ("phase", C.c_void_p),
    ]


def _ptr(a):
    return None if a is None else a.ctypes.data


def _f64(a, shape=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


@dataclass
class GeneralBatch:
    """n columns sharing (nstr, nlyr, pstride, output grid, flags).

    dtau, ssa: [n, L]; pmom: [n, L, P] (P >= nstr+1); scal: [n, NSCAL];
    temper / alt: [n, L+1]; rpv: [n, NRPV]; tau_user: [n, ntau];
    mu_user: [numu] increasing; phi_user: [nphi] degrees.
    """

    flags: int
    nstr: int
    dtau: np.ndarray
    ssa: np.ndarray
    pmom: np.ndarray
    scal: np.ndarray
    temper: np.ndarray | None = None
    alt: np.ndarray | None = None
    rpv: np.ndarray | None = None
    tau_user: np.ndarray | None = None
    mu_user: np.ndarray | None = None
    phi_user: np.ndarray | None = None
    accuracy: float = 0.0
    mu_phase: np.ndarray | None = None  # [nphase] tabulated phase function angles (Buras-Emde)
    phase: np.ndarray | None = None     # [n, L, nphase]
    ntau: int = field(init=False)
    numu: int = field(init=False)
    nphi: int = field(init=False)

    def __post_init__(self):
        self.dtau = _f64(self.dtau)
        n, L = self.dtau.shape
        self.ssa = _f64(self.ssa, (n, L))
        self.pmom = _f64(self.pmom)
        if self.pmom.ndim != 3 or self.pmom.shape[:2] != (n, L):
            raise ValueError("pmom must be [n, nlyr, pstride]")
        if self.pmom.shape[2] < self.nstr + 1:
            raise ValueError("pstride must be >= nstr + 1")
        self.scal = _f64(self.scal, (n, NSCAL))
        self.temper = _f64(self.temper, None if self.temper is None else (n, L + 1))
        self.alt = _f64(self.alt, None if self.alt is None else (n, L + 1))
        self.rpv = _f64(self.rpv, None if self.rpv is None else (n, NRPV))
        self.mu_user = _f64(self.mu_user)
        self.phi_user = _f64(self.phi_user)
        self.mu_phase = _f64(self.mu_phase)
        self.nphase = 0 if self.mu_phase is None else int(self.mu_phase.size)
        self.phase = _f64(self.phase, None if self.phase is None else (n, L, self.nphase))
        if self.flags & G_USER_TAU:
            self.tau_user = _f64(self.tau_user)
            if self.tau_user.ndim == 1:
                self.tau_user = np.ascontiguousarray(np.broadcast_to(self.tau_user, (n, self.tau_user.size)))
            self.ntau = int(self.tau_user.shape[1])
        else:
            self.tau_user = None
            self.ntau = L + 1
        only = bool(self.flags & G_ONLY_FLUXES)
        if self.flags & G_USER_MU and not only:
            if self.mu_user is None:
                raise ValueError("DISORT_G_USER_MU needs mu_user")
            self.numu = int(self.mu_user.size)
        else:
            self.numu = self.nstr
        if only:
            self.nphi = 0
        else:
            if self.phi_user is None:
                raise ValueError("intensities need phi_user")
            self.nphi = int(self.phi_user.size)
        self.flux = np.zeros((NOUT, n, self.ntau))
        self.inten = np.zeros((n, self.ntau, self.numu, max(self.nphi, 0)))
        self.inten_avg = np.zeros((n, self.ntau, self.numu))
        self.status = np.zeros(n, np.int32)

    @property
    def n(self) -> int:
        return int(self.dtau.shape[0])

    @property
    def nlyr(self) -> int:
        return int(self.dtau.shape[1])

    @property
    def pstride(self) -> int:
        return int(self.pmom.shape[2])

    def c_struct(self) -> CGeneralBatch:
        s = CGeneralBatch()
        s.n = self.n
        s.flags = int(self.flags)
        s.nstr = int(self.nstr)
        s.nlyr = self.nlyr
        s.pstride = self.pstride
        s.ntau, s.numu, s.nphi = self.ntau, self.numu, self.nphi
        s.reserved = 0
        s.accuracy = float(self.accuracy)
        for name in ("dtau", "ssa", "pmom", "temper", "alt", "scal", "rpv", "tau_user", "mu_user", "phi_user",
                     "flux", "inten", "inten_avg", "status", "mu_phase", "phase"):
            setattr(s, name, _ptr(getattr(self, name)))
        s.nphase = self.nphase
        return s

    def fluxes(self, i: int) -> dict:
        return {name: self.flux[f, i].copy() for f, name in enumerate(OUT_FIELDS)}


class DeviceGeneralBatch:
    """A GeneralBatch uploaded once to HBM; execute() solves it with
    disort_general_solve(..., DISORT_MEM_DEVICE), asynchronously on the
    context stream; download() copies the outputs back."""

    def __init__(self, ctx, gb: GeneralBatch):
        from . import _capi

        self.ctx, self.gb, self._bufs = ctx, gb, []
        st = gb.c_struct()

        def up(a):
            if a is None:
                return None
            p = ctx.alloc(a.nbytes)
            self._bufs.append(p)
            ctx.h2d(p, np.ascontiguousarray(a))
            return p

        for name in ("dtau", "ssa", "pmom", "temper", "alt", "scal", "rpv", "tau_user", "mu_user", "phi_user",
                     "mu_phase", "phase"):
            setattr(st, name, up(getattr(gb, name)))
        for name in ("flux", "inten", "inten_avg", "status"):
            a = getattr(gb, name)
            p = ctx.alloc(max(a.nbytes, 8))
            self._bufs.append(p)
            setattr(st, name, p)
        self.cstruct = st
        self._capi = _capi

    def execute(self):
        c = self._capi
        rc = c.lib().disort_general_solve(self.ctx.ptr, C.byref(self.cstruct), c.MEM_DEVICE)
        c._check(self.ctx.ptr, rc)

    def download(self) -> GeneralBatch:
        self.ctx.sync()
        for name in ("flux", "inten", "inten_avg", "status"):
            a = getattr(self.gb, name)
            if a.nbytes:
                self.ctx.d2h(a, getattr(self.cstruct, name))
        return self.gb

    def close(self):
        for p in self._bufs:
            self.ctx.free(p)
        self._bufs = []
