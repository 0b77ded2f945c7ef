"""ctypes access to the CPU oracle (oracle/liboracle.so) and, where it was
built, to the reference's own Eigen-free numerics (oracle/_ref).

Test infrastructure: the oracle checks the GPU results and is never the thing tested.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from disort_amd.batch import NOUT, FluxBatch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
_REF = os.path.join(ROOT, "oracle", "_ref", "libref_parts.so")

_lib = None
_ref = None

D = C.c_double
PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_ORACLE):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        l = C.CDLL(_ORACLE)
        l.oracle_planck.argtypes = [D, D, D, PD, C.c_char_p, C.c_int]
        l.oracle_planck_deriv.argtypes = [D, D, D, PD, C.c_char_p, C.c_int]
        l.oracle_legendre_flat.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        l.oracle_gauss.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
        l.oracle_band_factor.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
        l.oracle_band_solve.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        l.oracle_sym_eigen.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        l.oracle_cholesky.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
        l.oracle_chapman.argtypes = [C.c_int, D, C.c_void_p, C.c_void_p, C.c_int, D, D]
        l.oracle_chapman.restype = D
        l.oracle_flux_one.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_char_p, C.c_int]
        l.oracle_flux_jacobian_one.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p,
                                               C.c_char_p, C.c_int]
        l.oracle_ld_flux_jacobian_one.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
        l.oracle_ld_flux_one.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_char_p, C.c_int]
        l.oracle_flux_batch.argtypes = [C.c_void_p, C.c_int, PD]
        l.oracle_flux_batch.restype = C.c_int64
        l.oracle_general_one.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                         PI, C.c_char_p, C.c_int]
        l.oracle_general_jacobian_one.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p,
                                                  C.c_char_p, C.c_int]
        l.oracle_rpv.argtypes = [C.c_void_p, D, D, D]
        l.oracle_general_albtrans.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_char_p, C.c_int]
        l.oracle_rpv.restype = D
        _lib = l
    return _lib


def ref():
    """The reference's Planck/Legendre/BandSolver, or None when not built."""
    global _ref
    if _ref is None and os.path.exists(_REF):
        r = C.CDLL(_REF)
        r.ref_planck2.argtypes = [D, D, D, PD]
        r.ref_planck_deriv2.argtypes = [D, D, D, PD]
        r.ref_legendre_flat.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        r.ref_band_factor.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
        r.ref_band_solve.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        _ref = r
    return _ref


class OracleError(Exception):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def planck(lo, hi, t):
    out = D()
    err = C.create_string_buffer(256)
    rc = lib().oracle_planck(lo, hi, t, C.byref(out), err, 256)
    if rc:
        raise OracleError(rc, err.value.decode())
    return out.value


def planck_deriv(lo, hi, t):
    out = D()
    err = C.create_string_buffer(256)
    rc = lib().oracle_planck_deriv(lo, hi, t, C.byref(out), err, 256)
    if rc:
        raise OracleError(rc, err.value.decode())
    return out.value


def gauss(nn):
    cmu = np.zeros(2 * nn)
    cwt = np.zeros(2 * nn)
    assert lib().oracle_gauss(nn, cmu.ctypes.data, cwt.ctypes.data) == 0
    return cmu, cwt


def legendre_flat(mu, m, twonm1, ylm=None):
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    nmu = mu.shape[0]
    if ylm is None:
        ylm = np.zeros((twonm1 + 2, nmu))
    lib().oracle_legendre_flat(nmu, m, nmu - 1, twonm1, mu.ctypes.data, ylm.ctypes.data)
    return ylm


def flux_one(batch: FluxBatch, i: int, index_from_bottom=False):
    """Oracle solve of instance i; returns an [8, nlyr+1] array."""
    L = int(batch.nlyr[i])
    out = np.zeros((NOUT, L + 1))
    err = C.create_string_buffer(512)
    st = batch.c_struct()
    rc = lib().oracle_flux_one(C.byref(st), i, int(index_from_bottom), out.ctypes.data, err, 512)
    if rc:
        raise OracleError(rc, err.value.decode())
    return out


def flux_jacobian_one(batch: FluxBatch, i: int, index_from_bottom=False):
    """Oracle solve of instance i with temperature Jacobians: returns
    (out [8, L+1], jac [6, L+1, L+2]); jac fields flux_up, flux_down,
    mean_intensity, mean_intensity_down, mean_intensity_up,
    flux_tau_divergence; dof L+1 is the surface temperature."""
    L = int(batch.nlyr[i])
    out = np.zeros((NOUT, L + 1))
    jac = np.zeros((6, L + 1, L + 2))
    err = C.create_string_buffer(512)
    st = batch.c_struct()
    rc = lib().oracle_flux_jacobian_one(C.byref(st), i, int(index_from_bottom), out.ctypes.data, jac.ctypes.data,
                                        err, 512)
    if rc:
        raise OracleError(rc, err.value.decode())
    return out, jac


def flux_jacobian_one_extended(batch: FluxBatch, i: int):
    """flux_jacobian_one in x87 extended precision, rounded to FP64."""
    L = int(batch.nlyr[i])
    out = np.zeros((NOUT, L + 1))
    jac = np.zeros((6, L + 1, L + 2))
    err = C.create_string_buffer(512)
    st = batch.c_struct()
    rc = lib().oracle_ld_flux_jacobian_one(C.byref(st), i, out.ctypes.data, jac.ctypes.data, err, 512)
    if rc:
        raise OracleError(rc, err.value.decode())
    return out, jac


def flux_one_extended(batch: FluxBatch, i: int):
    """The same solve in x87 extended precision (64-bit significand),
    rounded to FP64: a higher-precision truth for rounding comparisons."""
    L = int(batch.nlyr[i])
    out = np.zeros((NOUT, L + 1))
    err = C.create_string_buffer(512)
    st = batch.c_struct()
    rc = lib().oracle_ld_flux_one(C.byref(st), i, out.ctypes.data, err, 512)
    if rc:
        raise OracleError(rc, err.value.decode())
    return out


def flux_batch(batch: FluxBatch, nthreads=0):
    """Oracle solve of the whole batch into batch.out; returns seconds."""
    st = batch.c_struct()
    secs = D()
    failed = lib().oracle_flux_batch(C.byref(st), nthreads, C.byref(secs))
    if failed:
        raise OracleError(2, f"{failed} oracle instances failed")
    return secs.value


def general_one(gb, i, index_from_bottom=False):
    """Oracle solve of column i of a disort_amd.general.GeneralBatch.
    Returns (fluxes [8][ntau], intensity [ntau][numu][nphi],
    azimuthal average [ntau][numu], modes computed)."""
    flux = np.zeros((NOUT, gb.ntau))
    inten = np.zeros((gb.ntau, gb.numu, max(gb.nphi, 0)))
    avg = np.zeros((gb.ntau, gb.numu))
    modes = C.c_int(0)
    err = C.create_string_buffer(512)
    st = gb.c_struct()
    rc = lib().oracle_general_one(C.byref(st), i, int(index_from_bottom), flux.ctypes.data, inten.ctypes.data,
                                  avg.ctypes.data, C.byref(modes), err, 512)
    if rc:
        raise OracleError(rc, err.value.decode())
    return flux, inten, avg, modes.value


def general_jacobian_one(gb, i, index_from_bottom=False):
    """Column i of a GeneralBatch with temperature Jacobians: returns
    (fluxes [8][ntau], jac [6][ntau][nlyr+2])."""
    flux = np.zeros((NOUT, gb.ntau))
    jac = np.zeros((6, gb.ntau, gb.nlyr + 2))
    err = C.create_string_buffer(512)
    st = gb.c_struct()
    rc = lib().oracle_general_jacobian_one(C.byref(st), i, int(index_from_bottom), flux.ctypes.data,
                                           jac.ctypes.data, err, 512)
    if rc:
        raise OracleError(rc, err.value.decode())
    return flux, jac


def general_all(gb):
    """Oracle solve of every column; fills gb.flux / gb.inten / gb.inten_avg."""
    for i in range(gb.n):
        f, u, a, _ = general_one(gb, i)
        gb.flux[:, i] = f
        gb.inten[i] = u
        gb.inten_avg[i] = a
    return gb


def general_albtrans(gb, i):
    """SPECIAL_BC of column i (albedo and transmissivity of the medium):
    returns (albedo [2h], transmissivity [2h], (spherical albedo, spherical
    transmissivity)) with h positive polar angles, the first h entries set."""
    h = gb.mu_user.size if (gb.flags & 32) else gb.nstr // 2
    alb, trn, sph = np.zeros(2 * h), np.zeros(2 * h), np.zeros(2)
    err = C.create_string_buffer(512)
    st = gb.c_struct()
    rc = lib().oracle_general_albtrans(C.byref(st), i, alb.ctypes.data, trn.ctypes.data, sph.ctypes.data, err, 512)
    if rc:
        raise OracleError(rc, err.value.decode())
    return alb, trn, (float(sph[0]), float(sph[1]))


def rpv(params, mu1, mu2, phi):
    p = np.ascontiguousarray(params, dtype=np.float64)
    return lib().oracle_rpv(p.ctypes.data, mu1, mu2, phi)
