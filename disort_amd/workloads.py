"""Seeded synthetic workloads of the flux path (counter-based Philox RNG).

c1  single column: nlyr=4, nstr=16, Henyey-Greenstein g=0.8 (17 moments),
    delta_tau=[0.1,0.5,1,2], single_scat_albedo=0.9, Lambertian albedo 0.1,
    beam flux pi at mu0=0.5, no thermal emission (README quick start of the
    reference, README.md:57-84).
c2  spectral bands: 10,000 bands [100+b, 101+b] cm^-1 x one column, nlyr=60,
    nstr=16, thermal emission with level temperatures linear from 200 K (top)
    to 300 K (surface); per band and layer delta_tau = 10^U(-4,1),
    single_scat_albedo ~ U(0,0.999), HG g ~ U(0,0.9); per band albedo
    ~ U(0,0.3); one beam cosine mu0 ~ U(0.1,1) for the column (the spectral
    driver cannot override mu0 per band, src/python_bindings.cpp:51-60);
    beam flux pi; temperature_bottom 300 K, temperature_top 0 K,
    emissivity_top 1 (choices not fixed by the workload statement).
c3  full solver: 4,096 columns, nlyr=80, nstr=32, nmom=33; per layer
    delta_tau = 10^U(-4,1), single_scat_albedo ~ U(0,0.999), phase moments
    w Rayleigh + (1-w) HG with w ~ U(0,1), g ~ U(0,0.95); per column
    mu0 ~ U(0.05,1), beam flux pi, beam azimuth 0; pseudo-spherical beam
    (radius 6371 km, level altitudes 80 -> 0 km in 1 km steps); an RPV
    surface per column, rho0 ~ U(0.01,0.3), k ~ U(0.5,1.5),
    theta ~ U(-0.3,0), sigma=t1=t2=0, scale=1; intensities at
    16 polar cosines +-{0.1,0.2,0.35,0.5,0.65,0.8,0.9,1} and 8 azimuths
    0..315 deg at the 81 layer boundaries; Delta-M+ requested;
    intensity_corr_buras and intensity_corr_nakajima on; every Fourier mode
    (accuracy_fourier_series = 0); no thermal emission (choices not fixed by
    the workload statement).
"""
from __future__ import annotations

import numpy as np

from .batch import F_THERMAL, FluxBatch, scalars, uniform_batch

C2_SEED = 20261203


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(seed))


def hg_moments(g: np.ndarray, nmom_plus1: int) -> np.ndarray:
    """pmom_k = g^k, k = 0..nmom_plus1-1 (src/PhaseFunction.cpp:6-18)."""
    g = np.asarray(g, dtype=np.float64)
    k = np.arange(nmom_plus1, dtype=np.float64)
    return np.power(g[..., None], k)


def c1() -> FluxBatch:
    nstr, L = 16, 4
    dtau = np.array([[0.1, 0.5, 1.0, 2.0]])
    ssa = np.full((1, L), 0.9)
    pmom = hg_moments(np.full((1, L), 0.8), nstr + 1)
    return uniform_batch(dtau, ssa, pmom, nstr, scalars(1, beam_flux=np.pi, beam_mu=0.5, albedo=0.1))


def c2(n_bands: int = 10000, nlyr: int = 60, nstr: int = 16, seed: int = C2_SEED, first_band: int = 0) -> FluxBatch:
    """Bands first_band .. first_band+n_bands-1 of the C2 workload.

    Every band's draws depend only on (seed, band index), so any subset of
    bands reproduces the same instances as the full 10,000-band run.
    """
    dtau = np.empty((n_bands, nlyr))
    ssa = np.empty((n_bands, nlyr))
    g = np.empty((n_bands, nlyr))
    alb = np.empty(n_bands)
    for k in range(n_bands):
        r = rng(seed + 1 + first_band + k)
        dtau[k] = 10.0 ** r.uniform(-4.0, 1.0, nlyr)
        ssa[k] = r.uniform(0.0, 0.999, nlyr)
        g[k] = r.uniform(0.0, 0.9, nlyr)
        alb[k] = r.uniform(0.0, 0.3)
    mu0 = rng(seed).uniform(0.1, 1.0)
    b = np.arange(first_band, first_band + n_bands, dtype=np.float64)
    wn_lo, wn_hi = 100.0 + b, 101.0 + b
    temper = np.tile(np.linspace(200.0, 300.0, nlyr + 1), (n_bands, 1))
    sc = scalars(
        n_bands,
        beam_flux=np.pi,
        beam_mu=mu0,
        albedo=alb,
        temp_top=0.0,
        temp_bot=300.0,
        emis_top=1.0,
        wn_lo=wn_lo,
        wn_hi=wn_hi,
    )
    return uniform_batch(dtau, ssa, hg_moments(g, nstr + 1), nstr, sc, flags=F_THERMAL, temper=temper)


def random_flux_batch(
    n: int,
    nlyr: int,
    nstr: int,
    seed: int,
    thermal: bool = True,
    beam: bool = True,
    flags_extra: int = 0,
    pstride: int | None = None,
) -> FluxBatch:
    """Small C2-distributed batch for parity tests (any nstr / nlyr)."""
    r = rng(seed)
    dtau = 10.0 ** r.uniform(-4.0, 1.0, (n, nlyr))
    ssa = r.uniform(0.0, 0.999, (n, nlyr))
    g = r.uniform(0.0, 0.9, (n, nlyr))
    P = pstride if pstride is not None else nstr + 1
    pmom = hg_moments(g, P)
    alb = r.uniform(0.0, 0.3, n)
    mu0 = r.uniform(0.1, 1.0, n)
    wn = r.uniform(100.0, 3000.0, n)
    temper = np.tile(np.linspace(200.0, 300.0, nlyr + 1), (n, 1)) + r.uniform(-5, 5, (n, nlyr + 1))
    sc = scalars(
        n,
        beam_flux=np.pi if beam else 0.0,
        beam_mu=mu0,
        albedo=alb,
        temp_top=0.0,
        temp_bot=300.0,
        emis_top=1.0,
        wn_lo=wn,
        wn_hi=wn + 1.0,
    )
    flags = (F_THERMAL if thermal else 0) | flags_extra
    return uniform_batch(dtau, ssa, pmom, nstr, sc, flags=flags, temper=temper if thermal else None)


C4_SEED = 20261204
C4_STREAMS = (4, 8, 16, 32, 64)


def c4(n: int = 50000, seed: int = C4_SEED, first: int = 0, streams=C4_STREAMS, lmin: int = 10, lmax: int = 120) -> FluxBatch:
    """Ragged batch: instance k has nstr ~ U{streams} and nlyr ~ U{lmin..lmax},
    beam only (flux pi, mu0 ~ U(0.1,1) per instance), Lambertian albedo
    ~ U(0,0.3), per-layer optical properties drawn as in c2.  Instance k's
    draws depend only on (seed, first+k), so subsets reproduce the full batch.
    """
    nstr = np.empty(n, np.int32)
    nlyr = np.empty(n, np.int32)
    dt, om, pm, sc = [], [], [], []
    for k in range(n):
        r = rng(seed + 1 + first + k)
        N = int(streams[r.integers(len(streams))])
        L = int(r.integers(lmin, lmax + 1))
        nstr[k], nlyr[k] = N, L
        dt.append(10.0 ** r.uniform(-4.0, 1.0, L))
        om.append(r.uniform(0.0, 0.999, L))
        pm.append(hg_moments(r.uniform(0.0, 0.9, L), N + 1).reshape(-1))
        sc.append(scalars(1, beam_flux=np.pi, beam_mu=r.uniform(0.1, 1.0), albedo=r.uniform(0.0, 0.3))[0])
    return FluxBatch(
        flags=0,
        nstr=nstr,
        nlyr=nlyr,
        pstride=nstr + 1,
        dtau=np.concatenate(dt),
        ssa=np.concatenate(om),
        pmom=np.concatenate(pm),
        scal=np.stack(sc),
    )


C3_SEED = 20261205
C3_MU = np.array([-1.0, -0.9, -0.8, -0.65, -0.5, -0.35, -0.2, -0.1, 0.1, 0.2, 0.35, 0.5, 0.65, 0.8, 0.9, 1.0])
C3_PHI = np.arange(8) * 45.0


def c3(n: int = 4096, nlyr: int = 80, nstr: int = 32, seed: int = C3_SEED, first: int = 0):
    """The general-solver workload (see the module docstring); column k draws
    from its own Philox stream, so any subset reproduces the same columns.

    Per column: dtau log-uniform[1e-4, 10], single-scattering albedo
    U[0, 0.999], phase function w * Rayleigh + (1 - w) * HG with w ~ U[0, 1],
    g ~ U[0, 0.95]; beam mu0 ~ U[0.05, 1]; its own RPV surface with
    rho0 ~ U[0.01, 0.3], k ~ U[0.5, 1.5], theta ~ U[-0.3, 0] (sigma = t1 =
    t2 = 0, scale 1)."""
    from . import general as G

    P = nstr + 2  # moments 0..nmom, nmom = nstr + 1
    dtau = np.empty((n, nlyr))
    ssa = np.empty((n, nlyr))
    pmom = np.empty((n, nlyr, P))
    mu0 = np.empty(n)
    rpv = np.zeros((n, 7))
    rpv[:, 6] = 1.0
    ray = np.zeros(P)
    ray[0], ray[2] = 1.0, 0.1
    for k in range(n):
        r = rng(seed + 1 + first + k)
        dtau[k] = 10.0 ** r.uniform(-4.0, 1.0, nlyr)
        ssa[k] = r.uniform(0.0, 0.999, nlyr)
        g = r.uniform(0.0, 0.95, nlyr)
        w = r.uniform(0.0, 1.0, nlyr)
        pmom[k] = w[:, None] * ray + (1.0 - w[:, None]) * hg_moments(g, P)
        mu0[k] = r.uniform(0.05, 1.0)
        rpv[k, 0] = r.uniform(0.01, 0.3)
        rpv[k, 1] = r.uniform(0.5, 1.5)
        rpv[k, 2] = r.uniform(-0.3, 0.0)
    sc = scalars(n, beam_flux=np.pi, beam_mu=mu0, radius=6371.0)
    alt = np.repeat(np.linspace(float(nlyr), 0.0, nlyr + 1)[None], n, axis=0)
    flags = (2 | 4 | G.G_USER_MU | G.G_BRDF_RPV | G.G_CORR_BURAS | G.G_CORR_NAKAJIMA)
    return G.GeneralBatch(flags=flags, nstr=nstr, dtau=dtau, ssa=ssa, pmom=pmom, scal=sc, alt=alt,
                          rpv=rpv, mu_user=C3_MU, phi_user=C3_PHI, accuracy=0.0)
