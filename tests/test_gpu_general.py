"""General solver on the GPU (disort_general_solve) against the CPU oracle
(oracle/general_oracle.cpp), which the reference's intensity tables pin
(tests/test_oracle_general.py).

Tolerance: FP64 on both sides with different operation orders (parallel
Jacobi, panel LU, regrouped layer sums): every output agrees with the oracle
to 1e-9 relative, with an absolute floor of 1e-12 times the largest
magnitude of that quantity (or of the incident beam mu0 F0, if larger) over
the batch.  Conservative scattering
(omega' = 1) is noise-dominated in every method (its smallest eigenvalue is
rounding noise); batches with such a layer use the reference's own
tolerance for it, 1e-8 relative, with a floor of 2e-9 of the incident beam
mu0 F0."""
import numpy as np
import pytest

import oracle_lib as ol
from disort_amd import general as G
from disort_amd.batch import OUT_FIELDS, scalars
from test_oracle_general import CASES, case_batch, check_expect

pytestmark = pytest.mark.gpu

RTOL, AFLOOR = 1e-9, 1e-12


@pytest.fixture(scope="module")
def ctx():
    from disort_amd._capi import Context

    return Context(0)


def compare(gpu, ref, label):
    rtol, afloor, aabs = RTOL, AFLOOR, 0.0
    beam = float((gpu.scal[:, 0] * gpu.scal[:, 1]).max())
    if (gpu.ssa >= 1.0).any():
        rtol, afloor = 1e-8, 0.0
        aabs = 2e-9 * max(1.0, float((gpu.scal[:, 0] * gpu.scal[:, 1]).max()))
    for name, a, b in (("flux", gpu.flux, ref.flux), ("intensity", gpu.inten, ref.inten),
                       ("azimuthal average", gpu.inten_avg, ref.inten_avg)):
        if b.size == 0:
            continue
        if name == "flux":
            for f, fname in enumerate(OUT_FIELDS):
                x, y = a[f], b[f]
                tol = rtol * np.abs(y) + afloor * max(np.abs(y).max(), beam, 1e-300) + aabs
                bad = np.abs(x - y) > tol
                assert not bad.any(), (f"{label}: {fname} differs at {np.argwhere(bad)[:5].tolist()}: "
                                       f"gpu {x[bad][:5]} oracle {y[bad][:5]}")
        else:
            tol = rtol * np.abs(b) + afloor * max(np.abs(b).max(), beam, 1e-300) + aabs
            bad = np.abs(a - b) > tol
            assert not bad.any(), (f"{label}: {name} differs at {np.argwhere(bad)[:5].tolist()}: "
                                   f"gpu {a[bad][:5]} oracle {b[bad][:5]}")


def solve_both(ctx, gb, label):
    import copy

    ref = copy.deepcopy(gb)
    ol.general_all(ref)
    print(f"[gpu general] {label}: n={gb.n} nstr={gb.nstr} nlyr={gb.nlyr} ntau={gb.ntau} numu={gb.numu} "
          f"nphi={gb.nphi} flags={gb.flags}", flush=True)
    ctx.general_solve(gb)
    assert (gb.status == 0).all()
    compare(gb, ref, label)
    return gb, ref


@pytest.mark.parametrize("c", CASES, ids=[c["name"] for c in CASES])
def test_reference_cases(ctx, c):
    gb, ref = solve_both(ctx, case_batch(c, n=2), c["name"])
    for i in range(2):
        check_expect(c, gb.flux[:, i], gb.inten[i], label="gpu ")


def random_general(n, L, nstr, seed, flags=0, numu=6, nphi=4, user_tau=False, rpv=False, spherical=False,
                   thermal=False, accuracy=0.0, corr=False, P=None, mu0=None):
    r = np.random.default_rng(seed)
    P = P or nstr + 2
    dtau = r.uniform(0.01, 1.5, (n, L))
    ssa = r.uniform(0.2, 0.99, (n, L))
    g = r.uniform(0.0, 0.85, (n, L))
    w = r.uniform(0.0, 1.0, (n, L))
    k = np.arange(P)
    hg = g[..., None] ** k
    ray = np.zeros(P)
    ray[0], ray[2] = 1.0, 0.1
    pmom = w[..., None] * ray + (1 - w[..., None]) * hg
    kw = dict(beam_flux=np.pi, beam_mu=r.uniform(0.15, 0.95, n) if mu0 is None else mu0,
              albedo=r.uniform(0.0, 0.4, n), iso_top=0.05)
    f = flags | G.G_USER_MU
    temper = None
    if thermal:
        f |= 1
        kw.update(temp_top=180.0, temp_bot=300.0, wn_lo=500.0, wn_hi=560.0)
        temper = np.linspace(200, 290, L + 1)[None] + r.uniform(-5, 5, (n, L + 1))
    alt = None
    if spherical:
        f |= 2
        kw["radius"] = 6371.0
        alt = np.repeat(np.linspace(L * 1.0, 0.0, L + 1)[None], n, axis=0)
    sc = scalars(n, **kw)
    sc[:, G.S_BEAM_PHI] = r.uniform(0, 30, n)
    rp = None
    if rpv:
        f |= G.G_BRDF_RPV
        rp = np.tile([0.027, 0.647, -0.169, 0.0, 0.0, 0.0, 1.0], (n, 1))
        rp[n // 2:, 0] = 0.1
    tu = None
    if user_tau:
        f |= G.G_USER_TAU
        tot = dtau.sum(axis=1)
        tu = np.stack([np.array([0.0, 0.3, 0.5, 0.9, 1.0]) * t for t in tot])
        tu[:, 2] = np.cumsum(dtau, axis=1)[:, L // 2]  # a layer boundary
    if corr:
        f |= G.G_CORR_BURAS | G.G_CORR_NAKAJIMA
    mu = np.sort(np.concatenate([-np.linspace(0.1, 1.0, numu // 2), np.linspace(0.12, 0.98, numu - numu // 2)]))
    phi = np.linspace(0.0, 300.0, nphi)
    return G.GeneralBatch(flags=f, nstr=nstr, dtau=dtau, ssa=ssa, pmom=pmom, scal=sc, temper=temper, alt=alt,
                          rpv=rp, tau_user=tu, mu_user=mu, phi_user=phi, accuracy=accuracy)


@pytest.mark.parametrize("nstr", [4, 6, 8, 16, 32])
def test_random_lambertian_beam_thermal(ctx, nstr):
    solve_both(ctx, random_general(4, 9, nstr, seed=10 + nstr, thermal=True, user_tau=True),
               f"lambertian beam+thermal nstr={nstr}")


@pytest.mark.parametrize("nstr", [8, 16, 32])
def test_random_rpv(ctx, nstr):
    solve_both(ctx, random_general(4, 7, nstr, seed=30 + nstr, rpv=True, corr=True, user_tau=True),
               f"RPV + corrections nstr={nstr}")


@pytest.mark.parametrize("nstr", [8, 32])
def test_random_spherical_rpv_nt(ctx, nstr):
    solve_both(ctx, random_general(3, 10, nstr, seed=50 + nstr, rpv=True, spherical=True, corr=True),
               f"spherical + RPV + NT nstr={nstr}")


def test_accuracy_cut_and_degenerate_angles(ctx):
    gb = random_general(4, 6, 16, seed=70, accuracy=1e-3, thermal=True)
    gb.scal[1, 1] = 1.0  # overhead sun: azimuth-independent column
    gb.scal[2, 0] = 0.0  # no beam
    gb.mu_user[:] = np.array([-1.0, -0.5, -0.2, 0.3, 0.6, 1.0])
    solve_both(ctx, gb, "accuracy cut, overhead sun, no beam")


@pytest.mark.parametrize("nstr,extra", [(16, {}), (32, {"rpv": True, "spherical": True, "corr": True}),
                                         (64, {})])
def test_layer_cut_off(ctx, nstr, extra):
    # Layers 3.. absorb (abstau >= 10 below layer 4): the tasks of the layers
    # below the cut-off do not run, their levels read zero.
    gb = random_general(2, 8, nstr, seed=80 + nstr, **extra)
    gb.dtau[:, 3:] = 30.0
    gb.ssa[:, 3:] = 0.05
    solve_both(ctx, gb, f"absorption cut-off nstr={nstr}")


@pytest.mark.parametrize("numu", [3, 20, 40, 70])
def test_user_angle_counts(ctx, numu):
    # Up to eight (column, mode, layer) tasks share a wavefront of the
    # user-angle coefficient kernel, or one task steps through 64-angle strides.
    gb = random_general(2, 5, 8, seed=95 + numu, numu=numu, nphi=3)
    solve_both(ctx, gb, f"numu={gb.numu}")


def test_delta_m_plus(ctx):
    gb = random_general(3, 6, 16, seed=90, flags=4, P=19, corr=True)
    solve_both(ctx, gb, "Delta-M+")


def test_only_fluxes_and_no_user_mu(ctx):
    gb = random_general(3, 6, 8, seed=95, thermal=True)
    gb2 = G.GeneralBatch(flags=(gb.flags & ~G.G_USER_MU) | G.G_ONLY_FLUXES, nstr=8, dtau=gb.dtau, ssa=gb.ssa,
                         pmom=gb.pmom, scal=gb.scal, temper=gb.temper)
    solve_both(ctx, gb2, "comp_only_fluxes")
    gb3 = G.GeneralBatch(flags=gb.flags & ~G.G_USER_MU, nstr=8, dtau=gb.dtau, ssa=gb.ssa, pmom=gb.pmom,
                         scal=gb.scal, temper=gb.temper, phi_user=[0.0, 90.0])
    g3, _ = solve_both(ctx, gb3, "no user angles")
    assert (g3.inten == 0).all()


def test_wide_64(ctx):
    solve_both(ctx, random_general(1, 4, 64, seed=99, rpv=True, corr=True, numu=4, nphi=2), "nstr=64")


@pytest.mark.parametrize("nstr", [32, 64])
def test_spherical_rpv_user_tau_wide(ctx, nstr):
    # One-wavefront boundary kernel at 32 streams, the two-block form
    # (3NN > 64 rows) at 64; spherical beam, RPV, corrections, user depths.
    gb = random_general(2, 5, nstr, seed=120 + nstr, rpv=True, spherical=True, corr=True, user_tau=True,
                        numu=4, nphi=3)
    solve_both(ctx, gb, f"spherical RPV user-tau nstr={nstr}")


@pytest.mark.parametrize("first", [0, 1000])
def test_c3_workload(ctx, first):
    # The C3 shape itself: 80 layers, 32 streams, Delta-M+ requested,
    # pseudo-spherical beam with mu0 down to 0.05, a distinct RPV surface per
    # column, 16 user polar angles x 8 azimuths at all 81 levels,
    # Nakajima-Tanaka correction, every Fourier mode.
    from disort_amd import workloads

    gb = workloads.c3(8, first=first)
    assert len({tuple(r) for r in gb.rpv}) == gb.n
    solve_both(ctx, gb, f"C3 columns {first}..{first + gb.n - 1}")


@pytest.mark.parametrize("nstr", [8, 16, 64])
def test_rpv_thermal_distinct_surfaces(ctx, nstr):
    # Thermal emission with an RPV surface: the directional emissivities
    # (200 x 100 azimuth double sums) of a distinct surface per column.
    gb = random_general(4, 5, nstr, seed=140 + nstr, rpv=True, thermal=True, numu=6, nphi=3)
    r = np.random.default_rng(nstr)
    gb.rpv[:, 0] = r.uniform(0.01, 0.3, gb.n)
    gb.rpv[:, 1] = r.uniform(0.5, 1.5, gb.n)
    gb.rpv[:, 2] = r.uniform(-0.3, 0.0, gb.n)
    solve_both(ctx, gb, f"RPV + thermal, distinct surfaces nstr={nstr}")


@pytest.mark.parametrize("name", ["disotest_4a", "disotest_4b", "disotest_4c"])
def test_garcia_siewert_haze(ctx, name):
    # Haze-L moments from the facade's fillHazeGarciaSiewert, 32 streams.
    from test_garcia_siewert import CASES, case_batch as gs_batch

    solve_both(ctx, gs_batch(next(c for c in CASES if c[0] == name), n=2), name)


def tabulated_hg(gb, g, nphase, seed):
    """Buras-Emde inputs: each layer's phase function tabulated at nphase
    increasing cosines (denser near forward scattering), HG of the layer's
    asymmetry g with a Rayleigh admixture, plus the matching moments."""
    r = np.random.default_rng(seed)
    x = np.sort(np.concatenate([np.linspace(-1.0, 0.89, nphase // 2), 1.0 - np.geomspace(1e-4, 0.1, nphase - nphase // 2)]))
    assert (np.diff(x) > 0).all()
    x[-1] = 1.0
    w = r.uniform(0.0, 0.3, g.shape)
    hg = (1 - g[..., None] ** 2) / (1 + g[..., None] ** 2 - 2 * g[..., None] * x) ** 1.5
    ray = 0.75 * (1 + x ** 2)
    phase = w[..., None] * ray + (1 - w[..., None]) * hg
    k = np.arange(gb.pstride)
    mom = (1 - w[..., None]) * g[..., None] ** k
    mom[..., 0] += w
    mom[..., 2] += 0.1 * w
    gb.pmom[:] = mom
    gb.mu_phase = x
    gb.nphase = x.size
    gb.phase = np.ascontiguousarray(phase)
    return gb


@pytest.mark.parametrize("nstr,extra", [(8, {}), (16, {"user_tau": True}), (32, {"spherical": True}),
                                         (16, {"thermal": True, "rpv": True})])
def test_buras_emde_random(ctx, nstr, extra):
    # Buras-Emde correction: tabulated phase function, intensity_corr_buras
    # without intensity_corr_nakajima; user angles in the 10-degree aureole of
    # the beam (secondary-scattering term) and away from it.
    n, L = 3, 5
    gb = random_general(n, L, nstr, seed=160 + nstr, corr=True, numu=8, nphi=3, mu0=np.array([0.8, 0.55, 1.0]),
                        P=nstr + 8, **extra)
    gb.flags &= ~G.G_CORR_NAKAJIMA
    gb.mu_user[:] = np.array([-1.0, -0.81, -0.56, -0.3, 0.2, 0.5, 0.8, 1.0])
    gb.phi_user[:] = np.array([0.0, 5.0, 120.0])
    g = np.random.default_rng(nstr).uniform(0.6, 0.9, (n, L))
    tabulated_hg(gb, g, 181, seed=nstr)
    solve_both(ctx, gb, f"Buras-Emde nstr={nstr} {sorted(extra)}")


def test_buras_emde_layer_cut(ctx):
    gb = random_general(2, 8, 16, seed=171, corr=True, numu=6, nphi=2, mu0=np.array([0.9, 0.7]), P=24)
    gb.flags &= ~G.G_CORR_NAKAJIMA
    gb.dtau[:, 3:] = 30.0
    gb.ssa[:, 3:] = 0.05
    gb.mu_user[:] = np.array([-0.95, -0.9, -0.7, 0.3, 0.6, 1.0])
    tabulated_hg(gb, np.full((2, 8), 0.8), 100, seed=3)
    solve_both(ctx, gb, "Buras-Emde with the absorption cut-off")


BRDF_KINDS = {"hapke": (G.G_BRDF_HAPKE, lambda r: [r.uniform(0.5, 1.5), r.uniform(0.02, 0.1), r.uniform(0.3, 0.9)]),
              "coxmunk": (G.G_BRDF_COXMUNK, lambda r: [r.uniform(2.0, 15.0), r.uniform(1.3, 1.4), float(r.integers(2))]),
              "rossli": (G.G_BRDF_ROSSLI, lambda r: [r.uniform(0.05, 0.2), r.uniform(0.0, 0.05), r.uniform(0.0, 0.03)])}


@pytest.mark.parametrize("kind,nstr,extra", [("hapke", 8, {"thermal": True}), ("coxmunk", 16, {"user_tau": True}),
                                              ("rossli", 32, {"spherical": True}), ("hapke", 64, {}),
                                              ("coxmunk", 8, {"thermal": True}), ("rossli", 16, {"thermal": True})])
def test_random_brdf_kinds(ctx, kind, nstr, extra):
    # Hapke, Cox-Munk and Ross-Li surfaces with their own parameters per
    # column: BRDF Fourier coefficients, emissivities (thermal), the BRDF
    # beam-reflection and Nakajima-Tanaka corrections.
    flag, draw = BRDF_KINDS[kind]
    gb = random_general(3, 4, nstr, seed=200 + nstr, rpv=True, corr=True, numu=6, nphi=3, **extra)
    gb.flags = (gb.flags & ~G.G_BRDF_RPV) | flag
    r = np.random.default_rng(nstr)
    gb.rpv[:] = 0.0
    for i in range(gb.n):
        p = draw(r)
        gb.rpv[i, :len(p)] = p
    solve_both(ctx, gb, f"{kind} surface nstr={nstr} {sorted(extra)}")


def test_validation_messages(ctx):
    from disort_amd._capi import DisortInvalid

    gb = random_general(2, 4, 8, seed=1)
    gb.mu_user[:] = gb.mu_user[::-1].copy()
    with pytest.raises(DisortInvalid, match="mu_user must be in increasing order"):
        ctx.general_solve(gb)
    gb = random_general(2, 4, 8, seed=1)
    gb.ssa[1, 2] = 1.5
    with pytest.raises(DisortInvalid, match=r"column 1: single_scat_albedo\[2\] must be in \[0, 1\]"):
        ctx.general_solve(gb)
