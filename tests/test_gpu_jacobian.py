"""GPU temperature Jacobians (disort_flux_jacobian) against the oracle's
analytic linearisation (DisortFluxSolver::computeThermalJacobian,
src/FluxSolver.cpp:1294-1497), which test_oracle_jacobian.py pins against
finite differences of the forward solve.

Criterion: |gpu - ref| <= 1e-9 |ref| + 1e-12 max|ref of the instance|, the
flux workload's 1e-9 relative bound with its absolute floor taken relative
to the instance's largest derivative (the entries span many decades: the
top-of-atmosphere flux barely depends on the deepest level).  An entry
outside it is arbitrated against the same linearisation in x87 extended
precision x: it passes only when the GPU is within the bound of x, or no
further from x than 4 times the FP64 oracle's largest error on the same
degree of freedom (the outputs of one unit-source solve).  Thin layers deep in the column make the
reference's algorithm ill-conditioned (the Planck intercept B_k - s1 tau_top
with s1 = 1/dtau cancels ~tau_top/dtau digits; unit-source solves hit it
hardest), and there FP64 itself cannot meet the bound."""
import numpy as np
import pytest

import oracle_lib as ol
from disort_amd import workloads as W
from disort_amd.batch import F_DIFFUSION_BC, F_THERMAL, FluxBatch
from test_oracle_general_jacobian import CASES as GENERAL_CASES, general_batch
from test_oracle_jacobian import CASES, thermal_batch

pytestmark = pytest.mark.gpu

RTOL = 1e-9
ATOL_REL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    from disort_amd._capi import Context

    return Context(0)


def assert_jac_parity(got, ref, label, batch=None, i=0):
    assert np.isfinite(ref).all() and np.isfinite(got).all(), label
    floor = ATOL_REL * np.abs(ref).max()
    bad = ~(np.abs(got - ref) <= RTOL * np.abs(ref) + floor)
    if bad.any() and batch is not None:
        _, x = ol.flux_jacobian_one_extended(batch, i)
        col_err = np.abs(ref - x).max(axis=(0, 1), keepdims=True)  # [1, 1, dof]
        ok = np.abs(got - x) <= np.maximum(RTOL * np.abs(x) + floor, 4.0 * col_err)
        bad &= ~ok
    if bad.any():
        f, lu, k = np.argwhere(bad)[0]
        raise AssertionError(f"{label}: {int(bad.sum())} entries off, e.g. field {f} level {lu} dof {k}: "
                             f"gpu {got[f, lu, k]!r} oracle {ref[f, lu, k]!r}")


def concat(batches):
    """One ragged batch from one-instance batches of any shape."""
    cat = lambda name: np.concatenate([getattr(b, name) for b in batches])  # noqa: E731
    return FluxBatch(flags=batches[0].flags, nstr=cat("nstr"), nlyr=cat("nlyr"), pstride=cat("pstride"),
                     dtau=cat("dtau"), ssa=cat("ssa"), pmom=cat("pmom"), scal=cat("scal"), temper=cat("temper"))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_gpu_jacobian_reference_problems(ctx, case):
    fb = thermal_batch(**case[1])
    print(f"[jacobian] case {case[0]} nstr {int(fb.nstr[0])} nlyr {int(fb.nlyr[0])}", flush=True)
    got = ctx.flux_jacobian(fb)[0]
    _, ref = ol.flux_jacobian_one(fb, 0)
    assert_jac_parity(got, ref, case[0], fb, 0)


def test_gpu_jacobian_ragged_batch(ctx):
    """Every plain-boundary case in one batch of mixed stream and layer counts."""
    bs = [thermal_batch(**c[1]) for c in CASES if not (c[1].get("flags", 0) & ~F_THERMAL)]
    fb = concat(bs)
    print(f"[jacobian] ragged batch of {fb.n}", flush=True)
    got = ctx.flux_jacobian(fb)
    for i, b in enumerate(bs):
        _, ref = ol.flux_jacobian_one(b, 0)
        assert_jac_parity(got[i], ref, f"instance {i}", b, 0)


@pytest.mark.parametrize("nstr,nlyr,extra", [(8, 20, 0), (16, 12, F_DIFFUSION_BC), (32, 6, 0)])
def test_gpu_jacobian_random_columns(ctx, nstr, nlyr, extra):
    """Random columns of the banded workload's distribution (optical depth
    1e-4..10, narrow Planck bands, beam on)."""
    fb = W.random_flux_batch(24, nlyr, nstr, seed=7 + nstr, flags_extra=extra)
    print(f"[jacobian] random nstr {nstr} nlyr {nlyr} flags {fb.flags}", flush=True)
    got = ctx.flux_jacobian(fb)
    for i in range(fb.n):
        _, ref = ol.flux_jacobian_one(fb, i)
        assert_jac_parity(got[i], ref, f"column {i}", fb, i)


def test_gpu_jacobian_forward_solve_unchanged(ctx):
    """The Jacobian call leaves the caller's batch and a later forward solve alone."""
    fb = W.random_flux_batch(8, 10, 8, seed=3)
    ctx.solve(fb)
    before = fb.out.copy()
    ctx.flux_jacobian(fb)
    ctx.solve(fb)
    assert np.array_equal(fb.out, before)


def test_gpu_jacobian_needs_thermal(ctx):
    from disort_amd._capi import DisortError

    fb = W.random_flux_batch(2, 4, 8, seed=1, thermal=False)
    with pytest.raises(DisortError, match="thermal"):
        ctx.flux_jacobian(fb)


@pytest.mark.parametrize("case", GENERAL_CASES, ids=[c[0] for c in GENERAL_CASES])
def test_gpu_general_jacobian(ctx, case):
    """disort_general_jacobian (DisortSolver::computeThermalJacobian) on the
    general oracle's problems: user optical depths, BRDF surfaces, diffusion
    lower boundary, Delta-M."""
    gb = general_batch(**case[1])
    print(f"[jacobian] general case {case[0]}", flush=True)
    got = ctx.general_jacobian(gb)[0]
    _, ref = ol.general_jacobian_one(gb, 0)
    assert_jac_parity(got, ref, case[0])


def test_gpu_general_jacobian_columns(ctx):
    """Several columns with their own temperatures and surfaces in one call."""
    r = np.random.default_rng(5)
    gb = general_batch([215.0, 240.0, 262.0, 281.0], 290.0, 16, 0.5, brdf="rpv", n=5,
                       tau_user=np.tile([0.0, 0.5, 0.9, 1.2], (5, 1)), beam=1.0)
    gb.temper[:] += r.uniform(-10, 10, gb.temper.shape)
    gb.scal[:, 5] = r.uniform(270, 310, gb.n)
    gb.rpv[:, 0] = r.uniform(0.02, 0.2, gb.n)
    print("[jacobian] general columns", flush=True)
    got = ctx.general_jacobian(gb)
    for i in range(gb.n):
        _, ref = ol.general_jacobian_one(gb, i)
        assert_jac_parity(got[i], ref, f"column {i}")
