"""Algorithmic work model of one flux-only DISORT instance (FP64).

One instance = one DisortFluxSolver<N>::solve (src/FluxSolver.cpp:23-171),
N = streams, L = layers, NN = N/2, n = N*L unknowns, w = 3NN-1 half
bandwidth, 1 FMA = 2 FLOP.  These are the operation counts of the
reference's algorithm (N x N LU per particular solution, LINPACK band LU
without pivoting growth), not of any particular implementation, so that a
roofline fraction computed from them is comparable across implementations.
Transcendentals (exp, sqrt) are not counted.
"""
from __future__ import annotations


def stage_flops(nstr: int, nlyr: int, thermal: bool = True, beam: bool = True) -> dict:
    N, L = nstr, nlyr
    NN = N // 2
    n = N * L
    w = 3 * NN - 1
    f = {
        # per (instance, layer): the eigen-system and particular solutions
        "phase_matrix": L * (2 * NN * N * N + 3 * NN * N),
        "eigenproblem": L * ((50.0 / 3.0) * NN**3 + 12 * NN * NN),
        "beam_source": L * ((2.0 / 3.0) * N**3 + 5 * N * N + N) if beam else 0.0,
        "thermal_source": L * ((2.0 / 3.0) * N**3 + 5 * N * N + 3 * N) if thermal else 0.0,
        "delta_m": 2 * L * 4 * N,
        # per instance: boundary-value problem
        "band_assembly": L * 2 * N * N,
        "band_lu": 2 * n * w * w,
        "band_solve": 6 * n * w,
        # per level
        "flux_assembly": (L + 1) * (2 * N * N + 8 * N),
        "planck": 40 * (L + 3) if thermal else 0.0,
    }
    return f


# Which kernel of the GPU path carries which model stages.
KERNEL_STAGES = {
    "k_column": ("planck",),
    "k_layer": ("phase_matrix", "eigenproblem", "beam_source", "thermal_source", "delta_m"),
    "k_boundary": ("band_assembly", "band_lu", "band_solve"),
    "k_flux": ("flux_assembly",),
}


def kernel_flops(nstr: int, nlyr: int, thermal: bool = True, beam: bool = True) -> dict:
    s = stage_flops(nstr, nlyr, thermal, beam)
    return {k: float(sum(s[x] for x in v)) for k, v in KERNEL_STAGES.items()}


def instance_flops(nstr: int, nlyr: int, thermal: bool = True, beam: bool = True) -> float:
    return float(sum(stage_flops(nstr, nlyr, thermal, beam).values()))


def instance_bytes(nstr: int, nlyr: int, pstride: int, thermal: bool = True) -> int:
    """Algorithmic HBM bytes per instance: inputs read once, outputs written once."""
    L = nlyr
    inputs = 8 * (2 * L + L * pstride + (L + 1 if thermal else 0) + 12)  # dtau, ssa, pmom, T, scalars
    outputs = 8 * 8 * (L + 1)
    return inputs + outputs


def general_mode_flops(nstr: int, nlyr: int, m: int, numu: int, ntau: int, ncut: int | None = None,
                       thermal: bool = False, beam: bool = True) -> dict:
    """Work of one azimuthal mode m of one column of the general solver
    (DisortSolver::solve, src/DisortSolver.cpp:218-350), by GPU kernel.
    Legendre sums run over l = m..N-1; the per-layer work (eigen-system,
    particular solutions, interpolation to the user angles) and the boundary
    problem over the ncut layers above the absorption cut-off (a layer below
    it enters no output); user intensities integrate on average half of the
    layers per (level, angle)."""
    N, L = nstr, nlyr
    NN = N // 2
    A = N - m  # active Legendre orders
    ncut = L if ncut is None else ncut
    n = N * ncut
    w = 3 * NN - 1
    layer = ncut * (2 * NN * NN * A + 3 * NN * N         # A_s, -B_s from the scaled tables
                 + (50.0 / 3.0) * NN**3 + 12 * NN * NN   # reduced eigenproblem
                 + ((2.0 / 3.0) * N**3 + 5 * N * N + N if beam else 0.0)
                 + ((2.0 / 3.0) * N**3 + 5 * N * N + 3 * N if thermal and m == 0 else 0.0)
                 + 2 * A * N * NN + 2 * numu * N * A      # eigenvectors at the user angles
                 + 2 * 3 * A * NN + 2 * 3 * numu * A)     # sources at the user angles
    boundary = ncut * 2 * N * N + 2 * n * w * w + 6 * n * w
    user = ncut * numu * (N * 10.0) + ntau * numu * ((L / 2.0) * 8.0 + N * 12.0)
    return {"k_glayer": float(layer), "k_gboundary": float(boundary), "k_guser": float(user)}


def general_column_flops(nstr: int, nlyr: int, numu: int, ntau: int, ncut: int | None = None,
                         thermal: bool = False, beam: bool = True) -> dict:
    tot = {"k_glayer": 0.0, "k_gboundary": 0.0, "k_guser": 0.0}
    for m in range(nstr):
        for k, v in general_mode_flops(nstr, nlyr, m, numu, ntau, ncut, thermal, beam).items():
            tot[k] += v
    return tot
