"""MI355X-native batched DISORT (flux path of DisORT++'s DisortFluxSolver<N>).

The compute path is the HIP library ``disort_amd/lib/libdisort_amd.so``
behind the C ABI of ``include/disort_amd.h``; Python reaches it through
ctypes (``disort_amd._capi``) and, as a drop-in for the reference's pybind11
module, through ``disortpp``.
"""
from __future__ import annotations

import os

from .batch import (  # noqa: F401
    F_DELTA_M_PLUS,
    F_DIFFUSION_BC,
    F_THERMAL,
    OUT_FIELDS,
    FluxBatch,
    scalars,
    uniform_batch,
)

__version__ = "0.1.0"

_HERE = os.path.dirname(os.path.abspath(__file__))


def library_path() -> str:
    return os.path.join(_HERE, "lib", "libdisort_amd.so")
