import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def printed_ulp(v: float) -> float:
    """One unit in the last printed digit of a fixture literal (the shortest
    repr of a float reproduces the decimal literal it was typed as)."""
    import math

    r = repr(abs(float(v)))
    mant, _, _ = r.partition("e")
    digits = len(mant.replace(".", "").lstrip("0"))
    return 10.0 ** (math.floor(math.log10(abs(v))) - (digits - 1))


def load_golden(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


def case_batch(case):
    """FluxBatch of one instance from a reference_cases.json entry."""
    from disort_amd.batch import scalars, uniform_batch

    L, N, P = case["nlyr"], case["nstr"], case["pstride"]
    pm = np.zeros((1, L, P))
    for lc, row in enumerate(case["pmom"]):
        pm[0, lc, : len(row)] = row[:P]
    temper = None if case["temper"] is None else np.array([case["temper"]], dtype=np.float64)
    return uniform_batch(
        np.array([case["dtau"]], dtype=np.float64),
        np.array([case["ssa"]], dtype=np.float64),
        pm,
        N,
        scalars(1, **case["scal"]),
        flags=case["flags"],
        temper=temper,
    )


@pytest.fixture(scope="session")
def reference_cases():
    return load_golden("reference_cases.json")


def max_rel_err(a, b, floor=1e-12):
    """max |a-b| / max(|b|, floor * max|b|), the parity metric of the tests."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.maximum(np.abs(b), floor * max(np.abs(b).max(), 1e-300))
    return float(np.max(np.abs(a - b) / scale))
