"""Shared helpers of the parity tests (GPU library vs. oracle)."""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CASES = os.path.join(ROOT, "cuibm_amd", "cases")


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def case(name):
    return os.path.join(CASES, name)


def pair(name, overrides=None, **kw):
    """GPU solver and oracle for the same case and overrides (dict of CLI flags)."""
    import cuibm_amd
    import pyoracle

    overrides = overrides or {}
    args = []
    for k, v in overrides.items():
        args += [f"-{k}", str(v)]
    g = cuibm_amd.Solver(case(name), *args, **kw)
    g.initialise()
    o = pyoracle.Oracle(pyoracle.read_case(case(name), {k: float(v) if k != "nt" else int(v)
                                                        for k, v in overrides.items()}))
    return g, o
