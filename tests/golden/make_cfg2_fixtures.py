#!/usr/bin/env python3
"""Write tests/golden/cfg2_<method>_series.json: the oracle's run of the 336x336 impulsively
started cylinder at Re = 40 (dt = 0.01, 2000 steps, solver tolerances 1e-8 as configured)
with the Luo external (cylinderRe40_luo) or the modified Fadlun (cylinderRe40_fadlun) method.

Kept: drag and lift (2 Fx, 2 Fy) every 5 steps, and every 200 steps the L2 norms of u and
of p minus its mean and u at 16 fixed nodes around the body and in the wake. The GPU tests
compare a full 2000-step run against these (the oracle needs about 12 minutes per method on
8 cores, longer than a GPU test may take).

    python tests/golden/make_cfg2_fixtures.py luo|fadlun
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import pyoracle as po  # noqa: E402
from helpers import case  # noqa: E402

STEPS, EVERY, SNAP = 2000, 5, 200


def probes(o):
    """indices of 16 u nodes: a ring at radius 0.6-1.0 around the body and the wake axis"""
    x, y, dx = o.get("x"), o.get("y"), o.get("dx")
    xu = x[:-1] + 0.5 * dx[:-1]
    nx = len(x)
    pts = [(0.8 * np.cos(a), 0.8 * np.sin(a)) for a in np.linspace(0, 2 * np.pi, 8, endpoint=False)]
    pts += [(1.0 + 0.7 * k, 0.0) for k in range(4)] + [(2.0, 0.5), (2.0, -0.5), (4.0, 1.0), (-2.0, 0.0)]
    return [int(np.searchsorted(y, py) * (nx - 1) + np.searchsorted(xu, px)) for px, py in pts]


def main(method):
    name = {"luo": "cylinderRe40_luo", "fadlun": "cylinderRe40_fadlun"}[method]
    o = po.Oracle(po.read_case(case(name)))
    idx = probes(o)
    out = {"case": name, "steps": STEPS, "every": EVERY, "snap": SNAP, "probe_index": idx,
           "cd": [], "cl": [], "u_norm": [], "p_norm": [], "u_probe": []}
    for k in range(1, STEPS + 1):
        o.step(1)
        if k % EVERY == 0:
            f = o.info()
            out["cd"].append(2 * f["forceX"])
            out["cl"].append(2 * f["forceY"])
        if k % SNAP == 0:
            u, p = o.get("u"), o.get("p")
            out["u_norm"].append(float(np.linalg.norm(u)))
            out["p_norm"].append(float(np.linalg.norm(p - p.mean())))
            out["u_probe"].append(u[idx].tolist())
            print(method, k, out["cd"][-1], flush=True)
    with open(os.path.join(HERE, f"cfg2_{method}_series.json"), "w") as fh:
        json.dump(out, fh)


if __name__ == "__main__":
    main(sys.argv[1])
