"""Write adjoint_small_fd.json: central differences of the oracle's mean drag for the
rotating, blowing cylinder of the adjoint tests (the oracle is the FP64 CPU solver).

Case rotatingCylinderRe100_small_luo (114 x 114, 256 points, Re = 100), 40 steps, mean
C_D over steps 21..40, solves to 1e-12; surface velocity alpha (D / 2) omega_ref along the
tangent (alpha = 0.5, omega_ref = 2, solid nodes turning at alpha omega_ref) plus normal
blowing from numpy's default_rng(7).normal(0, 0.01, 256). Stored: J, dJ/dalpha by
central differences with h = 1e-6 and 1e-5, and the derivatives along the unit vectors
default_rng(11).normal(size=256) / norm and default_rng(12).normal(size=256) / norm over
the blowing values (h = 1e-6).

Run from the repository root: python tests/golden/make_adjoint_fixtures.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "oracle")]
import pyoracle as po  # noqa: E402

CASE = os.path.join(ROOT, "cuibm_amd", "cases", "rotatingCylinderRe100_small_luo")
TIGHT = {"velocityTol": 1e-12, "poissonTol": 1e-12}
NSTEPS, N0, N1 = 40, 21, 40
ALPHA, OMEGA_REF, D = 0.5, 2.0, 1.0


def mean_drag(alpha, blow):
    case = po.read_case(CASE, TIGHT)
    o = po.Oracle(case)
    bx = np.concatenate([b[0] for b in case["bodies"]])
    by = np.concatenate([b[1] for b in case["bodies"]])
    th = np.arctan2(by - by.mean(), bx - bx.mean())
    vt = alpha * 0.5 * D * OMEGA_REF
    o.set_body_velocity(-vt * np.sin(th) + blow * np.cos(th), vt * np.cos(th) + blow * np.sin(th),
                        alpha * OMEGA_REF)
    J = 0.0
    for k in range(1, NSTEPS + 1):
        o.step()
        if N0 <= k <= N1:
            J += 2 * o.info()["forceX"]
    o.close()
    return J / (N1 - N0 + 1)


def main():
    blow = np.random.default_rng(7).normal(0.0, 0.01, 256)
    d = np.random.default_rng(11).normal(size=256)
    d /= np.linalg.norm(d)
    d2 = np.random.default_rng(12).normal(size=256)
    d2 /= np.linalg.norm(d2)
    out = {"case": "rotatingCylinderRe100_small_luo", "nsteps": NSTEPS, "n0": N0, "n1": N1, "alpha": ALPHA,
           "J": mean_drag(ALPHA, blow)}
    for h in (1e-6, 1e-5):
        out[f"dJ_dalpha_h{h:g}"] = (mean_drag(ALPHA + h, blow) - mean_drag(ALPHA - h, blow)) / (2 * h)
        print(h, out[f"dJ_dalpha_h{h:g}"], flush=True)
    h = 1e-6
    out["dJ_dir"] = (mean_drag(ALPHA, blow + h * d) - mean_drag(ALPHA, blow - h * d)) / (2 * h)
    out["dJ_dir2"] = (mean_drag(ALPHA, blow + h * d2) - mean_drag(ALPHA, blow - h * d2)) / (2 * h)
    print("directional", out["dJ_dir"], out["dJ_dir2"], flush=True)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "adjoint_small_fd.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
