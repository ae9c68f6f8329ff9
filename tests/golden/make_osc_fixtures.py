#!/usr/bin/env python3
"""Write tests/golden/osc_reference_series.json from the reference's committed output
of the in-line oscillating cylinder run (validation/osc/static/external: Re 100,
Luo external method, dt 0.001, x = -sin(2 pi t)/(2 pi), u = -cos(2 pi t)).

Data only, as the reference's run wrote it (text columns, 6 significant digits):

  midPosition_init   ts, X, Y, U, V of the body centre after each step
  forces_init        t, Fx, Fy after each step
  iterations_init    step, velocity iterations, Poisson iterations (each row twice)

The first NSTEPS steps are kept. Run from the repository root, in a checkout that has
the reference under /root/reference:  python tests/golden/make_osc_fixtures.py
"""
import json
import os

import numpy as np

REF = "/root/reference/validation/osc/static/external"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "osc_reference_series.json")
NSTEPS = 1000


def main():
    mid = np.loadtxt(os.path.join(REF, "midPosition_init"), skiprows=1)[:NSTEPS]
    frc = np.loadtxt(os.path.join(REF, "forces_init"), skiprows=1)[:NSTEPS]
    its = np.loadtxt(os.path.join(REF, "iterations_init"))[::2][:NSTEPS]
    assert np.array_equal(mid[:, 0], np.arange(1, NSTEPS + 1))
    assert np.array_equal(its[:, 0], np.arange(1, NSTEPS + 1))
    out = {
        "source": "validation/osc/static/external/{midPosition_init,forces_init,iterations_init}",
        "steps": NSTEPS,
        "midX": mid[:, 1].tolist(),
        "midU": mid[:, 3].tolist(),
        "forceX": frc[:, 1].tolist(),
        "forceY": frc[:, 2].tolist(),
        "iter_velocity": its[:, 1].astype(int).tolist(),
        "iter_poisson": its[:, 2].astype(int).tolist(),
    }
    with open(OUT, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
