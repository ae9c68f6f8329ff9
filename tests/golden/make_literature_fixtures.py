#!/usr/bin/env python3
"""Write tests/golden/literature.json from the benchmark data the reference validates
against (validation-data/, read as text; data only):

  cavity-GGS82.txt        Ghia, Ghia & Shin (1982), lid-driven cavity centre lines.
                          Column 0 = y and column 1 = u along x = 0.5 for Re = 100;
                          column 5 = x and column 6 = v along y = 0.5 for Re = 100
                          (the columns scripts/validation/validateCavity.py:24-40 picks)
  cylinderRe40-KL95.txt   Koumoutsakos & Leonard (1995), impulsively started cylinder,
  cylinderRe550-KL95.txt  drag coefficient against time in their units (the reference
                          plots them at time * 0.5: scripts/validation/validateCylinder.py:57)

Run from the repository root, in a checkout that has the reference under /root/reference:
    python tests/golden/make_literature_fixtures.py
"""
import json
import os

import numpy as np

REF = "/root/reference/validation-data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "literature.json")


def main():
    g = np.loadtxt(os.path.join(REF, "cavity-GGS82.txt"))
    out = {
        "source": "validation-data/{cavity-GGS82.txt, cylinderRe40-KL95.txt, cylinderRe550-KL95.txt}",
        "ghia82_re100": {"y": g[:, 0].tolist(), "u": g[:, 1].tolist(), "x": g[:, 5].tolist(), "v": g[:, 6].tolist()},
    }
    for re in (40, 550):
        k = np.loadtxt(os.path.join(REF, f"cylinderRe{re}-KL95.txt"))
        out[f"kl95_re{re}"] = {"t": k[:, 0].tolist(), "cd": k[:, 1].tolist()}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
