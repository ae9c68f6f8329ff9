#!/usr/bin/env python3
"""Write tests/golden/reference_grids.json from the reference's own case files.

Data only: for each case the domain parameters (start, mid_h, subdomain ends and
stretch ratios) and body options are read from the reference's YAML files, and the
expected cells per side are the numbers the reference itself records:

  scripts/validation/performance.py:22  grid_cylinder_750ti =
      [170,208,450,890, 135,208,450,890, 162,352,372,462]
  for validation/error/cylinder/{fadlun,external,embedded}{1..4}

and the 224 cells per side of the in-line oscillating cylinder runs
(scripts/validation/performance2.py:84-86, validation/error/oscflow2/*).

Run from the repository root, in a checkout that has the reference under
/root/reference:  python tests/golden/make_grid_fixtures.py
"""
import json
import os

import yaml

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_grids.json")

# scripts/validation/performance.py:22 (cells per side, in this order of cases)
PERF_GRID = [170, 208, 450, 890, 135, 208, 450, 890, 162, 352, 372, 462]
PERF_CASES = [f"validation/error/cylinder/{m}{k}" for m in ("fadlun", "external", "embedded") for k in (1, 2, 3, 4)]
OSC_CASES = [f"validation/error/oscflow2/{m}" for m in
             ("external0625", "external05", "external03125", "embedded0625", "embedded05", "embedded03125")]


def load(path):
    with open(path) as f:
        return yaml.safe_load(f)


def domain(case):
    out = {}
    for node in load(os.path.join(REF, case, "domain.yaml")):
        out[node["direction"]] = {
            "start": float(node["start"]),
            "mid_h": float(node["mid_h"]),
            "ends": [float(s["end"]) for s in node["subDomains"]],
            "ratios": [float(s["stretchRatio"]) for s in node["subDomains"]],
        }
    return out


def bodies(case):
    p = os.path.join(REF, case, "bodies.yaml")
    if not os.path.exists(p):
        return []
    out = []
    for node in load(p) or []:
        if node.get("type") == "circle":
            out.append([float(v) if not isinstance(v, bool) else v for v in node["circleOptions"]])
    return out


def main():
    cases = []
    for c, n in zip(PERF_CASES, PERF_GRID):
        cases.append({"case": c, "expected_n": n, "source": "scripts/validation/performance.py:22",
                      "domain": domain(c), "bodies": bodies(c)})
    for c in OSC_CASES:
        cases.append({"case": c, "expected_n": 224, "source": "scripts/validation/performance2.py:84-86",
                      "domain": domain(c), "bodies": bodies(c)})
    with open(OUT, "w") as f:
        json.dump({"cases": cases}, f, indent=1)
    print(f"wrote {len(cases)} cases to {OUT}")


if __name__ == "__main__":
    main()
