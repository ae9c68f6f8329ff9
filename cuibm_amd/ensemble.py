"""Case folders for ensembles of impulsively started cylinders of uneven size.

Each case is the Re=40 cylinder set-up of the reference's validation cases (D=1 at the
origin, Dirichlet (1, 0) inflow and sides, convective outflow, Luo external method) on
[-10,10]^2 with a uniform core [-0.6,0.6] of spacing h and geometric stretching outside
it. The edge length and the Reynolds number of every case are drawn from a seeded
generator; the stretch ratio is then chosen so that the reference's grid rule
(src/io/parseDomainFile.cu:41-222: int(log(1 + L (r-1)/h) / log r) cells per stretched
block of length L, int(1.2/h) in the core) gives the drawn edge, to within the rule's
odd totals.
"""
import math
import os

import numpy as np

CORE, OUTER = 0.6, 10.0


def cells_stretched(r, h, length=OUTER - CORE):
    """cells of a block of `length` stretched by r > 1 from spacing h (the reference rule)"""
    return int(math.log(1.0 / (h / (length * (r - 1.0))) + 1.0) / math.log(r))


def core_cells(h):
    return int(2 * CORE / h)


def ratio_for(m, h, lo=1.0 + 1e-9, hi=3.0):
    """a stretch ratio whose blocks have exactly m cells, from the middle of the interval
    of such ratios (cells_stretched decreases in r)"""
    def first_below(k):  # smallest r with cells_stretched(r) < k
        a, b = lo, hi
        for _ in range(200):
            c = 0.5 * (a + b)
            if cells_stretched(c, h) < k:
                b = c
            else:
                a = c
        return b
    r_hi, r_lo = first_below(m), first_below(m + 1)  # cells == m on [r_lo, r_hi)
    r = 0.5 * (r_lo + r_hi)
    if cells_stretched(r, h) != m:
        raise ValueError(f"no stretch ratio gives {m} cells at h={h}")
    return r


def draw(n=64, seed=2024, edges=(96, 512), re=(40.0, 200.0), h=0.025):
    """(edge, Re, stretch ratio) per case; edge = core + 2 x stretched cells"""
    rng = np.random.default_rng(seed)
    out = []
    nc = core_cells(h)
    for _ in range(n):
        target = int(rng.integers(edges[0], edges[1] + 1))
        reynolds = float(rng.uniform(re[0], re[1]))
        m = max(1, int(round((target - nc) / 2.0)))
        r = ratio_for(m, h)
        out.append((nc + 2 * m, reynolds, r))
    return out


def write_cases(folder, n=64, seed=2024, steps=500, dt=0.01, h=0.025, tol=1e-8, edges=(96, 512),
                re=(40.0, 200.0)):
    """write n case folders under `folder`; returns their paths and (edge, Re, ratio) list"""
    cases = draw(n, seed, edges, re, h)
    paths = []
    npts = int(2 * math.pi * 0.5 / h)
    npts += npts % 2
    for k, (edge, reynolds, r) in enumerate(cases):
        d = os.path.join(folder, f"case{k:03d}")
        os.makedirs(d, exist_ok=True)
        axis = lambda name: (f"- direction: {name}\n  start: {-OUTER}\n  mid_h: {h}\n  subDomains:\n"
                             f"    - end: {-CORE}\n      stretchRatio: {r!r}\n"
                             f"    - end: {CORE}\n      stretchRatio: 1.0\n"
                             f"    - end: {OUTER}\n      stretchRatio: {r!r}\n")
        with open(os.path.join(d, "domain.yaml"), "w") as f:
            f.write(axis("x") + "\n" + axis("y"))
        with open(os.path.join(d, "flow.yaml"), "w") as f:
            f.write(f"- type: flow\n  nu: {1.0 / reynolds!r}\n  initialVelocity: [1.0, 0.0]\n  boundaryConditions:\n")
            for loc, kind in (("xMinus", "DIRICHLET"), ("xPlus", "CONVECTIVE"), ("yMinus", "DIRICHLET"),
                              ("yPlus", "DIRICHLET")):
                f.write(f"    - location: {loc}\n      u: [{kind}, 1.0]\n      v: [{kind}, 0.0]\n")
        with open(os.path.join(d, "bodies.yaml"), "w") as f:
            f.write(f"- type: circle\n  circleOptions: [0.0, 0.0, 0.5, {npts}, false]\n")
        with open(os.path.join(d, "simParams.yaml"), "w") as f:
            f.write(f"- type: simulation\n  dt: {dt}\n  nt: {steps}\n  nsave: {steps + 1}\n  scaleCV: 2.0\n"
                    f"  SolverType: LUO\n  linearSolvers:\n"
                    f"    - system: velocity\n      solver: BICGSTAB\n      preconditioner: DIAGONAL\n"
                    f"      tolerance: {tol}\n      maxIterations: 10000\n"
                    f"    - system: Poisson\n      solver: BICGSTAB\n      preconditioner: SMOOTHED_AGGREGATION\n"
                    f"      tolerance: {tol}\n      maxIterations: 20000\n")
        paths.append(d)
    return paths, cases
