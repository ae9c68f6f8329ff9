"""Mixed-precision Poisson solve (-amgPrecision 32: the AMG V-cycle in FP32, its matrix
values, level vectors and smoother arithmetic, with the operator's pinned constant mode
removed and restored in FP64 around it; the Krylov vectors, dots and the projection in
FP64) against the FP64 oracle on the 336x336 Luo cylinder with the stencil finest level. Bar: |dC_D| <= 1e-6 per step (the precision
form's accuracy target); u within 1e-8 relative L2, since both forms solve to the same
residual tolerance and only the preconditioner differs.

A long run at tolerance 1e-12 (the 498x498 rotating cylinder with per-point body
velocities, 650 steps): the FP32 V-cycle's rounding makes it a slightly different map at
every apply, and without restarts BiCGSTAB's recurrence residual drifted from b - A x until
a solve near step 623 returned a wrong pressure and the flow blew up a few steps later.
With the reliable-update restarts (the FP32 V-cycle's default) the run follows the FP64
hierarchy's: drag within 1e-8 at every 50th step and Poisson iterations within 15% of the
FP64 run's in all."""
import numpy as np
import pytest

import cuibm_amd
from cuibm_amd.adjoint import RotatingCylinder
from helpers import case, pair, rel_l2

pytestmark = pytest.mark.gpu


def test_mixed_hierarchy_matches_fp64_oracle_on_cylinder_336():
    g, o = pair("cylinderRe40_luo", {"amgPrecision": 32, "velocityTol": 1e-10, "poissonTol": 1e-10})
    assert g.info()["amg_levels"] >= 3
    for k in range(10):
        g.step_time()
        o.step()
        gi, oi = g.info(), o.info()
        du = rel_l2(g.get("u"), o.get("u"))
        dcd = abs(2.0 * gi["forceX"] - 2.0 * oi["forceX"])
        print(f"step {k}: u {du:.2e} |dC_D| {dcd:.2e} its {gi['iter2']:.0f}/{oi['iter2']:.0f}", flush=True)
        assert dcd <= 1e-6, k
        assert du <= 1e-8, k


def _rotating_run(prec, steps):
    rc = RotatingCylinder(case("rotatingCylinderRe100_luo"), nsteps=steps, n0=0, n1=steps, every=20)
    ub, vb, om = rc.surface(0.5, np.random.default_rng(2024).normal(0.0, 0.01, rc.npts))
    cd = []
    with cuibm_amd.Solver(case("rotatingCylinderRe100_luo"), "-amgPrecision", prec, diagnostics=False) as s:
        s.initialise()
        s.set_body_velocity(ub, vb, om)
        for _ in range(steps // 50):
            s.step_time(50)
            cd.append(2.0 * s.info()["forceX"])
        return np.array(cd), s.info()["iter2_total"], s.info()["krylov_restarts"]


def test_fp32_vcycle_long_run_at_tight_tolerance_follows_fp64():
    steps = 650
    cd32, its32, restarts = _rotating_run("32", steps)
    cd64, its64, _ = _rotating_run("64", steps)
    print(f"{steps} steps at 1e-12: C_D {cd32[-1]:.12f} (FP32 V-cycle) {cd64[-1]:.12f} (FP64); largest |dC_D| "
          f"{np.max(np.abs(cd32 - cd64)):.2e}; Poisson iterations {its32:.0f} against {its64:.0f}, "
          f"{restarts:.0f} restarts", flush=True)
    assert np.max(np.abs(cd32 - cd64)) <= 1e-8
    assert abs(its32 - its64) <= 0.15 * its64
    assert restarts > 0
