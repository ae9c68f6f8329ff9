"""Mixed-precision Poisson solve (-amgPrecision 32: the AMG V-cycle in FP32, its matrix
values, level vectors and smoother arithmetic, with the operator's pinned constant mode
removed and restored in FP64 around it; the Krylov vectors, dots and the projection in
FP64) against the FP64 oracle on the 336x336 Luo cylinder with the stencil finest level. Bar: |dC_D| <= 1e-6 per step (the precision
form's accuracy target); u within 1e-8 relative L2, since both forms solve to the same
residual tolerance and only the preconditioner differs."""
import numpy as np
import pytest

from helpers import pair, rel_l2

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
