"""Ensembles on the GPU: cases run concurrently from several host threads, each on its
own stream, give exactly the drag and lift of the same cases run one at a time, and one
case agrees with the oracle (FP64 hierarchy) to the mixed-precision bar |dC_D| <= 1e-6."""
import numpy as np
import pytest

import cuibm_amd
import pyoracle as po
from cuibm_amd import ensemble as ens

pytestmark = pytest.mark.gpu

STEPS = 20


def test_concurrent_cases_equal_cases_run_alone(tmp_path):
    paths, cases = ens.write_cases(str(tmp_path), n=6, seed=11, steps=STEPS, edges=(96, 160))
    print("edges", [c[0] for c in cases], flush=True)
    r = cuibm_amd.ensemble_run(paths, "-amgPrecision", "32", steps=STEPS, workers=3)
    assert np.array_equal(r["cells"], [c[0] ** 2 for c in cases])
    for i, p in enumerate(paths):
        s = cuibm_amd.Solver(p, "-amgPrecision", "32", diagnostics=False)
        s.initialise()
        s.step_time(STEPS)
        info = s.info()
        s.close()
        assert 2.0 * info["forceX"] == r["cd"][i] and 2.0 * info["forceY"] == r["cl"][i], i
    o = po.Oracle(po.read_case(paths[0]))
    o.step(STEPS)
    cd_o = 2.0 * o.info()["forceX"]
    print("case 0 C_D: ensemble (FP32 hierarchy)", r["cd"][0], "oracle (FP64)", cd_o, flush=True)
    assert abs(r["cd"][0] - cd_o) <= 1e-6
