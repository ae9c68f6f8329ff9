"""Lockstep ensembles on the GPU (cuibm_ensemble_lockstep): every phase of the step is one
launch over all cases, each case's BiCGSTAB keeping its own scalars and stopping test.

The only difference from running each case alone is the grouping of the dot-product and
force partial sums (per-case blocks of the batch instead of a fixed grid of 2048 blocks).
At solver tolerance 1e-8 that can move a solve's stopping point by an iteration, so drag
and lift are compared within 1e-7 of the cases run alone with the same AMG precision over
50 steps of the impulsive start, for the FP32 V-cycle and for the FP64 hierarchy; with
solves to 1e-12 the FP64 batch's first case agrees with the oracle within 1e-8 after 20
steps. Graph replays of the solves' iteration blocks equal eager launches bit for bit. Cases
that the batch does not cover are refused with an error.
"""
import numpy as np
import pytest

import cuibm_amd
import pyoracle as po
from cuibm_amd import ensemble as ens
from helpers import case

pytestmark = pytest.mark.gpu


def _alone(paths, prec, steps):
    out = []
    for p in paths:
        with cuibm_amd.Solver(p, "-amgPrecision", prec, diagnostics=False) as s:
            s.initialise()
            s.step_time(steps)
            i = s.info()
            out.append((2 * i["forceX"], 2 * i["forceY"]))
    return np.array(out)


@pytest.mark.parametrize("prec", ["32", "64"])
def test_lockstep_equals_cases_run_alone(tmp_path, prec):
    steps = 50
    paths, cases = ens.write_cases(str(tmp_path), n=7, seed=5, steps=steps, edges=(96, 260))
    print("edges", [c[0] for c in cases], flush=True)
    r = cuibm_amd.ensemble_run(paths, "-amgPrecision", prec, steps=steps, workers=4, lockstep=True)
    assert np.array_equal(r["cells"], [c[0] ** 2 for c in cases])
    ref = _alone(paths, prec, steps)
    d = np.max(np.abs(np.c_[r["cd"], r["cl"]] - ref), axis=1)
    print(f"precision {prec}: lockstep {r['steps_wall_s']:.2f} s; C_D {np.round(r['cd'], 6)}; "
          f"largest |d| per case {d}", flush=True)
    assert np.all(d <= 1e-7)


def test_lockstep_fp64_first_case_matches_oracle(tmp_path):
    steps = 20
    paths, cases = ens.write_cases(str(tmp_path), n=3, seed=9, steps=steps, edges=(96, 200))
    tight = ["-velocityTol", "1e-12", "-poissonTol", "1e-12"]
    r = cuibm_amd.ensemble_run(paths, "-amgPrecision", "64", *tight, steps=steps, workers=3, lockstep=True)
    o = po.Oracle(po.read_case(paths[0], {"velocityTol": 1e-12, "poissonTol": 1e-12}))
    o.step(steps)
    oi = o.info()
    print("case 0: lockstep C_D", r["cd"][0], "oracle", 2 * oi["forceX"], flush=True)
    assert abs(r["cd"][0] - 2 * oi["forceX"]) <= 1e-8 * max(1.0, abs(2 * oi["forceX"]))
    assert abs(r["cl"][0] - 2 * oi["forceY"]) <= 1e-8


def test_lockstep_graphs_equal_eager_launches(tmp_path):
    """the solves' eight-iteration blocks replayed as captured graphs give the results of
    the same kernels launched one by one, bit for bit (every step replays each graph several
    times, with the batch's own launches and read-backs between the replays)"""
    steps = 30
    paths, cases = ens.write_cases(str(tmp_path), n=5, seed=11, steps=steps, edges=(96, 220))
    r1 = cuibm_amd.ensemble_run(paths, "-amgPrecision", "32", steps=steps, workers=4, lockstep=True)
    r0 = cuibm_amd.ensemble_run(paths, "-amgPrecision", "32", "-lockstepGraphs", "0", steps=steps, workers=4,
                                lockstep=True)
    print(f"graphs {r1['steps_wall_s']:.3f} s, eager {r0['steps_wall_s']:.3f} s; C_D {np.round(r1['cd'], 6)}",
          flush=True)
    assert np.array_equal(r1["cd"], r0["cd"]) and np.array_equal(r1["cl"], r0["cl"])


def test_lockstep_refuses_cases_it_does_not_cover(tmp_path):
    paths, _ = ens.write_cases(str(tmp_path), n=2, seed=3, steps=5, edges=(96, 120))
    with pytest.raises(cuibm_amd.CuibmError) as e:
        cuibm_amd.ensemble_run([paths[0], case("cylinderRe40_small_fadlun")], steps=5, lockstep=True)
    assert e.value.code == -5  # CUIBM_ERR_UNSUPPORTED
