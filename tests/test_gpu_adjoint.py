"""The discrete adjoint of the mean drag (cuibm_adjoint_run) on the GPU.

Case: a D = 1 cylinder of 256 points at Re = 100 on a 114 x 114 grid (h = 0.04 near the
body), impulsively started, 40 steps of dt = 0.01, objective the mean C_D over steps
21..40; surface velocity alpha (D / 2) omega_ref along the tangent (alpha = 0.5,
omega_ref = 2) plus normal blowing drawn from N(0, 0.01^2) (seed 7); solves to 1e-12.

- the objective of the forward pass equals the mean C_D of the ordinary time stepping
  with the same body velocities, and all-zero body velocities reproduce the plain run;
- dJ/dalpha and the directional derivative along a random unit vector over the 256
  blowing values against central differences with h = 1e-6, within 1e-5 relative;
- the same against central differences of the FP64 CPU oracle (fixture
  tests/golden/adjoint_small_fd.json, written by tests/golden/make_adjoint_fixtures.py):
  the objective within 1e-9 relative, dJ/dalpha and two directional derivatives over the
  blowing values within 1e-5 relative;
- the gradient does not depend on the checkpoint interval (1, 7 and 40 steps);
- the torch.autograd.Function returns the same gradient through J.backward();
- at the full size of the rotating-cylinder workload (498 x 498, 800 steps of dt = 0.005,
  J over steps 400..800, checkpoints every 20 steps): dJ/dalpha and the directional
  derivative over the blowing values against central differences (h = 1e-6) within 1e-5.
"""
import json
import os

import numpy as np
import pytest
import torch

import cuibm_amd
from cuibm_amd.adjoint import RotatingCylinder, mean_drag
from helpers import case

pytestmark = pytest.mark.gpu

SMALL = "rotatingCylinderRe100_small_luo"
ALPHA = 0.5


def _problem(every=7):
    return RotatingCylinder(case(SMALL), nsteps=40, n0=21, n1=40, every=every)


def _blow(n):
    return np.random.default_rng(7).normal(0.0, 0.01, n)


def test_forward_objective_matches_time_stepping():
    p = _problem()
    blow = _blow(p.npts)
    ub, vb, om = p.surface(ALPHA, blow)
    J, _, _ = p.evaluate(ALPHA, blow, backward=False)
    cd = []
    with cuibm_amd.Solver(case(SMALL), diagnostics=False) as s:
        s.initialise()
        s.set_body_velocity(ub, vb, om)
        for k in range(1, 41):
            s.step_time()
            if k >= 21:
                cd.append(2 * s.info()["forceX"])
    print(f"J {J:.15f} stepping {np.mean(cd):.15f}", flush=True)
    assert abs(J - np.mean(cd)) <= 1e-13 * abs(J)
    # all-zero body velocities: the per-point path against the plain static body
    z = np.zeros(p.npts)
    with cuibm_amd.Solver(case(SMALL), diagnostics=False) as a, cuibm_amd.Solver(case(SMALL), diagnostics=False) as b:
        a.initialise()
        b.initialise()
        a.set_body_velocity(z, z, 0.0)
        a.step_time(10)
        b.step_time(10)
        assert np.array_equal(a.get("u"), b.get("u"))
        assert a.info()["forceX"] == b.info()["forceX"]


def test_adjoint_gradient_matches_central_differences():
    p = _problem()
    blow = _blow(p.npts)
    J, ga, gb = p.evaluate(ALPHA, blow)
    r = p.last
    print(f"J {J:.12f} dJ/dalpha {ga:.10e} |dJ/db| {np.linalg.norm(gb):.6e} forward {r['forward_s']:.2f} s "
          f"backward {r['backward_s']:.2f} s adjoint its {r['adjoint_poisson_its']:.0f} "
          f"{r['adjoint_velocity_its']:.0f}", flush=True)
    h = 1e-6
    fa = (p.evaluate(ALPHA + h, blow, False)[0] - p.evaluate(ALPHA - h, blow, False)[0]) / (2 * h)
    print(f"dJ/dalpha adjoint {ga:.10e} central {fa:.10e} rel {abs(ga - fa) / abs(fa):.2e}", flush=True)
    assert abs(ga - fa) <= 1e-5 * abs(fa)
    d = np.random.default_rng(11).normal(size=p.npts)
    d /= np.linalg.norm(d)
    fd = (p.evaluate(ALPHA, blow + h * d, False)[0] - p.evaluate(ALPHA, blow - h * d, False)[0]) / (2 * h)
    print(f"dJ/db . d adjoint {gb @ d:.10e} central {fd:.10e} rel {abs(gb @ d - fd) / np.linalg.norm(gb):.2e}",
          flush=True)
    assert abs(gb @ d - fd) <= 1e-5 * np.linalg.norm(gb)
    # the single largest blowing component
    k = int(np.argmax(np.abs(gb)))
    e = np.zeros(p.npts)
    e[k] = h
    fk = (p.evaluate(ALPHA, blow + e, False)[0] - p.evaluate(ALPHA, blow - e, False)[0]) / (2 * h)
    print(f"dJ/db[{k}] adjoint {gb[k]:.10e} central {fk:.10e}", flush=True)
    assert abs(gb[k] - fk) <= 1e-5 * abs(fk)


def test_adjoint_gradient_matches_oracle_differences():
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "adjoint_small_fd.json")) as f:
        ref = json.load(f)
    p = _problem()
    blow = _blow(p.npts)
    J, ga, gb = p.evaluate(ALPHA, blow)
    print(f"J {J:.13f} oracle {ref['J']:.13f}; dJ/dalpha {ga:.10e} oracle differences {ref['dJ_dalpha_h1e-05']:.10e} "
          f"(h 1e-5) {ref['dJ_dalpha_h1e-06']:.10e} (h 1e-6)", flush=True)
    assert abs(J - ref["J"]) <= 1e-9 * abs(ref["J"])
    assert abs(ga - ref["dJ_dalpha_h1e-05"]) <= 1e-5 * abs(ref["dJ_dalpha_h1e-05"])
    for seed, key in ((11, "dJ_dir"), (12, "dJ_dir2")):
        d = np.random.default_rng(seed).normal(size=p.npts)
        d /= np.linalg.norm(d)
        print(f"direction {seed}: adjoint {gb @ d:.10e} oracle differences {ref[key]:.10e}", flush=True)
        assert abs(gb @ d - ref[key]) <= 1e-5 * np.linalg.norm(gb)


def test_gradient_independent_of_checkpoint_interval():
    blow = _blow(256)
    out = [_problem(every).evaluate(ALPHA, blow) for every in (1, 7, 40)]
    for J, ga, gb in out[1:]:
        assert abs(J - out[0][0]) <= 1e-14 * abs(J)
        assert abs(ga - out[0][1]) <= 1e-11 * abs(ga)
        assert np.max(np.abs(gb - out[0][2])) <= 1e-11 * np.max(np.abs(gb))


def test_autograd_function():
    p = _problem()
    blow = _blow(p.npts)
    a = torch.tensor(ALPHA, dtype=torch.float64, requires_grad=True)
    b = torch.tensor(blow, dtype=torch.float64, requires_grad=True)
    J = mean_drag(a, b, p)
    (3.0 * J).backward()
    J0, ga, gb = p.evaluate(ALPHA, blow)
    assert float(J) == J0
    assert abs(float(a.grad) - 3 * ga) <= 1e-12 * abs(3 * ga)
    assert np.allclose(b.grad.numpy(), 3 * gb, rtol=1e-12, atol=0)


def test_full_size_rotating_cylinder_against_central_differences():
    p = RotatingCylinder(case("rotatingCylinderRe100_luo"), nsteps=800, n0=400, n1=800, every=20)
    blow = np.random.default_rng(2024).normal(0.0, 0.01, p.npts)
    J, ga, gb = p.evaluate(ALPHA, blow)
    r = p.last
    print(f"J {J:.12f}; forward {r['forward_s']:.1f} s, backward {r['backward_s']:.1f} s "
          f"(recompute {r['backward_recompute_s']:.1f} s); Poisson its forward {r['forward_poisson_its']:.0f} "
          f"adjoint {r['adjoint_poisson_its']:.0f}", flush=True)
    h = 1e-6
    fa = (p.evaluate(ALPHA + h, blow, False)[0] - p.evaluate(ALPHA - h, blow, False)[0]) / (2 * h)
    print(f"dJ/dalpha adjoint {ga:.10e} central {fa:.10e} rel {abs(ga - fa) / abs(fa):.2e}", flush=True)
    assert abs(ga - fa) <= 1e-5 * abs(fa)
    d = np.random.default_rng(11).normal(size=p.npts)
    d /= np.linalg.norm(d)
    fd = (p.evaluate(ALPHA, blow + h * d, False)[0] - p.evaluate(ALPHA, blow - h * d, False)[0]) / (2 * h)
    print(f"dJ/db . d adjoint {gb @ d:.10e} central {fd:.10e} rel {abs(gb @ d - fd) / np.linalg.norm(gb):.2e}",
          flush=True)
    assert abs(gb @ d - fd) <= 1e-5 * np.linalg.norm(gb)

