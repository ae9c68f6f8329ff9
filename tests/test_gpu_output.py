"""Output files of a run with output on, in the reference's layouts (src/io/io.cu:406-488,
src/bodies.cu:459-471, luo_base.cu:119-206, NavierStokesSolver.cu:339-363, logger.h:216-222):
binary grid (int32 n, n+1 doubles per axis), binary q / lambda per saved step (int32 N, N
doubles), text bodies, forces, midPosition, iterations, run.info and time; plus forces_full,
the forces at 17 significant digits."""
import os

import numpy as np
import pytest

import cuibm_amd
from helpers import case

pytestmark = pytest.mark.gpu


def _read_block(f, dtype, n):
    a = np.fromfile(f, dtype=dtype, count=n)
    assert a.size == n
    return a


def test_output_files_follow_the_reference_layout(tmp_path):
    out = tmp_path / "run"
    g = cuibm_amd.Solver(case("cylinderRe40_small_luo"), "-nsave", "2", output=True, output_folder=out)
    g.initialise()
    for _ in range(2):
        g.step_time()
        g.write_data()
    info = g.info()
    uhat, p, bx, by = g.get("uhat"), g.get("p"), g.get("body_x"), g.get("body_y")
    g.shut_down()
    nx, ny = int(info["nx"]), int(info["ny"])

    with open(out / "grid", "rb") as f:
        assert _read_block(f, np.int32, 1)[0] == nx
        x = _read_block(f, np.float64, nx + 1)
        assert _read_block(f, np.int32, 1)[0] == ny
        y = _read_block(f, np.float64, ny + 1)
        assert f.read() == b""
    assert np.all(np.diff(x) > 0) and np.all(np.diff(y) > 0)

    step = out / "0000002"
    for name, ref in (("q", uhat), ("lambda", p)):
        with open(step / name, "rb") as f:
            n = _read_block(f, np.int32, 1)[0]
            assert n == ref.size
            assert np.array_equal(_read_block(f, np.float64, n), ref)
            assert f.read() == b""

    lines = (step / "bodies").read_text().splitlines()
    assert lines[0] == "#" + "x-coordinate".rjust(19) + "y-coordinate".rjust(20)
    pts = np.array([[float(v) for v in l.split("\t")] for l in lines[1:]])
    assert pts.shape == (bx.size, 2)
    np.testing.assert_allclose(pts[:, 0], bx, rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(pts[:, 1], by, rtol=1e-5, atol=1e-12)

    forces = (out / "forces").read_text().splitlines()
    assert forces[0] == "timestep\tFx\tFy" and len(forces) == 3
    full = [[float(v) for v in l.split("\t")] for l in (out / "forces_full").read_text().splitlines()]
    assert len(full) == 2
    assert full[-1][1] == info["forceX"] and full[-1][2] == info["forceY"]
    last = [float(v) for v in forces[-1].split("\t")]
    np.testing.assert_allclose(last[1:], full[-1][1:], rtol=1e-5)

    mid = (out / "midPosition").read_text().splitlines()
    assert mid[0] == "ts\tX\tY\tU\tV" and len(mid) == 3
    its = [l.split("\t") for l in (out / "iterations").read_text().splitlines()]
    assert [int(r[0]) for r in its] == [1, 2]
    assert int(its[-1][1]) == int(info["iter1"]) and int(its[-1][2]) == int(info["iter2"])

    runinfo = (out / "run.info").read_text().splitlines()
    assert runinfo[0].split("\t") == ["--nx".rjust(20), str(nx)]
    timers = dict(l.rsplit(" ", 1) for l in (out / "time").read_text().splitlines())
    assert "Poisson Solve" in timers and float(timers["Poisson Solve"]) >= 0
    assert not os.path.exists(out / "0000001")  # nsave = 2
