"""Ensemble case generation and the ensemble entry point's argument handling, on a
machine without a GPU. The generated grids are checked against the library's own case
front end (the reference's grid rule), not against the generator's arithmetic."""
import math
import os

import numpy as np
import pytest

import cuibm_amd
from cuibm_amd import ensemble as ens
from helpers import CASES


def test_drawn_edges_are_what_the_grid_rule_gives(tmp_path):
    paths, cases = ens.write_cases(str(tmp_path), n=64, seed=2024)
    edges = [c[0] for c in cases]
    assert min(edges) >= 95 and max(edges) <= 513  # drawn from 96..512, odd totals
    assert len(set(edges)) > 40
    for p, (edge, reynolds, r) in zip(paths, cases):
        s = cuibm_amd.Solver(p)  # parsing only: no device work before initialise()
        sc = s.case_scalars()
        assert int(sc["nx"]) == edge and int(sc["ny"]) == edge, p
        assert abs(sc["nu"] - 1.0 / reynolds) <= 1e-15 and 40.0 <= reynolds <= 200.0
        s.close()


def test_stretch_ratio_lies_inside_the_interval_of_its_cell_count():
    h = 0.025
    for m in (24, 57, 100, 232):
        r = ens.ratio_for(m, h)
        assert ens.cells_stretched(r, h) == m
        # a ratio 1% of the way to either neighbouring count keeps the count
        assert ens.cells_stretched(r * (1 + 1e-6), h) == m and ens.cells_stretched(r * (1 - 1e-6), h) == m
    # the uniform core of [-0.6, 0.6] at h = 0.025 is 47 cells by the rule's truncation
    assert ens.core_cells(0.025) == 47 and 1.2 / 0.025 < 48


def test_draw_is_seeded():
    assert ens.draw(8, seed=7) == ens.draw(8, seed=7)
    assert ens.draw(8, seed=7) != ens.draw(8, seed=8)


def test_ensemble_entry_point_rejects_bad_arguments():
    with pytest.raises(cuibm_amd.CuibmError) as e:
        cuibm_amd.ensemble_run([], steps=1, workers=1)
    assert e.value.code == -6  # CUIBM_ERR_ARGUMENT
    with pytest.raises(cuibm_amd.CuibmError) as e:
        cuibm_amd.ensemble_run([os.path.join(CASES, "cylinderRe40_small_luo")], steps=1, workers=0)
    assert e.value.code == -6
    with pytest.raises(cuibm_amd.CuibmError) as e:  # the case front end fails before any device work
        cuibm_amd.ensemble_run([os.path.join(CASES, "no_such_case")], steps=1, workers=1)
    assert e.value.code == -2 and "case 0" in str(e.value)
