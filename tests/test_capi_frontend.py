"""The C-ABI library on a machine without a GPU: it loads, exports every entry point
that include/cuibm_amd.h declares, and its case front end (YAML reader, stretched
grid generator, body points) agrees with the oracle's. No compute call is made.
"""
import json
import os
import re

import numpy as np
import pytest

import cuibm_amd
import pyoracle as po
from helpers import CASES, ROOT

HEADER = os.path.join(ROOT, "include", "cuibm_amd.h")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    with open(HEADER) as f:
        text = re.sub(r"/\*.*?\*/", "", f.read(), flags=re.S)
    return sorted(set(re.findall(r"\b(cuibm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_solver_lifecycle():
    syms = declared_symbols()
    for s in ("cuibm_create", "cuibm_initialise", "cuibm_step_time", "cuibm_write_data", "cuibm_finished",
              "cuibm_shut_down", "cuibm_destroy", "cuibm_bicgstab", "cuibm_amg_vcycle"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = cuibm_amd.lib()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_version_and_error_string():
    lib = cuibm_amd.lib()
    assert lib.cuibm_version() > 0
    with pytest.raises(cuibm_amd.CuibmError) as e:
        cuibm_amd.Solver(os.path.join(CASES, "no_such_case"))
    assert e.value.code == -2  # CUIBM_ERR_INPUT


LOCAL = sorted(d for d in os.listdir(CASES) if os.path.isdir(os.path.join(CASES, d)))


@pytest.mark.parametrize("name", LOCAL)
def test_front_end_matches_oracle(name):
    folder = os.path.join(CASES, name)
    s = cuibm_amd.Solver(folder)
    c = po.read_case(folder)
    sc = s.case_scalars()
    for d, (x, dx) in ((k, c["grid"][k]) for k in ("x", "y")):
        assert np.array_equal(s.case_array(d), x), d
        assert np.array_equal(s.case_array("d" + d), dx), d
    assert sc["nx"] == len(c["grid"]["x"][0]) and sc["ny"] == len(c["grid"]["y"][0])
    assert sc["dt"] == c["sim"]["dt"] and sc["nu"] == c["flow"]["nu"]
    assert sc["tol2"] == c["sim"]["Poisson"]["tolerance"]
    if c["bodies"]:
        X = np.concatenate([b[0] for b in c["bodies"]])
        Y = np.concatenate([b[1] for b in c["bodies"]])
        assert np.array_equal(s.case_array("body_X"), X)
        assert np.array_equal(s.case_array("body_Y"), Y)
    s.close()


def _write_case(tmp, dom, bodies):
    lines = []
    for d in ("x", "y"):
        p = dom[d]
        lines += [f"- direction: {d}", f"  start: {p['start']!r}", f"  mid_h: {p['mid_h']!r}", "  subDomains:"]
        for e, r in zip(p["ends"], p["ratios"]):
            lines += [f"    - end: {e!r}", f"      stretchRatio: {r!r}"]
    (tmp / "domain.yaml").write_text("\n".join(lines) + "\n")
    (tmp / "flow.yaml").write_text("- type: flow\n  nu: 0.025\n  initialVelocity: [1.0, 0.0]\n")
    (tmp / "simParams.yaml").write_text("- type: simulation\n  dt: 0.01\n  nt: 1\n  nsave: 1\n  SolverType: LUO\n")
    if bodies:
        b = bodies[0]
        opts = ", ".join(["true" if v is True else "false" if v is False else repr(v) for v in b])
        (tmp / "bodies.yaml").write_text(f"- type: circle\n  circleOptions: [{opts}]\n")


def _grid_cases():
    with open(os.path.join(GOLDEN, "reference_grids.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("c", _grid_cases(), ids=lambda c: c["case"].split("/", 2)[-1])
def test_front_end_on_reference_domains(c, tmp_path):
    """the product's grid generator reproduces the reference's recorded sizes and the
    oracle's coordinates on the reference's own validation domains"""
    _write_case(tmp_path, c["domain"], c["bodies"])
    s = cuibm_amd.Solver(str(tmp_path))
    sc = s.case_scalars()
    assert sc["nx"] == c["expected_n"] and sc["ny"] == c["expected_n"]
    p = c["domain"]["x"]
    x, dx = po.axis(p["start"], p["mid_h"], p["ends"], p["ratios"])
    assert np.array_equal(s.case_array("x"), x) and np.array_equal(s.case_array("dx"), dx)
    if c["bodies"]:
        b = c["bodies"][0]
        X, Y = po.circle(b[0], b[1], b[2], int(b[3]), bool(b[4]), b[5] if len(b) > 5 else 0.0)
        assert np.array_equal(s.case_array("body_X"), X)
        assert np.array_equal(s.case_array("body_Y"), Y)
    s.close()


def test_operators_switch_is_parsed():
    s = cuibm_amd.Solver(os.path.join(CASES, "cylinderRe40_small_luo"), "-operators", "csr")
    s.close()
    with pytest.raises(cuibm_amd.CuibmError) as e:
        cuibm_amd.Solver(os.path.join(CASES, "cylinderRe40_small_luo"), "-operators", "dense")
    assert e.value.code == -2 and "csr or stencil" in str(e.value)


def test_amg_coarse_max_is_parsed_and_bounded():
    s = cuibm_amd.Solver(os.path.join(CASES, "cylinderRe40_small_luo"), "-amgCoarseMax", "400")
    s.close()
    for bad in ("8", "10000"):
        with pytest.raises(cuibm_amd.CuibmError) as e:
            cuibm_amd.Solver(os.path.join(CASES, "cylinderRe40_small_luo"), "-amgCoarseMax", bad)
        assert e.value.code == -2 and "16..6144" in str(e.value)


def test_poisson_restart_is_parsed_and_bounded():
    for ok in ("0", "1e-5", "0.5"):
        cuibm_amd.Solver(os.path.join(CASES, "cylinderRe40_small_luo"), "-poissonRestart", ok).close()
    for bad in ("-1", "1", "x"):
        with pytest.raises(cuibm_amd.CuibmError) as e:
            cuibm_amd.Solver(os.path.join(CASES, "cylinderRe40_small_luo"), "-poissonRestart", bad)
        assert e.value.code == -2 and "[0, 1)" in str(e.value)


def test_amg_refresh_switch_is_parsed():
    for ok in ("0", "1"):
        cuibm_amd.Solver(os.path.join(CASES, "cylinderRe40_small_luo"), "-amgRefresh", ok).close()
    with pytest.raises(cuibm_amd.CuibmError) as e:
        cuibm_amd.Solver(os.path.join(CASES, "cylinderRe40_small_luo"), "-amgRefresh", "yes")
    assert e.value.code == -2 and "0 or 1" in str(e.value)


def test_csr_create_rejects_a_malformed_row_pointer():
    """row pointers that do not start at 0 or that decrease are refused before any upload"""
    import ctypes as C

    lib = cuibm_amd.lib()
    IP = C.POINTER(C.c_int)
    ci = np.zeros(4, dtype=np.int32)
    v = np.ones(4)
    h = C.c_void_p()
    for rp in ([1, 2, 3], [0, 2, 1], [0, 3, 2]):
        rp = np.array(rp, dtype=np.int32)
        rc = lib.cuibm_csr_create(2, rp.ctypes.data_as(IP), ci.ctypes.data_as(IP),
                                  v.ctypes.data_as(C.POINTER(C.c_double)), C.byref(h))
        assert rc == -6, rp
        assert b"row pointer" in lib.cuibm_last_error()
