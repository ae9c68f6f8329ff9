"""Plain Navier-Stokes step on the GPU against the oracle (lid-driven cavity).

Parity bar: per-step relative L2 of u and p within 1e-10 (solver tolerances are
tightened so that the BiCGSTAB stopping point does not dominate the difference).
"""
import numpy as np
import pytest
import pyoracle
from helpers import case, pair, rel_l2

pytestmark = pytest.mark.gpu

TIGHT = {"velocityTol": 1e-12, "poissonTol": 1e-12}


def test_matrices_match_oracle():
    g, o = pair("lidDrivenCavityRe100_32", TIGHT)
    for m in ("LHS1", "LHS2"):
        assert np.array_equal(g.get(m + "_rowptr"), o.get(m + "_rowptr"))
        assert np.array_equal(g.get(m + "_col"), o.get(m + "_col"))
        assert rel_l2(g.get(m + "_val"), o.get(m + "_val")) < 1e-15


def test_rhs1_pieces_match_oracle():
    g, o = pair("lidDrivenCavityRe100_32", TIGHT)
    rng = np.random.default_rng(7)
    u = rng.standard_normal(g.get("u").size)
    nold = rng.standard_normal(u.size)
    for s in (g, o):
        s.set("u", u)
        s.set("Nold", nold)
    g.phase("rhs1_only")
    for ph in ("generateN", "generateL", "generateBC1", "generateRHS"):
        o.phase(ph)
    for f in ("N", "L", "bc1", "rhs1"):
        assert rel_l2(g.get(f), o.get(f)) < 1e-14, f


@pytest.mark.parametrize("name", ["lidDrivenCavityRe100_32", "lidDrivenCavityRe100_32_jacobi"])
def test_steps_match_oracle(name):
    g, o = pair(name, TIGHT)
    for k in range(10):
        g.step_time()
        o.step()
        for f in ("u", "p"):
            e = rel_l2(g.get(f), o.get(f))
            assert e < 1e-10, (k, f, e)
    gi, oi = g.info(), o.info()
    assert gi["timeStep"] == oi["timeStep"] == 10


def test_cavity128_cfg1_1000_steps():
    """The 128x128 Re=100 cavity over its full 1000 steps (SA-AMG Poisson, rtol 1e-8 as
    configured), u and p compared every 100 steps."""
    g, o = pair("lidDrivenCavityRe100_128")
    for k in range(10):
        g.step_time(100)
        o.step(100)
        errs = {f: rel_l2(g.get(f), o.get(f)) for f in ("u", "p")}
        print(f"step {100 * (k + 1)}: u {errs['u']:.2e} p {errs['p']:.2e}", flush=True)
        for f, e in errs.items():
            assert e < 1e-10, (k, f, e)


@pytest.mark.parametrize("name", ["lidDrivenCavityRe100_32", "cylinderRe40_small_luo"])
def test_poisson_stencil_equals_csr_product(name):
    """the Poisson solve's stencil operator reproduces the assembled LHS2 bit for bit"""
    import cuibm_amd

    g, _ = pair(name, TIGHT)
    rng = np.random.default_rng(11)
    p = rng.standard_normal(g.get("p").size)
    g.set("p", p)
    g.phase("poisson_operator")
    y_stencil = g.get("rhs2")
    import scipy.sparse as sp

    rp, ci, v = g.get("LHS2_rowptr"), g.get("LHS2_col"), g.get("LHS2_val")
    A = cuibm_amd.Csr(sp.csr_matrix((v, ci, rp), shape=(p.size, p.size)))
    y_csr = A.spmv(p)
    assert np.array_equal(y_stencil, y_csr)


@pytest.mark.parametrize("name", ["lidDrivenCavityRe100_32", "cylinderRe40_small_luo"])
def test_velocity_stencil_equals_csr_product(name):
    """the velocity solve's stencil operator reproduces the assembled LHS1 bit for bit"""
    import cuibm_amd
    import scipy.sparse as sp

    g, _ = pair(name, TIGHT)
    rng = np.random.default_rng(12)
    u = rng.standard_normal(g.get("uhat").size)
    g.set("uhat", u)
    g.phase("velocity_operator")
    y_stencil = g.get("rhs1")
    rp, ci, v = g.get("LHS1_rowptr"), g.get("LHS1_col"), g.get("LHS1_val")
    A = cuibm_amd.Csr(sp.csr_matrix((v, ci, rp), shape=(u.size, u.size)))
    assert np.array_equal(y_stencil, A.spmv(u))


@pytest.mark.parametrize("name", ["lidDrivenCavityRe100_32", "cylinderRe40_small_luo"])
def test_cfl_monitor_and_phase_timers(name):
    """The CFL monitor (NavierStokes/CFL.cu:10-35) keeps the largest dt (|u|/dx + |v|/dy)
    over the cells (last column and row excluded), with the first cell attaining it and
    the step; the phase timers carry the reference's event names."""
    g, o = pair(name, TIGHT)
    dt = float(pyoracle.read_case(case(name))["sim"]["dt"])
    dx, dy = o.get("dx"), o.get("dy")
    nx, ny = len(dx), len(dy)
    best, cell, step = 0.0, (0, 0), 0
    for k in range(4):
        g.step_time()
        u = g.get("u")
        U = u[: (nx - 1) * ny].reshape(ny, nx - 1)
        V = u[(nx - 1) * ny:].reshape(ny - 1, nx)
        c = dt * (np.abs(U[: ny - 1, : nx - 1]) / dx[None, : nx - 1] + np.abs(V[:, : nx - 1]) / dy[: ny - 1, None])
        m = c.max()
        if m > best:
            best = m
            J, I = np.unravel_index(np.argmax(c), c.shape)
            cell, step = (int(I), int(J)), k
        gi = g.info()
        assert abs(gi["cfl_max"] - best) <= 1e-14 * best, (k, gi["cfl_max"], best)
        assert (gi["cfl_I"], gi["cfl_J"], gi["cfl_ts"]) == (cell[0], cell[1], step), k
    t = g.timers()
    for e in ("initialise", "Intermediate Velocity Setup", "Intermediate Velocity Solve", "Poisson Setup",
              "Poisson Solve", "Velocity Projection", "CFL"):
        assert e in t and t[e] >= 0, e


def test_csr_operator_form_matches_stencil_form():
    """-operators csr (stored CSR operators for both solves and every AMG level) against the
    default stencil forms on the Luo cylinder: the operators are the same numbers, only the
    V-cycle's finest-level arithmetic order differs, so u and C_D agree to rounding"""
    import cuibm_amd
    from helpers import case

    forms = []
    for extra in ((), ("-operators", "csr")):
        g = cuibm_amd.Solver(case("cylinderRe40_small_luo"), "-poissonTol", "1e-12", "-velocityTol", "1e-12", *extra,
                             diagnostics=False)
        g.initialise()
        g.step_time(5)
        forms.append((g.get("u"), 2.0 * g.info()["forceX"]))
        g.close()
    (u0, cd0), (u1, cd1) = forms
    print("u", np.linalg.norm(u1 - u0) / np.linalg.norm(u0), "C_D", cd0, cd1, flush=True)
    assert np.linalg.norm(u1 - u0) <= 1e-10 * np.linalg.norm(u0)
    assert abs(cd1 - cd0) <= 1e-9
