"""Pins of the oracle (the FP64 CPU restatement) against what does not depend on it.

Each test checks the oracle against a reference-held number, a closed form, an
invariant of the discretisation or an independent textbook computation (numpy /
scipy), chosen so that a dropped term, a wrong sign or index, or a transposed
operand fails. None of them re-types the oracle's own formula.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import pyoracle as po
from helpers import case

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ----------------------------------------------------------------------------- grid

def _grid_cases():
    with open(os.path.join(GOLDEN, "reference_grids.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("c", _grid_cases(), ids=lambda c: c["case"].split("/", 2)[-1])
def test_grid_sizes_match_reference_records(c):
    """cells per side recorded by the reference (performance.py:22, performance2.py:84-86)"""
    for d in ("x", "y"):
        p = c["domain"][d]
        x, dx = po.axis(p["start"], p["mid_h"], p["ends"], p["ratios"])
        assert len(dx) == c["expected_n"], (c["case"], d)


@pytest.mark.parametrize("c", _grid_cases()[:12], ids=lambda c: c["case"].split("/", 2)[-1])
def test_grid_closed_forms(c):
    p = c["domain"]["x"]
    x, dx = po.axis(p["start"], p["mid_h"], p["ends"], p["ratios"])
    # cell centres are the running sum of the widths
    edges = p["start"] + np.concatenate([[0.0], np.cumsum(dx)])
    assert np.allclose(x, 0.5 * (edges[:-1] + edges[1:]), rtol=0, atol=1e-11)
    # every subdomain is tiled exactly: the edges hit each subdomain end
    for e in p["ends"]:
        assert np.min(np.abs(edges - e)) < 1e-10 * max(1.0, abs(e))
    # uniform core: constant width (end-start)/n with n = int((end-start)/mid_h), never below mid_h
    starts = [p["start"]] + p["ends"][:-1]
    for s, e, r in zip(starts, p["ends"], p["ratios"]):
        m = (edges[:-1] >= s - 1e-12) & (edges[1:] <= e + 1e-12)
        w = dx[m]
        if r == 1.0:
            n = int((e - s) / p["mid_h"])
            assert w.size == n and np.allclose(w, (e - s) / n, rtol=1e-13)
        else:
            # geometric: ratio r between neighbours, smallest cell at the core and >= mid_h
            q = w[1:] / w[:-1]
            assert np.allclose(q, r, rtol=1e-12) or np.allclose(q, 1 / r, rtol=1e-12)
            assert min(w) >= p["mid_h"] * (1 - 1e-12)
            assert np.isclose(w.sum(), e - s, rtol=1e-12)


def test_circle_points_closed_form():
    for cx, cy, R, n, self_size, h in [(0, 0, 0.5, 314, False, 0), (0.25, -0.5, 0.5, 126, True, 0.03),
                                       (0, 0, 0.5, 126, True, 0.005)]:
        X, Y = po.circle(cx, cy, R, n, self_size, h)
        m = int(2 * R * np.pi / h) if self_size else n
        m += m % 2 if self_size else 0
        assert len(X) == m
        t = 2 * np.pi * np.arange(m) / m
        assert np.allclose(X, cx + R * np.cos(t), rtol=0, atol=2e-16)
        assert np.allclose(Y, cy + R * np.sin(t), rtol=0, atol=2e-16)


# ----------------------------------------------------------------------------- operators

def _cavity(n=32, **ov):
    return po.Oracle(po.read_case(case(f"lidDrivenCavityRe100_{n}"), ov or None))


def _uv_coords(o):
    x, y, dx, dy = (o.get(k) for k in ("x", "y", "dx", "dy"))
    nx, ny = len(dx), len(dy)
    xu = np.cumsum(dx)[:-1] + (x[0] - 0.5 * dx[0])  # u faces between cells
    yv = np.cumsum(dy)[:-1] + (y[0] - 0.5 * dy[0])
    XU, YU = np.meshgrid(xu, y)          # u at (xu[i], y[j]), row J = j
    XV, YV = np.meshgrid(x, yv)          # v at (x[i], yv[j])
    return nx, ny, XU, YU, XV, YV


def test_L_exact_for_quadratics_on_uniform_grid():
    o = _cavity()
    nu = o.case["flow"]["nu"]
    nx, ny, XU, YU, XV, YV = _uv_coords(o)
    a, b, c, d = 0.7, -1.3, 0.4, 0.2
    fu = a * XU**2 + b * YU**2 + c * XU * YU + d * XU
    fv = -b * XV**2 + a * YV**2 - c * XV * YV
    o.set("u", np.concatenate([fu.ravel(), fv.ravel()]))
    o.phase("generateL")
    L = o.get("L")
    Lu = L[: (nx - 1) * ny].reshape(ny, nx - 1)
    Lv = L[(nx - 1) * ny:].reshape(ny - 1, nx)
    # interior nodes (no wall ghost in the stencil)
    assert np.allclose(Lu[1:-1, 1:-1], nu * 2 * (a + b), rtol=1e-9, atol=1e-9)
    assert np.allclose(Lv[1:-1, 1:-1], nu * 2 * (a - b), rtol=1e-9, atol=1e-9)


def test_N_of_linear_solenoidal_field():
    """u = x, v = -y: (u.grad)u = (x, y) exactly for the central discretisation"""
    o = _cavity()
    nx, ny, XU, YU, XV, YV = _uv_coords(o)
    o.set("u", np.concatenate([XU.ravel(), (-YV).ravel()]))
    o.phase("generateN")
    N = o.get("N")
    Nu = N[: (nx - 1) * ny].reshape(ny, nx - 1)[1:-1, 1:-1]
    Nv = N[(nx - 1) * ny:].reshape(ny - 1, nx)[1:-1, 1:-1]
    xu, yv = XU[1:-1, 1:-1], YV[1:-1, 1:-1]
    s = np.sign(np.sum(Nu * xu))  # one sign convention for both components
    assert s != 0
    assert np.allclose(Nu, s * xu, rtol=0, atol=1e-12)
    assert np.allclose(Nv, s * yv, rtol=0, atol=1e-12)
    # a constant field convects to exactly zero away from the walls
    o.set("u", np.concatenate([np.full(XU.size, 0.3), np.full(XV.size, -1.1)]))
    o.phase("generateN")
    N = o.get("N")
    assert np.all(N[: (nx - 1) * ny].reshape(ny, nx - 1)[1:-1, 1:-1] == 0)
    assert np.all(N[(nx - 1) * ny:].reshape(ny - 1, nx)[1:-1, 1:-1] == 0)


def _csr(o, m):
    rp, ci, v = o.get(m + "_rowptr"), o.get(m + "_col"), o.get(m + "_val")
    n = len(rp) - 1
    return sp.csr_matrix((v, ci, rp), shape=(n, n))


@pytest.mark.parametrize("name", ["lidDrivenCavityRe100_32", "cylinderRe40_small_luo"])
def test_lhs_nnz_and_row_sums(name):
    o = po.Oracle(po.read_case(case(name)))
    nx, ny = len(o.get("dx")), len(o.get("dy"))
    numUV = (nx - 1) * ny + nx * (ny - 1)
    A1, A2 = _csr(o, "LHS1"), _csr(o, "LHS2")
    # nnz closed forms (5-point rows minus the missing wall neighbours)
    assert A1.nnz == 5 * numUV - 4 * nx - 4 * ny + 4
    assert A2.nnz == 5 * nx * ny - 2 * nx - 2 * ny
    # LHS1 = I - 0.5 dt nu Lap: rows away from the walls sum to 1
    s1 = np.asarray(A1.sum(axis=1)).ravel()
    Iu = np.arange((nx - 1) * ny)
    interior_u = Iu[(Iu % (nx - 1) > 0) & (Iu % (nx - 1) < nx - 2) & (Iu // (nx - 1) > 0) & (Iu // (nx - 1) < ny - 1)]
    assert np.allclose(s1[interior_u], 1.0, rtol=0, atol=1e-12)
    # LHS2 = -dt Lap with Neumann walls: zero row sums except the pinned rows,
    # whose sum equals their undoubled diagonal
    s2 = np.asarray(A2.sum(axis=1)).ravel()
    d2 = A2.diagonal()
    pinned = np.where(np.abs(s2) > 1e-12 * np.abs(d2))[0]
    assert 0 in pinned
    solver = o.case["sim"]["SolverType"]
    assert len(pinned) == (2 if solver == "NAVIER_STOKES_SOLVER" else 1)
    assert np.allclose(s2[pinned], d2[pinned] / 2, rtol=1e-13)


def test_lhs2_equals_kronecker_laplacian_on_uniform_grid():
    o = _cavity()
    nx, ny = len(o.get("dx")), len(o.get("dy"))
    h = o.get("dx")[0]
    dt = o.case["sim"]["dt"]

    def neumann(n):
        T = sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tolil()
        T[0, 0] = T[n - 1, n - 1] = 1
        return T.tocsr()

    K = (sp.kron(sp.identity(ny), neumann(nx)) + sp.kron(neumann(ny), sp.identity(nx))) * (dt / h**2)
    A2 = _csr(o, "LHS2").tolil()
    s2 = np.asarray(A2.sum(axis=1)).ravel()
    for k in np.where(np.abs(s2) > 1e-9)[0]:  # undo the pins
        A2[k, k] = A2[k, k] / 2
    assert abs(A2.tocsr() - K).max() < 1e-12 * abs(K).max()


def test_projection_makes_velocity_solenoidal():
    """after an (effectively exact) Poisson solve, div(u) vanishes except at pinned rows"""
    o = _cavity(poissonTol=1e-14)
    rng = np.random.default_rng(3)
    nuv = o.get("uhat").size
    o.set("uhat", rng.standard_normal(nuv))
    o.phase("rhs2")
    r0 = o.get("rhs2").copy()
    o.set("p", np.zeros(r0.size))
    o.phase("solve2")
    o.phase("project")
    o.set("uhat", o.get("u"))
    o.phase("rhs2")
    r1 = o.get("rhs2")
    A2 = _csr(o, "LHS2")
    s2 = np.asarray(A2.sum(axis=1)).ravel()
    free = np.abs(s2) < 1e-9
    assert np.max(np.abs(r1[free])) < 1e-9 * np.max(np.abs(r0))
    assert np.max(np.abs(r1[~free])) > 0


# ----------------------------------------------------------------------------- Krylov / AMG

def _random_nonsym(n, seed):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=0.03, random_state=seed, data_rvs=rng.standard_normal).tolil()
    for i in range(n):
        A[i, i] = 4.0 + abs(A[i, :].toarray()).sum()
    return A.tocsr()


@pytest.mark.parametrize("precond", ["NONE", "DIAGONAL", "SMOOTHED_AGGREGATION"])
def test_bicgstab_matches_dense_solve(precond):
    A = _random_nonsym(300, 11)
    b = np.random.default_rng(5).standard_normal(300)
    M = po.OrcCsr(A)
    x, it, res = M.bicgstab(b, np.zeros(300), precond=precond, rtol=1e-13, maxit=500)
    xe = np.linalg.solve(A.toarray(), b)
    assert np.linalg.norm(x - xe) <= 1e-11 * np.linalg.norm(xe)
    assert np.linalg.norm(b - A @ x) <= 1e-13 * np.linalg.norm(b) * 1.0001
    # a hierarchy of one level is an exact (dense) inverse: converged at the first half step
    assert it < 100 and (it > 0 or precond == "SMOOTHED_AGGREGATION")


def test_bicgstab_monitor_semantics():
    """cusp::monitor: the test ||r|| <= rtol ||b|| is made before the first iteration"""
    A = _random_nonsym(120, 2)
    b = np.random.default_rng(1).standard_normal(120)
    M = po.OrcCsr(A)
    _, it, _ = M.bicgstab(b, np.zeros(120), rtol=1.0)
    assert it == 0
    xe = np.linalg.solve(A.toarray(), b)
    _, it, _ = M.bicgstab(b, xe, rtol=1e-8)
    assert it == 0


def _poisson_lhs2(n=32):
    return _csr(_cavity(n), "LHS2")


@pytest.fixture
def small_coarse():
    """a coarsest level of <= 256 unknowns, so that a 32x32 problem has several levels"""
    po.set_amg_coarse_max(256)
    yield
    po.set_amg_coarse_max(0)


def test_amg_galerkin_identity(small_coarse):
    A = _poisson_lhs2(32)
    M = po.OrcCsr(A)
    sizes, nnz = M.amg_levels()
    assert sizes[0] == A.shape[0] and len(sizes) >= 2 and sizes[1] < sizes[0] / 4
    Ac = M.coarse_dense(1)
    P = M.prolongator_dense(sizes[1])
    G = P.T @ A.toarray() @ P
    assert np.abs(Ac - G).max() <= 1e-12 * np.abs(G).max()
    assert nnz[1] == np.count_nonzero(Ac)


def test_amg_vcycle_is_symmetric_and_contracts(small_coarse):
    """equal pre/post Jacobi sweeps and R = P^T make the V-cycle a symmetric operator;
    as a stationary iteration it reduces the error on the pinned Poisson problem"""
    A = _poisson_lhs2(32)
    M = po.OrcCsr(A)
    rng = np.random.default_rng(9)
    x, y = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    Mx, My = M.vcycle(x), M.vcycle(y)
    assert abs(x @ My - y @ Mx) <= 1e-12 * np.linalg.norm(x) * np.linalg.norm(My)
    e = rng.standard_normal(A.shape[0])
    rates = []
    for _ in range(8):
        en = e - M.vcycle(A @ e)
        rates.append(np.sqrt((en @ (A @ en)) / (e @ (A @ e))))
        e = en
    assert max(rates[3:]) < 0.9


# ----------------------------------------------------------------------------- IBM

def _point_in_polygon(px, py, X, Y):
    """even-odd rule against the closed polyline (X, Y), vectorised over points"""
    inside = np.zeros(px.shape, dtype=bool)
    n = len(X)
    for k in range(n):
        x1, y1, x2, y2 = X[k], Y[k], X[(k + 1) % n], Y[(k + 1) % n]
        cond = (y1 > py) != (y2 > py)
        with np.errstate(divide="ignore", invalid="ignore"):
            xc = x1 + (py - y1) * (x2 - x1) / (y2 - y1)
        inside ^= cond & (px < xc)
    return inside


@pytest.mark.parametrize("name", ["cylinderRe40_small_luo", "cylinderRe40_luo"])
def test_velocity_tags_against_point_in_polygon(name):
    o = po.Oracle(po.read_case(case(name)))
    X, Y = o.get("body_x"), o.get("body_y")
    x, y, xu, yv = o.get("x"), o.get("y"), o.get("xu"), o.get("yv")
    nx, ny = len(x), len(y)
    g, h = o.get("ghostTagsUV"), o.get("hybridTagsUV")
    XU, YU = np.meshgrid(xu, y)
    XV, YV = np.meshgrid(x, yv)
    px = np.concatenate([XU.ravel(), XV.ravel()])
    py = np.concatenate([YU.ravel(), YV.ravel()])
    inside = _point_in_polygon(px, py, X, Y)
    # nodes on the polyline itself are ambiguous for any rule
    near = np.zeros_like(inside)
    for k in range(len(X)):
        near |= np.hypot(px - X[k], py - Y[k]) < 1e-9
    ok = ~near
    assert np.array_equal(g[ok] != -1, inside[ok])
    numU = (nx - 1) * ny
    ghosts = np.where(g > 0)[0]
    assert np.array_equal(g[ghosts], ghosts)  # a ghost carries its own index
    assert len(ghosts) > 0 and np.all(inside[ghosts])
    # every ghost has a fluid 4-neighbour; every hybrid node is fluid and has a ghost 4-neighbour

    def nbrs(k):
        if k < numU:
            i, j, w, off, hgt = k % (nx - 1), k // (nx - 1), nx - 1, 0, ny
        else:
            i, j, w, off, hgt = (k - numU) % nx, (k - numU) // nx, nx, numU, ny - 1
        out = []
        for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
            if 0 <= i + di < w and 0 <= j + dj < hgt:
                out.append(off + (j + dj) * w + i + di)
        return out

    for k in ghosts:
        assert any(g[m] == -1 for m in nbrs(k)), k
    hyb = np.where(h >= 0)[0]
    assert len(hyb) > 0 and np.array_equal(h[hyb], hyb)
    for k in hyb:
        assert g[k] == -1 and not inside[k]
        assert any(g[m] > 0 for m in nbrs(k)), k


def test_body_intercepts_lie_on_the_body():
    o = po.Oracle(po.read_case(case("cylinderRe40_luo")))
    X, Y = o.get("body_x"), o.get("body_y")
    R = np.mean(np.hypot(X - X.mean(), Y - Y.mean()))
    g = o.get("ghostTagsUV")
    bx, by = o.get("body_intercept_x"), o.get("body_intercept_y")
    ix, iy = o.get("image_point_x"), o.get("image_point_y")
    x, y, xu, yv = o.get("x"), o.get("y"), o.get("xu"), o.get("yv")
    nx, ny = len(x), len(y)
    XU, YU = np.meshgrid(xu, y)
    XV, YV = np.meshgrid(x, yv)
    px = np.concatenate([XU.ravel(), XV.ravel()])
    py = np.concatenate([YU.ravel(), YV.ravel()])
    k = np.where(g > 0)[0]
    # the intercept is the foot of the perpendicular from the ghost node to the line
    # through one body segment: on that line, and node -> intercept normal to it
    n = len(X)
    ex, ey = np.roll(X, -1) - X, np.roll(Y, -1) - Y
    ln = np.hypot(ex, ey)
    for m in k:
        on_line = np.abs((bx[m] - X) * ey - (by[m] - Y) * ex) / ln
        s = np.argmin(on_line)
        assert on_line[s] < 1e-12
        dn = np.hypot(px[m] - bx[m], py[m] - by[m])
        assert abs((px[m] - bx[m]) * ex[s] + (py[m] - by[m]) * ey[s]) / ln[s] <= 1e-10 * max(dn, 1e-3)
        # and within one grid cell of the circle
        assert abs(np.hypot(bx[m] - X.mean(), by[m] - Y.mean()) - R) < 0.05
    # image point = mirror of the ghost node through its intercept, outside the body
    assert np.allclose(ix[k], 2 * bx[k] - px[k], rtol=0, atol=1e-13)
    assert np.allclose(iy[k], 2 * by[k] - py[k], rtol=0, atol=1e-13)
    assert not np.any(_point_in_polygon(ix[k], iy[k], X, Y))


def test_cv_force_vanishes_for_uniform_flow():
    o = po.Oracle(po.read_case(case("cylinderRe40_small_luo")))
    nuv, np_ = o.get("u").size, o.get("p").size
    nx, ny = len(o.get("dx")), len(o.get("dy"))
    numU = (nx - 1) * ny
    u = np.concatenate([np.ones(numU), np.zeros(nuv - numU)])
    o.set("u", u)
    o.set("uold", u)
    o.set("p", np.full(np_, 0.37))
    o.phase("force")
    f = o.info()
    assert abs(f["forceX"]) < 1e-12 and abs(f["forceY"]) < 1e-12
    # p = -2x: the pressure part of the drag is the pressure drop between the cell
    # centres on the control-volume faces times the face height
    x, dy = o.get("x"), o.get("dy")
    o.set("p", np.tile(-2.0 * x, ny))
    o.phase("force")
    f = o.info()
    I, J, W, H = (int(f[k]) for k in ("startI", "startJ", "numCellsX", "numCellsY"))
    assert np.isclose(f["fxx"], 2.0 * (x[I + W] - x[I]) * dy[J:J + H].sum(), rtol=1e-13)
    assert f["fxu"] == 0 and abs(f["forceY"]) < 1e-12


def test_oscillating_body_motion_matches_reference_run():
    """The reference's committed run of the in-line oscillating cylinder
    (validation/osc/static/external, fixture written by tests/golden/make_osc_fixtures.py)
    fixes the body motion after every step: x = -sin(2 pi t)/(2 pi) and u = -cos(2 pi t)
    evaluated at t = dt * (step - 1), because the move at the end of a step uses the step
    counter before it advances (luo_base.cu:255-264, luo_base/move.cu:11-36). Printed to
    six significant digits. The velocity solve's iteration counts are in the same file and
    must agree over the first 40 steps (2 per step)."""
    with open(os.path.join(GOLDEN, "osc_reference_series.json")) as f:
        ref = json.load(f)
    o = po.Oracle(po.read_case(case("oscCylinderRe100_luo")))
    n = 40
    mid, its = [], []
    for _ in range(n):
        o.step(1)
        i = o.info()
        mid.append(i["midX"])
        its.append(int(i["iter1"]))
    rx = np.array(ref["midX"][:n])
    assert np.all(np.abs(np.array(mid) - rx) <= 6e-6 * np.abs(rx) + 1e-15)
    # an off-by-one in the motion's time would shift x by dt * u = 1e-3 per step
    assert abs(mid[5] - ref["midX"][4]) > 1e-4
    assert its == ref["iter_velocity"][:n]


# ---- Luo embedded method (OSC_CYLINDER) -----------------------------------------------
# The embedded rows fold the ghost/hybrid interpolation into the matrices. What pins them
# independently of the assembly formulas: the interpolation weights reproduce bilinear
# fields (Neumann corners included), and constant fields are preserved row by row.

def _embedded(name="cylinderRe40_small_embedded", steps=1):
    o = po.Oracle(po.read_case(case(name)))
    o.step(steps)
    st = o.embedded_stats()
    assert all(v == 0 for d in st.values() for v in d.values()), st
    return o


def _emb_csr(o, which):
    rp, ci, v = o.get(f"LHS{which}_rowptr"), o.get(f"LHS{which}_col"), o.get(f"LHS{which}_val")
    return sp.csr_matrix((v, ci, rp), shape=(len(rp) - 1, len(rp) - 1))


def _grid(o):
    return {k: o.get(k) for k in ("x", "y", "xu", "yv")}


def _vel_pos(g, i, nx, ny):
    numU = (nx - 1) * ny
    if i < numU:
        return g["xu"][i % (nx - 1)], g["y"][i // (nx - 1)]
    k = i - numU
    return g["x"][k % nx], g["yv"][k // nx]


def test_embedded_velocity_weights_reproduce_bilinear_fields():
    """Ghost nodes: the corner equal to the node, or else (no ghost corner at all) the
    corner nearest the body intercept, sits at the intercept (biLinearInterpolation.cu:
    60-110); the weights evaluate at the image point. Hybrid nodes: the corner equal to
    the node sits at its intercept; the weights evaluate at the node."""
    o = _embedded()
    nx, ny = len(o.get("x")), len(o.get("y"))
    g = _grid(o)
    gt, ht = o.get("ghostTagsUV"), o.get("hybridTagsUV")
    bix, biy, ipx, ipy = o.get("body_intercept_x"), o.get("body_intercept_y"), o.get("image_point_x"), o.get("image_point_y")
    C = np.stack([o.get(f"qcoef{l}") for l in range(4)])
    Q = np.stack([o.get(f"qidx{l}") for l in range(4)])
    rng = np.random.default_rng(7)
    a, b, c, d = rng.standard_normal(4)
    f = lambda x, y: a + b * x + c * y + d * x * y
    nodes = np.nonzero((gt > 0) | (ht > 0))[0]
    assert len(nodes) > 20
    worst = 0.0
    for i in nodes:
        pos = [_vel_pos(g, int(Q[l, i]), nx, ny) for l in range(4)]
        if gt[i] > 0:
            inflag = any(gt[Q[l, i]] > 0 for l in range(4))
            mn, close = 1.0, -1
            for l in range(4):
                s = np.hypot(pos[l][0] - bix[i], pos[l][1] - biy[i])
                if s < mn:
                    mn, close = s, Q[l, i]
            for l in range(4):
                if Q[l, i] == i or (Q[l, i] == close and not inflag):
                    pos[l] = (bix[i], biy[i])
            target = f(ipx[i], ipy[i])
        else:
            for l in range(4):
                if Q[l, i] == i:
                    pos[l] = (bix[i], biy[i])
            target = f(*_vel_pos(g, i, nx, ny))
        val = sum(C[l, i] * f(*pos[l]) for l in range(4))
        worst = max(worst, abs(val - target) / (1 + abs(target)))
    assert worst < 1e-9, worst


def test_embedded_pressure_weights_reproduce_bilinear_fields_with_neumann_corners():
    """Replaced pressure corners carry the normal derivative along (image point - body
    intercept) of that corner (biLinearInterpolation.cu:740-770, 930-970): a bilinear
    field f gives the value grad f . n there, and the weights must return f at the node
    (hybrid) or at the image point (ghost)."""
    o = _embedded()
    x, y = o.get("x"), o.get("y")
    nx = len(x)
    gt, ht = o.get("ghostTagsP"), o.get("hybridTagsP")
    bx, by = o.get("body_intercept_p_x"), o.get("body_intercept_p_y")
    ix, iy = o.get("image_point_p_x"), o.get("image_point_p_y")
    C = np.stack([o.get(f"pqcoef{l}") for l in range(4)])
    Q = np.stack([o.get(f"pqidx{l}") for l in range(4)])
    rng = np.random.default_rng(8)
    a, b, c, d = rng.standard_normal(4)
    f = lambda X, Y: a + b * X + c * Y + d * X * Y
    grad = lambda X, Y: (b + d * Y, c + d * X)
    nodes = np.nonzero((gt > 0) | (ht > 0))[0]
    assert len(nodes) > 10
    worst = 0.0
    for i in nodes:
        val = 0.0
        for l in range(4):
            j = int(Q[l, i])
            neumann = (gt[j] > 0) if gt[i] > 0 else (j == i)
            if neumann:
                n = np.array([ix[j] - bx[j], iy[j] - by[j]])
                n /= np.linalg.norm(n)
                gx, gy = grad(bx[j], by[j])
                val += C[l, i] * (gx * n[0] + gy * n[1])
            else:
                val += C[l, i] * f(x[j % nx], y[j // nx])
        target = f(ix[i], iy[i]) if gt[i] > 0 else f(x[i % nx], y[i // nx])
        worst = max(worst, abs(val - target) / (1 + abs(target)))
    assert worst < 1e-9, worst


def test_embedded_rows_preserve_constants():
    """A constant field with the body moving at that constant is a solution of every
    row: interior velocity rows sum to ns_rhs + alpha * (weight of the node's own corner),
    ghost velocity rows sum to 2 - (weights of the corners kept explicit), interior and
    wall pressure rows sum to 0 (row 0 carries the doubled-diagonal pin). A dropped or
    sign-flipped interpolation entry breaks one of these."""
    o = _embedded()
    x, y = o.get("x"), o.get("y")
    nx, ny = len(x), len(y)
    numU = (nx - 1) * ny
    A1, A2 = _emb_csr(o, 1), _emb_csr(o, 2)
    s1, s2 = np.asarray(A1.sum(axis=1)).ravel(), np.asarray(A2.sum(axis=1)).ravel()
    gt, ht = o.get("ghostTagsUV"), o.get("hybridTagsUV")
    ns, al = o.get("ns_rhs1"), o.get("alpha1")
    C = np.stack([o.get(f"qcoef{l}") for l in range(4)])
    Q = np.stack([o.get(f"qidx{l}") for l in range(4)])
    checked = {"hybrid": 0, "ghost": 0, "plain": 0}
    for i in range(len(s1)):
        if i < numU:
            I, J, ni, nj = i % (nx - 1), i // (nx - 1), nx - 1, ny
        else:
            I, J, ni, nj = (i - numU) % nx, (i - numU) // nx, nx, ny - 1
        if I == 0 or I == ni - 1 or J == 0 or J == nj - 1:
            continue
        if ht[i] > 0:
            own = sum(C[l, i] for l in range(4) if Q[l, i] == i)
            expect = ns[i] + al[i] * own
            checked["hybrid"] += 1
        elif gt[i] > 0:
            expect = 2 - sum(C[l, i] for l in range(4) if gt[Q[l, i]] > 0 or Q[l, i] == i)
            checked["ghost"] += 1
        else:
            expect = ns[i]
            checked["plain"] += 1
        assert abs(s1[i] - expect) < 1e-12 * (1 + abs(expect)), (i, s1[i], expect)
    assert min(checked.values()) > 10, checked
    scale = np.asarray(abs(A2).sum(axis=1)).ravel()
    assert np.all(np.abs(s2[1:]) <= 1e-12 * scale[1:])
    assert s2[0] > 0


def test_embedded_plain_rows_are_the_navier_stokes_rows_over_their_diagonal():
    """Velocity rows away from the body are the plain LHS1 rows (pinned above) divided by
    their diagonal (LHS1.cu:162-207); pressure rows away from the body are the plain LHS2
    rows (LHS2.cu:171-222)."""
    o = _embedded()
    c = po.read_case(case("cylinderRe40_small_embedded"))
    c["sim"]["SolverType"] = "LUO"
    p = po.Oracle(c)
    E1, P1 = _emb_csr(o, 1), _emb_csr(p, 1)
    E2, P2 = _emb_csr(o, 2), _emb_csr(p, 2)
    gt, ht = o.get("ghostTagsUV"), o.get("hybridTagsUV")
    x, y = o.get("x"), o.get("y")
    nx, ny = len(x), len(y)
    numU = (nx - 1) * ny
    n_checked = 0
    for i in range(0, E1.shape[0], 7):
        if gt[i] != -1 or ht[i] != -1:
            continue
        if i < numU:
            I, J, ni, nj = i % (nx - 1), i // (nx - 1), nx - 1, ny
        else:
            I, J, ni, nj = (i - numU) % nx, (i - numU) // nx, nx, ny - 1
        er, pr = E1.getrow(i), P1.getrow(i)
        assert list(er.indices) == list(pr.indices)
        if I == 0 or I == ni - 1 or J == 0 or J == nj - 1:
            assert np.array_equal(er.data, pr.data)
        else:
            np.testing.assert_allclose(er.data, pr.data / pr[0, i], rtol=1e-14, atol=0)
        n_checked += 1
    assert n_checked > 100
    gp, hp = o.get("ghostTagsP"), o.get("hybridTagsP")
    far = np.nonzero((gp == -1) & (hp == -1))[0]
    d = (E2[far] - P2[far])
    assert abs(d).max() == 0


def test_embedded_poisson_hierarchy_restricts_with_scaled_transpose():
    """The embedded Poisson matrix mixes unit-diagonal interpolation rows with Poisson rows
    of diagonal 4 dt/h^2 (40 on the reference's 462x462 validation/error/cylinder/
    embedded4 case). With R = P^T its Galerkin level 1 has negative diagonals and the
    V-cycle amplifies; the embedded hierarchy restricts its finest level with P^T D^-1.
    Pins: A_1 equals P^T D^-1 A P formed here from the level's own P (a dropped or
    inverted D fails), every coarse diagonal is positive, and BiCGSTAB with it converges
    where R = P^T does not."""
    c = po.read_case(case("cylinderRe40_462_embedded"))
    o = po.Oracle(c)
    o.step(1)
    assert o.info()["iter2"] < 200
    A = _emb_csr(o, 2)
    d = A.diagonal()
    assert (np.abs(d - 1.0) < 1e-12).sum() > 100 and d.max() > 30  # rows of both scales
    try:
        po.set_amg_scale0(False)
        M0 = po.OrcCsr(A)
        assert (M0.level_csr(1).diagonal() <= 0).any()
        po.set_amg_scale0(True)
        M = po.OrcCsr(A)
        P = M.level_csr(0, 1)
        A1 = M.level_csr(1)
        ref = (P.T @ sp.diags(1.0 / d) @ A @ P).tocsr()
        diff = (A1 - ref).tocsr()
        assert np.abs(diff.data).max(initial=0.0) <= 1e-12 * np.abs(ref.data).max()
        for lev in range(1, len(M.amg_levels()[0])):
            assert (M.level_csr(lev).diagonal() > 0).all(), lev
        b = A @ np.random.default_rng(3).standard_normal(A.shape[0])
        x, it, r = M.bicgstab(b, precond="SMOOTHED_AGGREGATION", rtol=1e-6, maxit=300)
        assert 0 <= it < 100 and np.linalg.norm(b - A @ x) <= 1e-6 * np.linalg.norm(b) * 1.0001
    finally:
        po.set_amg_scale0(False)
