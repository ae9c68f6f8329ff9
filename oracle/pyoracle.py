"""ORACLE — test infrastructure only.

ctypes front end of ``liboracle.so`` (the FP64 CPU restatement of cuIBM's hot
path, see ``orc.hpp``).  Only the tests, ``__graft_entry__.smoke()`` and the
bench's ``cpu_baseline`` leg import this module, and only to compare against it.

Case files are read with PyYAML here (the product has its own C++ reader); the
keys and defaults follow src/io/parseSimulationFile.cu:71-182,
src/io/parseFlowFile.cu:79-128, src/io/parseDomainFile.cu:28-110 and
src/io/parseBodiesFile.cu:30-178.
"""
import ctypes as C
import os

import numpy as np
import yaml

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

SOLVER_TYPES = {"NAVIER_STOKES_SOLVER": 0, "FADLUN": 1, "OSC_CYLINDER": 2, "FSI": 3, "LUO": 4}
PRECONDS = {"NONE": 0, "DIAGONAL": 1, "SMOOTHED_AGGREGATION": 2, "AINV": 3}
LOCATIONS = {"xMinus": 0, "xPlus": 1, "yMinus": 2, "yPlus": 3}

_DP = C.POINTER(C.c_double)


class OrcParams(C.Structure):
    _fields_ = [
        ("solver_type", C.c_int), ("nx", C.c_int), ("ny", C.c_int),
        ("x", _DP), ("dx", _DP), ("y", _DP), ("dy", _DP),
        ("mid_h", C.c_double), ("nu", C.c_double), ("dt", C.c_double), ("nt", C.c_int),
        ("uInitial", C.c_double), ("vInitial", C.c_double),
        ("uPerturb", C.c_double), ("vPerturb", C.c_double),
        ("bc_u", C.c_double * 4), ("bc_v", C.c_double * 4),
        ("precond1", C.c_int), ("precond2", C.c_int),
        ("tol1", C.c_double), ("tol2", C.c_double),
        ("maxit1", C.c_int), ("maxit2", C.c_int),
        ("scaleCV", C.c_double), ("nbody", C.c_int), ("bx", _DP), ("by", _DP),
        ("xCoeff", C.c_double), ("uCoeff", C.c_double), ("xPhase", C.c_double), ("uPhase", C.c_double),
        ("xfreq", C.c_double), ("viv", C.c_int), ("ured", C.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        L.orc_axis_cells.restype = C.c_int
        L.orc_axis_cells.argtypes = [C.c_double, C.c_double, C.c_int, _DP, _DP]
        L.orc_axis_fill.argtypes = [C.c_double, C.c_double, C.c_int, _DP, _DP, _DP, _DP]
        L.orc_circle_points.restype = C.c_int
        L.orc_circle_points.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                        C.c_double, _DP, _DP]
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_phase.restype = C.c_int
        L.orc_phase.argtypes = [C.c_void_p, C.c_char_p]
        L.orc_get.restype = C.c_long
        L.orc_get.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_long]
        L.orc_set.restype = C.c_long
        L.orc_set.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_long]
        L.orc_info.argtypes = [C.c_void_p, _DP]
        L.orc_csr.restype = C.c_void_p
        L.orc_csr.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _DP]
        L.orc_csr_free.argtypes = [C.c_void_p]
        L.orc_bicgstab.restype = C.c_int
        L.orc_bicgstab.argtypes = [C.c_void_p, _DP, _DP, C.c_int, C.c_double, C.c_int, _DP]
        L.orc_amg_levels.restype = C.c_int
        L.orc_amg_levels.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int]
        L.orc_amg_vcycle.restype = C.c_int
        L.orc_amg_vcycle.argtypes = [C.c_void_p, _DP, _DP]
        L.orc_set_amg_coarse_max.argtypes = [C.c_int]
        L.orc_set_amg_coarse_max.restype = None
        L.orc_amg_scale0.restype = None
        L.orc_amg_scale0.argtypes = [C.c_int]
        L.orc_amg_level_csr.restype = C.c_int
        L.orc_amg_level_csr.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _DP,
                                        C.c_int, C.POINTER(C.c_int)]
        L.orc_amg_coarse_dense.restype = C.c_int
        L.orc_amg_coarse_dense.argtypes = [C.c_void_p, C.c_int, _DP, C.c_int]
        L.orc_amg_prolongator_dense.restype = C.c_int
        L.orc_amg_prolongator_dense.argtypes = [C.c_void_p, _DP, C.c_int]
        L.orc_viv_update.restype = None
        L.orc_viv_update.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, C.c_double,
                                     _DP, _DP]
        L.orc_set_body_velocity.restype = C.c_int
        L.orc_set_body_velocity.argtypes = [C.c_void_p, _DP, _DP, C.c_int, C.c_double]
        L.orc_embedded_stats.restype = None
        L.orc_embedded_stats.argtypes = [C.c_void_p, C.POINTER(C.c_long)]
        _lib = L
    return _lib


def dptr(a):
    return a.ctypes.data_as(_DP)


def axis(start, mid_h, ends, ratios):
    """Cell centres and widths of one direction (parseDomainFile.cu:28-222)."""
    e = np.ascontiguousarray(ends, dtype=np.float64)
    r = np.ascontiguousarray(ratios, dtype=np.float64)
    n = lib().orc_axis_cells(start, mid_h, len(e), dptr(e), dptr(r))
    x = np.zeros(n)
    dx = np.zeros(n)
    lib().orc_axis_fill(start, mid_h, len(e), dptr(e), dptr(r), dptr(x), dptr(dx))
    return x, dx


def circle(cx, cy, R, n, self_size=False, mid_h=0.0):
    """Circle body points (parseBodiesFile.cu:145-178)."""
    m = lib().orc_circle_points(cx, cy, R, n, int(bool(self_size)), mid_h, None, None)
    X = np.zeros(m)
    Y = np.zeros(m)
    lib().orc_circle_points(cx, cy, R, n, int(bool(self_size)), mid_h, dptr(X), dptr(Y))
    return X, Y


def _load(folder, name):
    with open(os.path.join(folder, name)) as f:
        return yaml.safe_load(f)


def read_case(folder, overrides=None):
    """Parse the four case files into a plain dict (reference keys and defaults)."""
    sim = {"dt": 0.02, "scaleCV": 2.0, "nt": 100, "nsave": 100, "SolverType": "NAVIER_STOKES_SOLVER",
           "VIV": 0, "Ured": 3,
           "velocity": {"preconditioner": "DIAGONAL", "tolerance": 1e-5, "maxIterations": 10000},
           "Poisson": {"preconditioner": "DIAGONAL", "tolerance": 1e-5, "maxIterations": 20000}}
    for node in _load(folder, "simParams.yaml"):
        for k in ("dt", "scaleCV", "nt", "nsave", "SolverType", "VIV", "Ured"):
            if k in node:
                sim[k] = node[k]
        prec, tol, mit = "DIAGONAL", 1e-5, 10000
        for s in node.get("linearSolvers", []):
            prec = s.get("preconditioner", prec)
            tol = float(s.get("tolerance", tol))
            mit = int(s.get("maxIterations", mit))
            sim[s["system"]] = {"preconditioner": prec, "tolerance": tol, "maxIterations": mit}
    flow = {"nu": 0.01, "uInitial": 1.0, "vInitial": 0.0, "uPerturb": 0.0, "vPerturb": 0.0,
            "bc_u": [0.0] * 4, "bc_v": [0.0] * 4}
    for node in _load(folder, "flow.yaml"):
        flow["nu"] = float(node["nu"])
        flow["uInitial"], flow["vInitial"] = map(float, node["initialVelocity"][:2])
        if "initialPerturbation" in node:
            flow["uPerturb"], flow["vPerturb"] = map(float, node["initialPerturbation"][:2])
        for b in node.get("boundaryConditions", []):
            loc = LOCATIONS[b["location"]]
            flow["bc_u"][loc] = float(b["u"][1])
            flow["bc_v"][loc] = float(b["v"][1])
    grid = {}
    for node in _load(folder, "domain.yaml"):
        ends = [float(s["end"]) for s in node["subDomains"]]
        ratios = [float(s["stretchRatio"]) for s in node["subDomains"]]
        x, dx = axis(float(node["start"]), float(node["mid_h"]), ends, ratios)
        grid[node["direction"]] = (x, dx)
        grid["mid_h"] = float(node["mid_h"])
    bodies = []
    motion = {"xCoeff": 0.0, "uCoeff": 0.0, "xPhase": 0.0, "uPhase": 0.0, "xfreq": 0.0}
    bpath = os.path.join(folder, "bodies.yaml")
    if os.path.exists(bpath):
        for ib, node in enumerate(_load(folder, "bodies.yaml") or []):
            if ib == 0:  # the motion of the first body drives all points (src/bodies.cu:27-35)
                for key, k in (("xCoefficient", "xCoeff"), ("uCoefficient", "uCoeff"), ("xPhase", "xPhase"),
                               ("uPhase", "uPhase"), ("xfrequency", "xfreq")):
                    if key in node:
                        motion[k] = float(node[key])
            if node.get("type") == "circle":
                o = node["circleOptions"]
                self_size = bool(o[4]) if len(o) > 4 else False
                mid_h = float(o[5]) if len(o) > 5 else 0.0
                bodies.append(circle(float(o[0]), float(o[1]), float(o[2]), int(o[3]), self_size, mid_h))
    case = {"sim": sim, "flow": flow, "grid": grid, "bodies": bodies, "motion": motion}
    for k, v in (overrides or {}).items():
        if k in ("nu",):
            flow[k] = v
        elif k in ("velocityTol",):
            sim["velocity"]["tolerance"] = v
        elif k in ("poissonTol",):
            sim["Poisson"]["tolerance"] = v
        else:
            sim[k] = v
    return case


def viv_update(yold, vold, cy, ured, dt, strong=False, ycur=0.0):
    """one spring-mass step of a VIV body (luo_base/move.cu:38-93): (vnew, ynew)"""
    v, y = C.c_double(0), C.c_double(0)
    lib().orc_viv_update(yold, vold, cy, int(ured), dt, int(bool(strong)), ycur, C.byref(v), C.byref(y))
    return v.value, y.value


def set_amg_scale0(on):
    """hierarchies of OrcCsr (inspection and bicgstab) restrict the finest level with
    P^T D^-1 while on (the embedded method's hierarchy)"""
    lib().orc_amg_scale0(1 if on else 0)


def set_amg_coarse_max(n):
    """coarsest AMG level size for hierarchies set up from now on (n <= 0: 2500)"""
    lib().orc_set_amg_coarse_max(int(n))


class Oracle:
    """One oracle solver instance for a parsed case (see read_case)."""

    def __init__(self, case):
        self.case = case
        g = case["grid"]
        self._keep = []
        x, dx = [np.ascontiguousarray(a) for a in g["x"]]
        y, dy = [np.ascontiguousarray(a) for a in g["y"]]
        self._keep += [x, dx, y, dy]
        s, f = case["sim"], case["flow"]
        p = OrcParams()
        p.solver_type = SOLVER_TYPES.get(s["SolverType"], 0)
        p.nx, p.ny = len(x), len(y)
        p.x, p.dx, p.y, p.dy = dptr(x), dptr(dx), dptr(y), dptr(dy)
        p.mid_h = g["mid_h"]
        p.nu, p.dt, p.nt = f["nu"], float(s["dt"]), int(s["nt"])
        p.uInitial, p.vInitial = f["uInitial"], f["vInitial"]
        p.uPerturb, p.vPerturb = f["uPerturb"], f["vPerturb"]
        for i in range(4):
            p.bc_u[i] = f["bc_u"][i]
            p.bc_v[i] = f["bc_v"][i]
        p.precond1 = PRECONDS[s["velocity"]["preconditioner"]]
        p.precond2 = PRECONDS[s["Poisson"]["preconditioner"]]
        p.tol1 = s["velocity"]["tolerance"]
        p.tol2 = s["Poisson"]["tolerance"]
        p.maxit1 = s["velocity"]["maxIterations"]
        p.maxit2 = s["Poisson"]["maxIterations"]
        p.scaleCV = float(s["scaleCV"])
        m = case.get("motion", {})
        p.xCoeff, p.uCoeff = m.get("xCoeff", 0.0), m.get("uCoeff", 0.0)
        p.xPhase, p.uPhase, p.xfreq = m.get("xPhase", 0.0), m.get("uPhase", 0.0), m.get("xfreq", 0.0)
        p.viv, p.ured = int(s.get("VIV", 0)), int(s.get("Ured", 3))
        if case["bodies"]:
            bx = np.concatenate([b[0] for b in case["bodies"]])
            by = np.concatenate([b[1] for b in case["bodies"]])
            self._keep += [bx, by]
            p.nbody, p.bx, p.by = len(bx), dptr(bx), dptr(by)
        else:
            p.nbody = 0
        self.nx, self.ny = p.nx, p.ny
        self.h = lib().orc_create(C.byref(p))

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    __del__ = close

    def step(self, n=1):
        lib().orc_step(self.h, n)

    def phase(self, name):
        if lib().orc_phase(self.h, name.encode()) != 0:
            raise ValueError(name)

    _INT = {"ghostTagsUV", "hybridTagsUV", "hybridTagsUV2", "ghostTagsP", "hybridTagsP",
            "LHS1_rowptr", "LHS1_col", "LHS2_rowptr", "LHS2_col",
            "qidx0", "qidx1", "qidx2", "qidx3", "pqidx0", "pqidx1", "pqidx2", "pqidx3"}

    def get(self, name):
        n = lib().orc_get(self.h, name.encode(), None, 0)
        if n < 0:
            raise KeyError(name)
        a = np.zeros(n, dtype=np.int32 if name in self._INT else np.float64)
        lib().orc_get(self.h, name.encode(), a.ctypes.data_as(C.c_void_p), n)
        return a

    def set(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.int32 if name in self._INT else np.float64)
        if lib().orc_set(self.h, name.encode(), a.ctypes.data_as(C.c_void_p), a.size) < 0:
            raise KeyError(name)

    def info(self):
        out = np.zeros(32)
        lib().orc_info(self.h, dptr(out))
        keys = ["timeStep", "iter1", "iter2", "res1", "res2", "forceX", "forceY", "fxx", "fxy", "fxu",
                "startI", "startJ", "numCellsX", "numCellsY", "midX", "midY", "amg_levels", "sc_passes", "centerV"]
        return dict(zip(keys, out[:len(keys)]))

    def set_body_velocity(self, ub, vb, omega=0.0):
        """per-point body velocities and the solid nodes' rotation rate (LUO cases)"""
        ub = np.ascontiguousarray(ub, dtype=np.float64)
        vb = np.ascontiguousarray(vb, dtype=np.float64)
        if lib().orc_set_body_velocity(self.h, dptr(ub), dptr(vb), ub.size, float(omega)) != 0:
            raise ValueError("per-point body velocities need a LUO case and one value per body point")

    def embedded_stats(self):
        """Slot anomalies of the embedded assembly summed over all steps: for LHS1 then
        LHS2, 5-point slots left unwritten by ghost rows, hybrid tail slots never written,
        tail slots written more than once, interpolation corners outside the 3x3 stencil."""
        out = (C.c_long * 8)()
        lib().orc_embedded_stats(self.h, out)
        keys = ["unwritten", "tail_unwritten", "tail_multi", "dropped"]
        return {"LHS1": dict(zip(keys, out[0:4])), "LHS2": dict(zip(keys, out[4:8]))}


class OrcCsr:
    """A CSR matrix held by the oracle (for the linear-solver tests)."""

    def __init__(self, A):
        A = A.tocsr()
        A.sort_indices()
        self.n = A.shape[0]
        self.rp = np.ascontiguousarray(A.indptr, dtype=np.int32)
        self.ci = np.ascontiguousarray(A.indices, dtype=np.int32)
        self.v = np.ascontiguousarray(A.data, dtype=np.float64)
        IP = C.POINTER(C.c_int)
        self.h = lib().orc_csr(self.n, self.rp.ctypes.data_as(IP), self.ci.ctypes.data_as(IP), dptr(self.v))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_csr_free(self.h)
            self.h = None

    def bicgstab(self, b, x0=None, precond="NONE", rtol=1e-10, maxit=1000):
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        r = C.c_double(0)
        it = lib().orc_bicgstab(self.h, dptr(x), dptr(b), PRECONDS[precond], rtol, maxit, C.byref(r))
        return x, it, r.value

    def amg_levels(self):
        s = np.zeros(64, dtype=np.int32)
        z = np.zeros(64, dtype=np.int32)
        IP = C.POINTER(C.c_int)
        L = lib().orc_amg_levels(self.h, s.ctypes.data_as(IP), z.ctypes.data_as(IP), 64)
        return s[:L], z[:L]

    def vcycle(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n)
        lib().orc_amg_vcycle(self.h, dptr(b), dptr(x))
        return x

    def level_csr(self, level, which=0, cap=50_000_000):
        """hierarchy matrix of level `level` as scipy CSR (which: 0 A, 1 P, 2 R)"""
        import scipy.sparse as sp
        IP = C.POINTER(C.c_int)
        rp = np.zeros(cap, dtype=np.int32)
        ci = np.zeros(cap, dtype=np.int32)
        v = np.zeros(cap)
        nnz = C.c_int(0)
        n = lib().orc_amg_level_csr(self.h, level, which, rp.ctypes.data_as(IP), ci.ctypes.data_as(IP), dptr(v), cap,
                                    C.byref(nnz))
        if n < 0 or nnz.value > cap:
            raise ValueError(f"level {level} unavailable (n={n}, nnz={nnz.value})")
        ncols = int(ci[: nnz.value].max()) + 1 if nnz.value else 0
        return sp.csr_matrix((v[: nnz.value], ci[: nnz.value], rp[: n + 1]), shape=(n, max(ncols, n if which == 0 else ncols)))

    def coarse_dense(self, level=1, cap=4_000_000):
        """dense Galerkin operator of hierarchy level `level` (-1: coarsest)"""
        out = np.zeros(cap)
        n1 = lib().orc_amg_coarse_dense(self.h, level, dptr(out), cap)
        return out[: n1 * n1].reshape(n1, n1)

    def prolongator_dense(self, n1, cap=4_000_000):
        out = np.zeros(cap)
        lib().orc_amg_prolongator_dense(self.h, dptr(out), cap)
        return out[: self.n * n1].reshape(self.n, n1)
