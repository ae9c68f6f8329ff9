"""cuibm_amd — MI355X (gfx950) rebuild of cuIBM's per-time-step hot path.

Python mirror of the reference's solver-class interface over the C ABI in
``include/cuibm_amd.h`` (``libcuibm_amd.so``, built in-tree).  The compute runs in
the hand-written HIP kernels of that library; there is no Python or CPU fallback:
importing the package works anywhere, but creating a solver needs the built
library and a GPU, and fails loudly otherwise.

Reference interfaces mirrored (src/solvers/NavierStokes/NavierStokesSolver.h:145-186,
src/cuIBM.cu:42-74)::

    s = Solver("cases/cylinderRe40_luo")        # io::readInputs + factory
    s.initialise()
    while not s.finished():
        s.step_time()
        s.write_data()
    s.shut_down()
"""
import ctypes as C
import os

import numpy as np

__all__ = ["Solver", "Csr", "CuibmError", "lib", "LIB_PATH", "CASES"]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcuibm_amd.so")
CASES = os.path.join(HERE, "cases")

PRECONDITIONERS = {"NONE": 0, "DIAGONAL": 1, "SMOOTHED_AGGREGATION": 2}
INT_FIELDS = {"ghostTagsUV", "hybridTagsUV", "hybridTagsUV2", "ghostTagsP", "hybridTagsP",
              "LHS1_rowptr", "LHS1_col", "LHS2_rowptr", "LHS2_col", "tag_records_own", "tag_records_flags"}
INFO_KEYS = ["timeStep", "iter1", "iter2", "res1", "res2", "forceX", "forceY", "fxx", "fxy", "fxu",
             "startI", "startJ", "numCellsX", "numCellsY", "midX", "midY", "amg_levels", "nx", "ny",
             "tol1", "tol2", "n_ghost_uv", "n_hybrid_uv", "n_ghost_p", "n_hybrid_p", "n_solid",
             "iter1_total", "iter2_total", "velocity_bytes_total", "poisson_bytes_total", "poisson_solves",
             "krylov_restarts",
             # Luo embedded assembly, slot anomalies summed over all assemblies (LHS1, then LHS2):
             # 5-point slots a ghost row leaves unwritten, hybrid tail slots never written,
             # tail slots written twice, interpolation corners outside the 3x3 stencil
             "emb1_unwritten", "emb1_tail_unwritten", "emb1_tail_multi", "emb1_dropped",
             "emb2_unwritten", "emb2_tail_unwritten", "emb2_tail_multi", "emb2_dropped",
             # passes of the last step (strongly coupled VIV repeats the step), transverse
             # velocity of the body centre (VIV)
             "sc_passes", "centerV",
             # largest CFL number so far (NavierStokes/CFL.cu:10-35), its cell and time step
             "cfl_max", "cfl_I", "cfl_J", "cfl_ts",
             # moving embedded body: hierarchies set up again on the device (same pattern)
             "amg_refreshes"]


class CuibmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cuibm error {code}: {msg}")
        self.code = code


_lib = None
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)


def lib():
    """Load libcuibm_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C cuibm_amd/csrc")
    L = C.CDLL(LIB_PATH)
    L.cuibm_version.restype = C.c_int
    L.cuibm_last_error.restype = C.c_char_p
    L.cuibm_create.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_void_p)]
    for f in ("cuibm_initialise", "cuibm_step_time", "cuibm_write_data", "cuibm_finished", "cuibm_shut_down"):
        getattr(L, f).argtypes = [C.c_void_p]
        getattr(L, f).restype = C.c_int
    L.cuibm_destroy.argtypes = [C.c_void_p]
    L.cuibm_destroy.restype = None
    L.cuibm_name.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
    L.cuibm_phase.argtypes = [C.c_void_p, C.c_char_p]
    L.cuibm_case_array.argtypes = [C.c_void_p, C.c_char_p, _DP, C.c_long]
    L.cuibm_case_array.restype = C.c_long
    L.cuibm_case_scalars.argtypes = [C.c_void_p, _DP, C.c_int]
    L.cuibm_field_size.argtypes = [C.c_void_p, C.c_char_p]
    L.cuibm_field_size.restype = C.c_long
    L.cuibm_get_field.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_long]
    L.cuibm_get_field.restype = C.c_long
    L.cuibm_set_field.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_long]
    L.cuibm_set_field.restype = C.c_long
    L.cuibm_info.argtypes = [C.c_void_p, _DP, C.c_int]
    L.cuibm_timers.restype = C.c_long
    L.cuibm_timers.argtypes = [C.c_void_p, C.c_char_p, C.c_long]
    L.cuibm_bytes.argtypes = [C.c_void_p, C.c_int]
    L.cuibm_device_synchronize.argtypes = []
    L.cuibm_device_synchronize.restype = C.c_int
    L.cuibm_bytes.restype = C.c_double
    L.cuibm_time_steps.argtypes = [C.c_void_p, C.c_int, _DP]
    L.cuibm_time_kernel.argtypes = [C.c_void_p, C.c_char_p, C.c_int, _DP, _DP]
    L.cuibm_csr_create.argtypes = [C.c_int, _IP, _IP, _DP, C.POINTER(C.c_void_p)]
    L.cuibm_csr_destroy.argtypes = [C.c_void_p]
    L.cuibm_csr_destroy.restype = None
    L.cuibm_csr_spmv.argtypes = [C.c_void_p, _DP, _DP]
    L.cuibm_bicgstab.argtypes = [C.c_void_p, _DP, _DP, C.c_int, C.c_int, C.c_double, C.c_int, _IP, _DP]
    L.cuibm_bicgstab_noalloc.argtypes = [C.c_void_p, _DP, _DP, C.c_int, C.c_double, C.c_int, _IP, _DP]
    L.cuibm_amg_levels.argtypes = [C.c_void_p, _IP, _IP, C.c_int]
    L.cuibm_amg_vcycle.argtypes = [C.c_void_p, C.c_int, _DP, _DP]
    L.cuibm_ensemble_run.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.c_int, C.POINTER(C.c_char_p), C.c_int, C.c_int,
                                     C.POINTER(C.c_longlong), _DP, _DP, _DP, _DP, C.c_char_p, C.c_int]
    L.cuibm_ensemble_run.restype = C.c_int
    L.cuibm_ensemble_lockstep.argtypes = L.cuibm_ensemble_run.argtypes
    L.cuibm_ensemble_lockstep.restype = C.c_int
    L.cuibm_launch_bytes.argtypes = [C.c_void_p, C.c_char_p, _DP, C.c_long]
    L.cuibm_launch_bytes.restype = C.c_long
    L.cuibm_set_body_velocity.argtypes = [C.c_void_p, _DP, _DP, C.c_int, C.c_double]
    L.cuibm_adjoint_run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _DP, _DP, _DP, _DP, _DP]
    _lib = L
    return L


def _check(rc):
    if rc < 0:
        raise CuibmError(rc, lib().cuibm_last_error().decode(errors="replace"))
    return rc


def device_synchronize():
    """wait for all work on the current device (hipDeviceSynchronize)"""
    _check(lib().cuibm_device_synchronize())


def _dp(a):
    return a.ctypes.data_as(_DP)


def ensemble_run(case_folders, *args, steps=0, workers=8, lockstep=False):
    """Run independent cases to completion on one GPU; `args` are common command-line
    overrides. Default (cuibm_ensemble_run): each case on its own solver and HIP stream,
    driven by `workers` host threads. lockstep=True (cuibm_ensemble_lockstep): every phase
    of the step is one launch over all cases, each case's solves stopping at its own
    tolerance (`workers` threads run the set-up). Returns a dict with per-case cells, cd,
    cl and the wall times."""
    n = len(case_folders)
    folders = (C.c_char_p * n)(*[os.fsencode(str(f)) for f in case_folders])
    argv = (C.c_char_p * max(len(args), 1))(*[os.fsencode(str(a)) for a in args])
    cells = np.zeros(n, dtype=np.int64)
    cd, cl = np.zeros(n), np.zeros(n)
    wall, wall_steps = C.c_double(0), C.c_double(0)
    err = C.create_string_buffer(1024)
    fn = lib().cuibm_ensemble_lockstep if lockstep else lib().cuibm_ensemble_run
    rc = fn(n, folders, len(args), argv, int(steps), int(workers),
                                  cells.ctypes.data_as(C.POINTER(C.c_longlong)), _dp(cd), _dp(cl), C.byref(wall),
                                  C.byref(wall_steps), err, len(err))
    if rc < 0:
        raise CuibmError(rc, err.value.decode(errors="replace"))
    return {"cells": cells, "cd": cd, "cl": cl, "wall_s": wall.value, "steps_wall_s": wall_steps.value}


class Solver:
    """One cuIBM case on the GPU (NavierStokesSolver / luoIBM / fadlunModified)."""

    def __init__(self, case_folder, *overrides, output=False, output_folder=None, diagnostics=True, **kw):
        args = [str(a) for a in overrides]
        for k, v in kw.items():
            args += [f"-{k}", str(v)]
        args += ["-output", "1" if output else "0", "-diagnostics", "1" if diagnostics else "0"]
        if output_folder:
            args += ["-outputFolder", str(output_folder)]
        argv = (C.c_char_p * len(args))(*[a.encode() for a in args])
        h = C.c_void_p()
        _check(lib().cuibm_create(str(case_folder).encode(), len(args), argv, C.byref(h)))
        self._h = h
        self.case_folder = str(case_folder)

    def close(self):
        if getattr(self, "_h", None):
            lib().cuibm_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # reference lifecycle
    def initialise(self):
        _check(lib().cuibm_initialise(self._h))
        return self

    def step_time(self, n=1):
        for _ in range(n):
            _check(lib().cuibm_step_time(self._h))

    def write_data(self):
        _check(lib().cuibm_write_data(self._h))

    def finished(self):
        return bool(_check(lib().cuibm_finished(self._h)))

    def shut_down(self):
        _check(lib().cuibm_shut_down(self._h))

    def name(self):
        buf = C.create_string_buffer(64)
        _check(lib().cuibm_name(self._h, buf, 64))
        return buf.value.decode()

    def phase(self, name):
        _check(lib().cuibm_phase(self._h, name.encode()))

    # parsed case (host side, usable before initialise and without a GPU)
    def case_array(self, name):
        n = _check(lib().cuibm_case_array(self._h, name.encode(), None, 0))
        a = np.zeros(n)
        _check(lib().cuibm_case_array(self._h, name.encode(), _dp(a), n))
        return a

    CASE_KEYS = ["nx", "ny", "nu", "dt", "nt", "nsave", "SolverType", "scaleCV", "startStep", "tol1", "tol2",
                 "maxit1", "maxit2", "precond1", "precond2", "amgPrecision", "uInitial", "vInitial", "uPerturb",
                 "vPerturb", "bc_xm_u", "bc_xm_v", "bc_xp_u", "bc_xp_v", "bc_ym_u", "bc_ym_v", "bc_yp_u", "bc_yp_v",
                 "mid_h", "nbodies"]

    def case_scalars(self):
        out = np.zeros(64)
        n = _check(lib().cuibm_case_scalars(self._h, _dp(out), 64))
        return dict(zip(self.CASE_KEYS, out[:n]))

    # data access
    def get(self, name):
        n = _check(lib().cuibm_field_size(self._h, name.encode()))
        a = np.zeros(n, dtype=np.int32 if name in INT_FIELDS else np.float64)
        _check(lib().cuibm_get_field(self._h, name.encode(), a.ctypes.data_as(C.c_void_p), n))
        return a

    def set(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.int32 if name in INT_FIELDS else np.float64)
        _check(lib().cuibm_set_field(self._h, name.encode(), a.ctypes.data_as(C.c_void_p), a.size))

    def info(self):
        out = np.zeros(47)
        _check(lib().cuibm_info(self._h, _dp(out), 47))
        return dict(zip(INFO_KEYS, out[: len(INFO_KEYS)]))

    def timers(self):
        """Accumulated wall-clock seconds per phase, with the reference's event names."""
        n = lib().cuibm_timers(self._h, None, 0)
        _check(n if n < 0 else 0)
        buf = C.create_string_buffer(int(n) + 1)
        lib().cuibm_timers(self._h, buf, n + 1)
        out = {}
        for line in buf.value.decode().splitlines():
            name, sec = line.rsplit("\t", 1)
            out[name] = float(sec)
        return out

    def bytes(self, which=0):
        return lib().cuibm_bytes(self._h, which)

    def time_steps(self, n):
        ms = C.c_double(0)
        _check(lib().cuibm_time_steps(self._h, n, C.byref(ms)))
        return ms.value

    # discrete adjoint (LUO solver, static body)
    def set_body_velocity(self, ub, vb, omega=0.0):
        """per-point body velocities ub[k], vb[k] and the rotation rate of the solid nodes
        (cuibm_set_body_velocity), after initialise()"""
        ub = np.ascontiguousarray(ub, dtype=np.float64)
        vb = np.ascontiguousarray(vb, dtype=np.float64)
        if ub.shape != vb.shape or ub.ndim != 1:
            raise ValueError("ub and vb must be 1-D arrays of one length")
        _check(lib().cuibm_set_body_velocity(self._h, _dp(ub), _dp(vb), ub.size, float(omega)))

    def adjoint_run(self, nsteps, n0, n1, every, backward=True):
        """J = mean of C_D over steps n0..n1 of an nsteps run from time step 0 and, with
        backward, dJ/dub, dJ/dvb (per body point) and dJ/domega (cuibm_adjoint_run)"""
        npts = self.case_array("body_X").size
        J = C.c_double(0)
        gu, gv = np.zeros(npts), np.zeros(npts)
        gom = C.c_double(0)
        st = np.zeros(8)
        _check(lib().cuibm_adjoint_run(self._h, int(nsteps), int(n0), int(n1), int(every), int(bool(backward)),
                                       C.byref(J), _dp(gu), _dp(gv), C.byref(gom), _dp(st)))
        out = {"J": J.value, "forward_s": st[0], "forward_poisson_its": st[2], "forward_velocity_its": st[3]}
        if backward:
            out.update({"grad_ub": gu, "grad_vb": gv, "grad_omega": gom.value, "backward_s": st[1],
                        "adjoint_poisson_its": st[4], "adjoint_velocity_its": st[5], "backward_setup_s": st[6],
                        "backward_recompute_s": st[7]})
        return out

    def launch_bytes(self, what="poisson_iteration"):
        """algorithmic bytes of each launch of one Poisson BiCGSTAB iteration"""
        n = _check(lib().cuibm_launch_bytes(self._h, what.encode(), None, 0))
        a = np.zeros(n)
        _check(lib().cuibm_launch_bytes(self._h, what.encode(), _dp(a), n))
        return a

    def time_kernel(self, kernel, reps=20):
        ms = C.c_double(0)
        b = C.c_double(0)
        _check(lib().cuibm_time_kernel(self._h, kernel.encode(), reps, C.byref(ms), C.byref(b)))
        return ms.value, b.value


class Csr:
    """A square CSR matrix on the GPU with the BiCGSTAB / AMG entry points."""

    def __init__(self, A):
        A = A.tocsr()
        A.sort_indices()
        self.n = A.shape[0]
        rp = np.ascontiguousarray(A.indptr, dtype=np.int32)
        ci = np.ascontiguousarray(A.indices, dtype=np.int32)
        v = np.ascontiguousarray(A.data, dtype=np.float64)
        h = C.c_void_p()
        _check(lib().cuibm_csr_create(self.n, rp.ctypes.data_as(_IP), ci.ctypes.data_as(_IP), _dp(v), C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            lib().cuibm_csr_destroy(self._h)
            self._h = None

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        _check(lib().cuibm_csr_spmv(self._h, _dp(x), _dp(y)))
        return y

    def bicgstab(self, b, x0=None, precond="NONE", rtol=1e-10, maxit=1000, amg_precision=64):
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        it = C.c_int(0)
        r = C.c_double(0)
        rc = lib().cuibm_bicgstab(self._h, _dp(x), _dp(b), PRECONDITIONERS[precond], amg_precision, rtol, maxit,
                                  C.byref(it), C.byref(r))
        if rc < 0 and rc != -4:
            _check(rc)
        return x, it.value, r.value, rc == 0

    def bicgstab_noalloc(self, b, x0=None, precond="NONE", rtol=1e-10, maxit=1000):
        """the same solve as bicgstab() (NONE or DIAGONAL) with every device buffer in the
        library's fixed scratch array: no device allocation, n <= 2**18"""
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        it = C.c_int(0)
        r = C.c_double(0)
        rc = lib().cuibm_bicgstab_noalloc(self._h, _dp(x), _dp(b), PRECONDITIONERS[precond], rtol, maxit,
                                          C.byref(it), C.byref(r))
        if rc < 0 and rc != -4:
            _check(rc)
        return x, it.value, r.value, rc == 0

    def amg_levels(self):
        s = np.zeros(64, dtype=np.int32)
        z = np.zeros(64, dtype=np.int32)
        L = _check(lib().cuibm_amg_levels(self._h, s.ctypes.data_as(_IP), z.ctypes.data_as(_IP), 64))
        return s[:L], z[:L]

    def vcycle(self, b, amg_precision=64):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n)
        _check(lib().cuibm_amg_vcycle(self._h, amg_precision, _dp(b), _dp(x)))
        return x
