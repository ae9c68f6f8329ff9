#!/usr/bin/env python3
"""Time-step throughput of the cuIBM per-step hot path on one MI355X.

Workload: impulsively started circular cylinder at Re=550 (case
``cuibm_amd/cases/cylinderRe550_1024_luo``): D=1 at the origin, [-15,15]^2 with
h=0.005 on [-0.6,0.6]^2 stretched by 1.008178 to 1024x1024 cells, Luo external
ghost-node IBM with 628 body points, dt=0.0025. Velocity solve: BiCGSTAB + Jacobi,
rtol 1e-9. Poisson solve: FP64 BiCGSTAB, rtol 1e-9, preconditioned by the
smoothed-aggregation V-cycle (2+2 Jacobi sweeps, omega 0.8) run in FP32 (matrix values,
level vectors and smoother arithmetic), with the operator's pinned constant mode removed
and restored in FP64 around it (the headline "mixed" form); the outer residual, the
Krylov dot products and the projection stay FP64. The same steps are also run with an
FP64 hierarchy and with FP64 CSR operators everywhere, and the drag coefficients of the
forms are compared.

The discrete adjoint (``adjoint``): the gradient of the mean drag over t = 2..4 of the
rotating cylinder at Re=100 (case ``cuibm_amd/cases/rotatingCylinderRe100_luo``: 498x498
stretched grid with h=0.01 near the body, 256 body points with tangential velocity
alpha (D/2) omega_ref, alpha=0.5, omega_ref=2, plus seeded N(0, 0.01^2) normal blowing,
solves to 1e-12) with respect to alpha and the 256 blowing values, checkpoints every 20
steps; seconds per gradient = the forward run plus the backward sweep. At dt=0.01 this
grid puts the convective CFL number at 2.2 and the start diverges within 50 steps; the
case runs dt=0.005 (CFL 1.1), 800 steps, mean over steps 400..800.

A step is one full ``stepTime`` (rhs1, velocity solve, rhs2, Poisson solve, IBM
interpolation/blending, projection, control-volume force) on fields resident in
HBM. W untimed warm-up steps are followed by exactly K timed steps, bracketed by
device synchronisation and measured with HIP events on the solver's stream. With
dt=0.0025 and h=0.005 the convective CFL number reaches about 1 at the cylinder
shoulders, beyond the Adams-Bashforth limit, and the impulsive start diverges after
about 55 steps in this method (GPU and CPU restatement alike); the default window
(steps 6-45) lies before that.

Prints one JSON line. ``roofline`` reports the dominant kernel of the step (the
fused finest-level pre-smoothing + residual kernel of the Poisson V-cycle, one per
V-cycle and two V-cycles per BiCGSTAB iteration), timed with a HIP start/stop event
pair attached to each of its launches (hipExtLaunchKernel, which stamps the kernel's
own begin and end) over a few live time steps of a second instance of the same form,
after the same warm-up (the timed window itself carries no probes); ``poisson`` is the
Poisson solve alone, timed likewise with events before and after each whole solve;
``kernels`` lists back-to-back timings of single kernels on the live inputs for
comparison; ``traffic`` is the HBM byte count per launch of that kernel from hardware
counters, read from the committed summary under profiles/ when that summary was taken on
a kernel with the same algorithmic bytes per launch (null otherwise). ``cpu_baseline``
is the FP64 C++ CPU restatement of the same step (the oracle) timed on this host over a
short sample of the same workload, with the host's CPU count and model.
"""
import argparse
import json
import os
import re
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = ("time steps/s + HBM GB/s at 1024² (mixed vs FP64, same C_D accuracy); adjoint s/grad")
CASE = "cylinderRe550_1024_luo"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(HERE, "profiles", "pmc_vc_pre_cylinder1024.md")
# kernels timed beside the roofline kernel: the Poisson operator product with the
# BiCGSTAB dots, one whole V-cycle, the fused intermediate-velocity right-hand side
EXTRA_KERNELS = ["vc_post", "spmv_dot_lhs2", "vcycle", "rhs1", "jacobi0"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_form(cu, case, amg_precision, warmup, steps, kernel, reps, probe_steps=5, extra=(), median=False):
    """initialise, warm up, time `steps` steps; return a dict of figures"""
    t0 = time.perf_counter()
    s = cu.Solver(case, "-amgPrecision", str(amg_precision), *extra, diagnostics=False)
    s.initialise()
    t_init = time.perf_counter() - t0
    if warmup > 0:
        s.step_time(warmup)
    i0 = s.info()
    cu.device_synchronize()
    w0 = time.perf_counter()
    ms = s.time_steps(steps)
    cu.device_synchronize()
    wall = time.perf_counter() - w0
    i1 = s.info()
    d = lambda k: i1[k] - i0[k]
    out = {
        "amg_precision": amg_precision,
        "init_s": round(t_init, 3),
        "ms_per_step": ms / steps,
        "wall_ms_per_step": 1e3 * wall / steps,
        "steps_per_s": 1e3 * steps / ms,
        "velocity_iters_per_step": d("iter1_total") / steps,
        "poisson_iters_per_step": d("iter2_total") / steps,
        "poisson_bytes_per_step": d("poisson_bytes_total") / steps,
        "velocity_bytes_per_step": d("velocity_bytes_total") / steps,
        "cd": 2.0 * i1["forceX"],
        "cl": 2.0 * i1["forceY"],
        "time_step": int(i1["timeStep"]),
    }
    # step bytes = the two solves (counted per iteration as run) + the stencil sweeps
    out["step_bytes"] = s.bytes(0) - s.bytes(1) - s.bytes(2) + out["poisson_bytes_per_step"] + \
        out["velocity_bytes_per_step"]
    out["hbm_gbps_step"] = out["step_bytes"] / (out["ms_per_step"] * 1e-3) / 1e9
    s.close()
    if median:
        # the same window once more on a fresh instance, each step timed on its own (HIP
        # events around the step, the stream drained between steps): median and spread
        s = cu.Solver(case, "-amgPrecision", str(amg_precision), *extra, diagnostics=False)
        s.initialise()
        if warmup > 0:
            s.step_time(warmup)
        per = sorted(s.time_steps(1) for _ in range(steps))
        s.close()
        out["per_step_ms"] = {"median": per[len(per) // 2], "min": per[0], "max": per[-1],
                              "p90": per[int(0.9 * (len(per) - 1))]}
    if kernel:
        # probes on a second instance of the same form, over the same early steps (W + 1
        # onwards), so that the timed window above carries no events: the roofline kernel
        # (HIP events attached to each of its launches) and every whole Poisson solve
        # (events before and after each solve, host waits inside it included)
        s = cu.Solver(case, "-amgPrecision", str(amg_precision), *extra, diagnostics=False)
        s.initialise()
        if warmup > 0:
            s.step_time(warmup)
        kms, kbytes = s.time_kernel("live:" + kernel, probe_steps)
        out["kernel"] = {"name": kernel, "ms": kms, "bytes": kbytes, "steps": probe_steps}
        i0 = s.info()
        pms, pbytes = s.time_kernel("live:poisson", probe_steps)
        its = (s.info()["iter2_total"] - i0["iter2_total"]) / max(probe_steps, 1)
        out["poisson"] = {"ms_per_solve": pms, "bytes_per_solve": pbytes, "steps": probe_steps,
                          "iterations_per_solve": its, "gbps": pbytes / (pms * 1e-3) / 1e9 if pms > 0 else 0.0}
        out["poisson"].update(copy_ceiling(s, s.launch_bytes(), pms, its))
        out["kernels"] = {}
        if reps > 0:  # back-to-back launches of single kernels on the live inputs
            for k in dict.fromkeys([kernel] + EXTRA_KERNELS):
                kms, kbytes = s.time_kernel(k, reps)
                out["kernels"][k] = {"us": kms * 1e3, "bytes": kbytes, "gbps": kbytes / (kms * 1e-3) / 1e9}
    s.close()
    return out


def copy_ceiling(s, launches, ms_per_solve, its):
    """The rate a plain copy kernel reaches at each launch size (back-to-back launches of
    copy:N), and the time one Poisson iteration would take if each of its launches ran at
    the copy's rate for its size: the attainable floor for launches of these sizes, set by
    the dependent-dispatch gap at the small ones and by HBM at the large ones."""
    import numpy as np

    curve = []
    for k in range(10, 27):  # 2^10 .. 2^26 doubles per array: 16 KB .. 1 GB moved
        ms, b = s.time_kernel(f"copy:{2 ** k}", 50)
        curve.append((b, ms))
    cb = np.log([c[0] for c in curve])
    ct = np.log([c[1] for c in curve])
    floor = min(c[1] for c in curve)

    def t_of(b):  # ms of a copy moving b bytes (log-log interpolation; the smallest copy's time below)
        return max(floor, float(np.exp(np.interp(np.log(max(b, 1.0)), cb, ct))))

    ceil_ms = sum(t_of(b) for b in launches)
    meas_ms = ms_per_solve / its if its > 0 else 0.0
    return {
        "launches_per_iteration": len(launches),
        "bytes_per_iteration": float(sum(launches)),
        "ms_per_iteration": meas_ms,
        "copy_ceiling_ms_per_iteration": ceil_ms,
        "frac_of_copy_ceiling": ceil_ms / meas_ms if meas_ms > 0 else None,
        "copy_ceiling_gbps": sum(launches) / (ceil_ms * 1e-3) / 1e9,
        "copy_curve": [{"bytes": b, "us": round(ms * 1e3, 3), "gbps": round(b / (ms * 1e-3) / 1e9, 1)}
                       for b, ms in curve],
    }


def read_profile(key, bytes_per_launch=None):
    """a figure of the roofline kernel from the committed counter / trace summary; None
    when the summary was taken on a kernel whose algorithmic bytes per launch differ from
    the live kernel's (a stale summary)"""
    try:
        with open(TRAFFIC_FILE) as f:
            text = f.read()
    except OSError:
        return None
    if bytes_per_launch is not None:
        m = re.search(r"algorithmic_bytes_per_launch\s*[:=]\s*([0-9.eE+]+)", text)
        if not m or abs(float(m.group(1)) - bytes_per_launch) > 1e-6 * bytes_per_launch:
            return None
    m = re.search(key + r"\s*[:=]\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def host_cpu():
    """logical CPUs and the model name of this host (what lscpu prints)"""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def cpu_baseline(case_dir, steps):
    """time the FP64 oracle (the CPU restatement of the same step) on this host"""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import pyoracle as po

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    c = po.read_case(case_dir)  # the oracle's hierarchy is always FP64
    o = po.Oracle(c)
    o.step(1)  # warm-up step (first-touch of the work arrays)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    o.close()
    return {"value": steps / dt, "unit": "time steps/s", "cores": threads, "kind": "port", "host": host_cpu(),
            "sample": f"{steps} time steps (after 1 warm-up step) of the same 1024x1024 Luo cylinder case with "
                      f"the FP64 C++ oracle (FP64 SA-AMG), OpenMP threads={threads}; "
                      f"{dt:.1f} s of CPU work; initialisation excluded"}


# The reference's own published run times (scripts/validation/performance.py:22-24 on a Tesla
# K20; scripts/validation/performance2.py:24-25, hardware not stated): whole runs of the
# same step, solver tolerance 1e-5, FP64 smoothed aggregation. Case folders restate
# validation/error/cylinder/{fadlun4,external4,embedded4} and
# validation/error/oscflow2/{external05,embedded05}.
PUBLISHED = [
    ("cylinderRe40_890_fadlun", "modified Fadlun, Re=40 cylinder, 890x890, 1000 steps", 563.0, "Tesla K20"),
    ("cylinderRe40_890_luo", "Luo external, Re=40 cylinder, 890x890, 1000 steps", 174.0, "Tesla K20"),
    ("cylinderRe40_462_embedded", "Luo embedded, Re=40 cylinder, 462x462, 1000 steps", 762.0, "Tesla K20"),
    ("oscFlowRe100_luo", "Luo external, oscillating cylinder in flow, 224x224, 2000 steps", 61.0, "not stated"),
    ("oscFlowRe100_embedded", "Luo embedded, oscillating cylinder in flow, 224x224, 2000 steps", 581.0,
     "not stated"),
]


def published_runs(cu, only=None):
    """each published workload run in full (nt steps from the impulsive start), FP64 hierarchy;
    the per-step phase times come from the solver's own timers"""
    rows = []
    for name, what, ref_s, hw in PUBLISHED:
        if only and name not in only:
            continue
        t0 = time.perf_counter()
        s = cu.Solver(os.path.join(cu.CASES, name), "-amgPrecision", "64", diagnostics=False)
        s.initialise()
        t_init = time.perf_counter() - t0
        nt = int(s.case_scalars()["nt"])
        cu.device_synchronize()
        w0 = time.perf_counter()
        ms, done = 0.0, 0
        while done < nt:  # in tenths, with a progress line each
            k = min(max(nt // 10, 1), nt - done)
            ms += s.time_steps(k)
            done += k
            log(f"[bench]   {name}: {done}/{nt} steps, {time.perf_counter() - w0:.1f} s")
        cu.device_synchronize()
        wall = time.perf_counter() - w0
        i1 = s.info()
        phases = {k: round(1e3 * v / nt, 4) for k, v in s.timers().items() if v > 0}
        s.close()
        run_s = t_init + wall
        rows.append({"case": name, "workload": what, "steps": nt, "init_s": round(t_init, 3),
                     "steps_s": round(wall, 3), "steps_per_s": round(1e3 * nt / ms, 2), "run_s": round(run_s, 3),
                     "published_s": ref_s, "published_hardware": hw, "speedup_vs_published": round(ref_s / run_s, 1),
                     "poisson_iters_per_step": round(i1["iter2_total"] / nt, 2), "cd": 2.0 * i1["forceX"],
                     "amg_refreshes": int(i1["amg_refreshes"]), "phase_ms_per_step": phases})
        log(f"[bench] {name}: {nt} steps in {wall:.2f} s (+{t_init:.2f} s set-up), published {ref_s:.0f} s; "
            f"Update Solver {phases.get('Update Solver', 0.0):.3f} ms/step, Poisson Solve "
            f"{phases.get('Poisson Solve', 0.0):.3f} ms/step")
    return rows


def adjoint_bench(cu, nsteps=800):
    """one objective + gradient of the rotating cylinder (cuibm_amd.adjoint); J over the
    second half of the nsteps steps"""
    import numpy as np
    from cuibm_amd.adjoint import RotatingCylinder

    p = RotatingCylinder(os.path.join(cu.CASES, "rotatingCylinderRe100_luo"), nsteps=nsteps, n0=nsteps // 2,
                         n1=nsteps, every=20)
    blow = np.random.default_rng(2024).normal(0.0, 0.01, p.npts)
    t0 = time.perf_counter()
    J, ga, gb = p.evaluate(0.5, blow)
    wall = time.perf_counter() - t0
    r = p.last
    log(f"[bench] adjoint: {wall:.2f} s per gradient (forward {r['forward_s']:.2f} s, backward {r['backward_s']:.2f} s)")
    return {
        "s_per_grad": round(wall, 3),
        "forward_s": round(r["forward_s"], 3),
        "backward_s": round(r["backward_s"], 3),
        "backward_setup_s": round(r["backward_setup_s"], 3),
        "backward_recompute_s": round(r["backward_recompute_s"], 3),
        "forward_poisson_its": r["forward_poisson_its"],
        "adjoint_poisson_its": r["adjoint_poisson_its"],
        "adjoint_velocity_its": r["adjoint_velocity_its"],
        "J": J,
        "dJ_dalpha": ga,
        "dJ_dblow_norm": float(np.linalg.norm(gb)),
        "workload": f"rotatingCylinderRe100_luo: 498x498, 256 points, alpha=0.5 + seeded blowing, {nsteps} steps "
                    f"of dt=0.005, J = mean C_D over steps {nsteps // 2}..{nsteps}, gradient w.r.t. alpha and 256 "
                    f"blowing values, checkpoints every 20 steps, FP64, solves to 1e-12",
        "timing": "wall clock of one call (fresh solver set-up, forward run, backward sweep)",
    }


def ensemble_bench(cu, n, steps, workers, check, prec="32", modes=("lockstep", "eager", "concurrent")):
    """n seeded cylinder cases of 96^2..512^2 cells and Re 40..200 (cuibm_amd.ensemble),
    run in lockstep (every phase one launch over all cases, the solves' iteration blocks as
    captured graphs), in lockstep with every kernel launched one by one ("eager") and, for
    comparison, as concurrent solvers on their own streams; every (n // check)-th case is
    then run alone with the FP64 hierarchy and its drag coefficient compared with them"""
    import tempfile

    import numpy as np
    from cuibm_amd import ensemble as ens

    folder = tempfile.mkdtemp(prefix="cuibm_ensemble_")
    paths, cases = ens.write_cases(folder, n=n, steps=steps)
    cells = sum(c[0] ** 2 for c in cases)
    log(f"[bench] ensemble: {n} cases, {cells} cells, {steps} steps each, AMG precision {prec}")
    r = rc = None
    rate = rate_c = None
    if "lockstep" in modes:
        r = cu.ensemble_run(paths, "-amgPrecision", prec, steps=steps, workers=workers, lockstep=True)
        rate = cells * steps / r["steps_wall_s"]
        log(f"[bench] lockstep: {rate / 1e6:.1f} M cell-updates/s over the steps ({r['steps_wall_s']:.2f} s), "
            f"{r['wall_s']:.2f} s with set-up")
    re_ = rate_e = None
    if "eager" in modes:
        re_ = cu.ensemble_run(paths, "-amgPrecision", prec, "-lockstepGraphs", "0", steps=steps, workers=workers,
                              lockstep=True)
        rate_e = cells * steps / re_["steps_wall_s"]
        same = r is not None and np.array_equal(r["cd"], re_["cd"]) and np.array_equal(r["cl"], re_["cl"])
        log(f"[bench] lockstep, kernels launched one by one: {rate_e / 1e6:.1f} M cell-updates/s "
            f"({re_['steps_wall_s']:.2f} s); C_D and C_L {'equal' if same else 'differ from'} the graph run's")
    if "concurrent" in modes:
        rc = cu.ensemble_run(paths, "-amgPrecision", prec, steps=steps, workers=workers)
        rate_c = cells * steps / rc["steps_wall_s"]
        log(f"[bench] concurrent solvers ({workers} host threads): {rate_c / 1e6:.1f} M cell-updates/s "
            f"({rc['steps_wall_s']:.2f} s)")
    if r is None:
        r, rate = (re_, rate_e) if re_ is not None else (rc, rate_c)
    picked = list(range(0, n, max(1, n // check)))[:check] if check > 0 else []
    dcd, dcd_c, t_alone = [], [], 0.0
    for i in picked:
        s = cu.Solver(paths[i], "-amgPrecision", "64", diagnostics=False)
        s.initialise()
        ms = s.time_steps(steps)
        t_alone += ms * 1e-3
        cd64 = 2.0 * s.info()["forceX"]
        dcd.append(abs(cd64 - r["cd"][i]))
        if rc is not None:
            dcd_c.append(abs(cd64 - rc["cd"][i]))
        s.close()
        log(f"[bench]   case {i} ({cases[i][0]}^2, Re {cases[i][1]:.1f}) alone in FP64: |dC_D| {dcd[-1]:.2e}")
    alone_rate = sum(cases[i][0] ** 2 for i in picked) * steps / t_alone if picked else None
    import shutil
    shutil.rmtree(folder, ignore_errors=True)
    return {
        "metric": "aggregate cell-updates/s, ensemble of independent cylinder cases on one GPU",
        "value": round(rate, 1), "unit": "cell-updates/s", "higher_is_better": True, "n_gpus": 1,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{n} impulsively started cylinders, edge 96..512 cells (seeded), Re 40..200, "
                               f"h=D/40 on [-0.6,0.6] stretched to [-10,10], dt=0.01, Luo external, Poisson FP64 "
                               f"BiCGSTAB rtol 1e-8 + SA-AMG V-cycle (precision {prec}), velocity BiCGSTAB+Jacobi "
                               f"rtol 1e-8; lockstep: every phase one launch over all cases",
                   "cases": n, "steps": steps, "cells": cells, "workers": workers, "amg_precision": prec},
        "steps_wall_s": round(r["steps_wall_s"], 3), "wall_s_with_setup": round(r["wall_s"], 3),
        "modes": list(modes),
        "eager_lockstep_cell_updates_per_s": round(rate_e, 1) if rate_e else None,
        "graphs_equal_eager": bool(np.array_equal(r["cd"], re_["cd"]) and np.array_equal(r["cl"], re_["cl"]))
        if re_ is not None and "lockstep" in modes else None,
        "concurrent_cell_updates_per_s": round(rate_c, 1) if rate_c else None,
        "concurrent_steps_wall_s": round(rc["steps_wall_s"], 3) if rc else None,
        "speedup_lockstep_vs_concurrent": round(rate / rate_c, 3) if rate_c and "lockstep" in modes else None,
        "alone_fp64_cell_updates_per_s": round(alone_rate, 1) if alone_rate else None,
        "check": {"cases": picked, "max_abs_dcd_vs_alone_fp64": max(dcd) if dcd else None,
                  "max_abs_dcd_concurrent_vs_alone_fp64": max(dcd_c) if dcd_c else None},
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--kernel", default="vc_pre", help="roofline kernel (see cuibm_time_kernel)")
    ap.add_argument("--kernel-reps", type=int, default=100,
                    help="back-to-back launches per kernel for the kernel table (0: skip the table)")
    ap.add_argument("--probe-steps", type=int, default=5,
                    help="live time steps (after the warm-up, on a second instance) over which the roofline "
                         "kernel and the Poisson solve are timed")
    ap.add_argument("--cpu-steps", type=int, default=10, help="oracle steps for the CPU baseline (0: skip)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-hierarchy comparison run")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the discrete-adjoint gradient timing")
    ap.add_argument("--kernels-only", action="store_true",
                    help="initialise the case, run two time steps and the copy calibration kernel "
                         "(for hardware-counter runs)")
    ap.add_argument("--kernels-case", default="",
                    help="with --kernels-only: this case folder instead, and its operator products timed")
    ap.add_argument("--published", action="store_true",
                    help="run the reference's published whole-run workloads in full and print their times "
                         "beside the published ones (one JSON line; not the headline metric)")
    ap.add_argument("--published-cases", default="",
                    help="comma-separated case folders of --published to run (default: all five)")
    ap.add_argument("--ensemble", type=int, default=0, metavar="N",
                    help="run N seeded cylinder cases concurrently and print the aggregate cell-update rate "
                         "(one JSON line; not the headline)")
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--ensemble-workers", type=int, default=16)
    ap.add_argument("--ensemble-check", type=int, default=8, help="cases re-run alone in FP64 for the C_D check")
    ap.add_argument("--ensemble-precision", default="32", help="AMG precision of the ensemble cases (64, 32, 3264)")
    ap.add_argument("--ensemble-modes", default="lockstep,eager,concurrent",
                    help="comma list of ensemble forms to run: lockstep (iteration blocks as graphs), eager "
                         "(lockstep, kernels launched one by one), concurrent")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="time one discrete-adjoint gradient only (one JSON line; not the headline)")
    ap.add_argument("--adjoint-steps", type=int, default=800, help="time steps of the adjoint workload")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only: the 1024x1024 step runs on a single MI355X "
              f"(--gpus {args.gpus}, WORLD_SIZE {world} requested); multi-GPU sharding is not built.",
              file=sys.stderr)
        sys.exit(2)

    import cuibm_amd as cu

    case = os.path.join(cu.CASES, CASE)
    if args.ensemble > 0:
        print(json.dumps(ensemble_bench(cu, args.ensemble, args.ensemble_steps, args.ensemble_workers,
                                        args.ensemble_check, args.ensemble_precision,
                                        tuple(args.ensemble_modes.split(",")))), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps({"metric": "seconds per discrete-adjoint gradient", "unit": "s", "higher_is_better": False,
                          "n_gpus": 1, "dtype": "f64", "adjoint": adjoint_bench(cu, args.adjoint_steps)}), flush=True)
        return
    if args.published:
        rows = published_runs(cu, args.published_cases.split(",") if args.published_cases else None)
        print(json.dumps({"metric": "seconds per published whole run (set-up + all steps, no file output)",
                          "unit": "s", "higher_is_better": False, "n_gpus": 1, "dtype": "f64",
                          "runs": rows}), flush=True)
        return
    if args.kernels_only:
        # two live time steps (their kernels are the ones the counters should see), then,
        # for another case, the two operator products with their dots as launched in the
        # solves (1 + reps launches each, in this order), then the copy calibration kernel
        other = bool(args.kernels_case)
        s = cu.Solver(os.path.join(cu.CASES, args.kernels_case) if other else case, diagnostics=False)
        s.initialise()
        s.step_time(2)
        out = {}
        for k in (("spmv_dot_lhs1", "spmv_dot_lhs2") if other else ()) + ("copy",):
            ms, b = s.time_kernel(k, args.kernel_reps)
            out[k] = {"us": ms * 1e3, "bytes": b, "gbps": b / (ms * 1e-3) / 1e9}
        s.close()
        print(json.dumps(out), flush=True)
        return
    log(f"[bench] host: {host_cpu()}")
    log(f"[bench] mixed form (FP32 V-cycle, FP64 deflation and Krylov): warmup {args.warmup}, steps {args.steps}")
    mixed = run_form(cu, case, 32, args.warmup, args.steps, args.kernel, args.kernel_reps, args.probe_steps,
                     median=True)
    log(f"[bench] mixed: {mixed['steps_per_s']:.2f} steps/s, Poisson its/step {mixed['poisson_iters_per_step']:.1f}")
    fp64 = None
    if not args.no_fp64:
        log("[bench] FP64 form")
        fp64 = run_form(cu, case, 64, args.warmup, args.steps, None, 0, median=True)
        log(f"[bench] fp64: {fp64['steps_per_s']:.2f} steps/s, Poisson its/step {fp64['poisson_iters_per_step']:.1f}")
    csr = None
    if not args.no_fp64:
        log("[bench] FP64 form with stored CSR operators at every level (no stencil forms)")
        csr = run_form(cu, case, 64, args.warmup, args.steps, None, 0, extra=("-operators", "csr"))
        log(f"[bench] fp64 csr: {csr['steps_per_s']:.2f} steps/s, Poisson its/step {csr['poisson_iters_per_step']:.1f}")
    adj = None
    if not args.no_adjoint:
        log("[bench] discrete adjoint of the rotating cylinder (498x498, 800 steps)")
        adj = adjoint_bench(cu, args.adjoint_steps)
    cpu = None
    if args.cpu_steps > 0:
        log(f"[bench] CPU baseline: oracle, {args.cpu_steps} steps")
        cpu = cpu_baseline(case, args.cpu_steps)
        log(f"[bench] cpu: {cpu['value']:.3f} steps/s")

    k = mixed["kernel"]
    achieved = k["bytes"] / (k["ms"] * 1e-3) / 1e9 if k["ms"] > 0 else 0.0
    line = {
        "metric": METRIC,
        "value": round(mixed["steps_per_s"], 4),
        "unit": "time steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms_per_step"], 4),
        "per_step_ms": {k: round(v, 4) for k, v in mixed["per_step_ms"].items()},
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "cylinderRe550_1024_luo: Re=550 cylinder, 1024x1024 stretched MAC grid, Luo external "
                        "IBM, dt=0.0025; Poisson FP64 BiCGSTAB rtol 1e-9 preconditioned by an FP32 SA-AMG V-cycle "
                        "(values, vectors, smoothers) with FP64 removal of the pinned constant mode; velocity FP64 "
                        "BiCGSTAB+Jacobi rtol 1e-9",
            "grid": "1024x1024",
            "time_step_range": [mixed["time_step"] - args.steps + 1, mixed["time_step"]],
        },
        "hbm_gbps_step": round(mixed["hbm_gbps_step"], 1),
        "poisson": {
            "gbps": round(mixed["poisson"]["gbps"], 1),
            "frac_of_peak": round(mixed["poisson"]["gbps"] / HBM_PEAK_GBS, 4),
            "ms_per_solve": round(mixed["poisson"]["ms_per_solve"], 4),
            "bytes_per_solve": mixed["poisson"]["bytes_per_solve"],
            "iterations_per_solve": mixed["poisson"]["iterations_per_solve"],
            "ms_per_iteration": round(mixed["poisson"]["ms_per_iteration"], 4),
            "copy_ceiling_ms_per_iteration": round(mixed["poisson"]["copy_ceiling_ms_per_iteration"], 4),
            "frac_of_copy_ceiling": mixed["poisson"]["frac_of_copy_ceiling"],
            "copy_ceiling_gbps": round(mixed["poisson"]["copy_ceiling_gbps"], 1),
            "launches_per_iteration": mixed["poisson"]["launches_per_iteration"],
            "copy_curve": mixed["poisson"]["copy_curve"],
            "timing": f"HIP events before and after each whole Poisson solve (host waits inside it included) "
                      f"over {mixed['poisson']['steps']} live time steps after the warm-up, on a second instance",
        },
        "poisson_iters_per_step": mixed["poisson_iters_per_step"],
        "velocity_iters_per_step": mixed["velocity_iters_per_step"],
        "cd": mixed["cd"],
        "cl": mixed["cl"],
        "roofline": {
            "kernel": k["name"],
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": read_profile("traffic_bytes_per_launch", k["bytes"]),
            "rocprof_us_per_launch": read_profile("rocprof_us_per_launch", k["bytes"]),
            "bytes_per_launch": k["bytes"],
            "us_per_launch": round(k["ms"] * 1e3, 2),
            "timing": f"HIP start/stop events attached to each in-step launch (hipExtLaunchKernel: the kernel's own "
                      f"begin and end) over {k['steps']} live time steps after the warm-up, on a second instance",
        },
        "kernels": {n: {kk: round(vv, 2) for kk, vv in d.items()} for n, d in mixed["kernels"].items()},
        "cpu_baseline": cpu,
        "adjoint": adj,
    }
    if fp64:
        line["fp64"] = {
            "steps_per_s": round(fp64["steps_per_s"], 4),
            "ms_per_step": round(fp64["ms_per_step"], 4),
            "per_step_ms": {k: round(v, 4) for k, v in fp64["per_step_ms"].items()},
            "poisson_iters_per_step": fp64["poisson_iters_per_step"],
            "hbm_gbps_step": round(fp64["hbm_gbps_step"], 1),
            "cd": fp64["cd"],
            "cl": fp64["cl"],
        }
        line["speedup_mixed_vs_fp64"] = round(mixed["steps_per_s"] / fp64["steps_per_s"], 4)
        line["speedup_mixed_vs_fp64_median"] = round(fp64["per_step_ms"]["median"] / mixed["per_step_ms"]["median"], 4)
        line["speedup_target"] = 1.7
        line["abs_dcd_mixed_vs_fp64"] = abs(mixed["cd"] - fp64["cd"])
        # where the time goes: the Poisson solve's share of each form's step, and its bytes
        line["poisson_bytes_per_step"] = {"mixed": mixed["poisson_bytes_per_step"], "fp64": fp64["poisson_bytes_per_step"]}
        line["step_bytes"] = {"mixed": mixed["step_bytes"], "fp64": fp64["step_bytes"]}
    if csr:
        line["fp64_csr"] = {
            "steps_per_s": round(csr["steps_per_s"], 4),
            "ms_per_step": round(csr["ms_per_step"], 4),
            "poisson_iters_per_step": csr["poisson_iters_per_step"],
            "cd": csr["cd"],
            "form": "FP64 hierarchy, CSR operators for both solves and every AMG level (-operators csr)",
        }
        line["speedup_mixed_vs_fp64_csr"] = round(mixed["steps_per_s"] / csr["steps_per_s"], 4)
        line["abs_dcd_mixed_vs_fp64_csr"] = abs(mixed["cd"] - csr["cd"])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
