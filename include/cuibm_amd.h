/*
 * cuibm_amd — C ABI of the MI355X (gfx950) rebuild of cuIBM's per-time-step hot path.
 *
 * The reference has no FFI: its hot path sits behind two C++ interfaces, the
 * solver-class API (src/solvers/NavierStokes/NavierStokesSolver.h:145-186,
 * luo_base.h:216-243, the factory switch in src/cuIBM.cu:42-60 and the driver loop
 * src/cuIBM.cu:62-74) and the CUSP linear-operator API
 * (cusp::krylov::bicgstab(A, x, b, monitor, M) at NavierStokesSolver.cu:227-228 and
 * :256-257, preconditioner<coo_matrix> in src/preconditioner.h:35-195). Each entry
 * point below names the reference interface it replaces.
 *
 * Conventions: every function returns 0 (or a count) on success and a negative
 * code on failure (the reference calls exit(-1) instead); cuibm_last_error() gives
 * the message. The library owns device memory, callers own host buffers. One HIP
 * stream per solver; a solver handle is not thread-safe. Field layouts are the
 * reference's (packed u then v velocities, iu = J*(nx-1)+I, iv = numU + J*nx+I;
 * pressure ip = J*nx+I).
 */
#ifndef CUIBM_AMD_H
#define CUIBM_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
#define CUIBM_OK 0
#define CUIBM_ERR_GENERIC (-1)
#define CUIBM_ERR_INPUT (-2)
#define CUIBM_ERR_HIP (-3)
#define CUIBM_ERR_NOT_CONVERGED (-4)  /* NavierStokesSolver.cu:232-240, 260-268 exit(-1) */
#define CUIBM_ERR_UNSUPPORTED (-5)
#define CUIBM_ERR_ARGUMENT (-6)

typedef struct cuibm_solver cuibm_solver;
typedef struct cuibm_csr cuibm_csr;

/* library version (major*10000 + minor*100 + patch) */
int cuibm_version(void);

/* message of the last failed call on this thread */
const char *cuibm_last_error(void);

/* ---- solver lifecycle ----------------------------------------------------------- */

/* io::readInputs (src/io/io.cu:108-138) + the SolverType factory (src/cuIBM.cu:42-60):
 * read <case_folder>/{simParams,flow,domain,bodies}.yaml, apply the command-line
 * overrides in argv (-nu -uPerturb -vPerturb -scaleCV -nsave -nt -dt -velocityTol
 * -poissonTol, plus -amgPrecision 64|32|3264 (as cuibm_bicgstab's amg_precision; in a
 * solver whose Poisson operator has its stencil form, 32 also removes the pinned constant
 * mode in FP64 around the FP32 V-cycle), -amgCoarseMax N (largest coarsest AMG level, solved
 * with a dense inverse: 16..6144, default 2500; 60 for a moving embedded body),
 * -amgRefresh 0|1 (a moving embedded body's Poisson hierarchy is built on a pattern padded
 * around the body's rows and, while each new matrix keeps to it, set up again numerically
 * on the device with its aggregates kept: 1, the default; or built anew on the host every
 * step, as the reference does: 0), -poissonRestart F (the Poisson BiCGSTAB restarts from its
 * true residual each time its residual has fallen by F since the last restart, 0 <= F < 1,
 * 0: never; default 1e-3 with the FP32 V-cycle, whose rounding makes it a slightly different
 * map at every apply, and never otherwise), -poissonPreconditioner NAME,
 * -operators csr|stencil, -outputFolder DIR and -output 0|1) and construct the solver
 * (not yet initialised). */
int cuibm_create(const char *case_folder, int argc, const char *const *argv, cuibm_solver **out);

/* NavierStokesSolver::initialise() (NavierStokesSolver.cu:34-40, luoIBM.cu:27-52,
 * fadlunModified.cu:28-43): fields, grid, bodies, tags, matrices, preconditioners */
int cuibm_initialise(cuibm_solver *s);

/* NavierStokesSolver::stepTime() (NavierStokesSolver.cu:163-188, luo_base.cu:208-233);
 * a LUO body with a prescribed oscillation is moved and re-tagged at the end of the
 * step (luo_base.cu:137-174, luo_base/move.cu:11-36) */
int cuibm_step_time(cuibm_solver *s);

/* NavierStokesSolver::writeData() (NavierStokesSolver.cu:356-363, luoIBM.cu:69-80,
 * fadlunModified.cu:64-77): q/lambda every nsave steps, iterations, forces */
int cuibm_write_data(cuibm_solver *s);

/* NavierStokesSolver::finished() (NavierStokesSolver.cu:195-199): 1 when done, 0 if not */
int cuibm_finished(cuibm_solver *s);

/* NavierStokesSolver::shutDown() (NavierStokesSolver.cu:368-372) */
int cuibm_shut_down(cuibm_solver *s);

/* release the solver and its device memory */
void cuibm_destroy(cuibm_solver *s);

/* NavierStokesSolver::name() */
int cuibm_name(cuibm_solver *s, char *buf, int len);

/* ---- case front end (host only; valid right after cuibm_create, no GPU needed) --- */

/* parsed case arrays (src/domain.h:17-34, src/io/parseBodiesFile.cu:145-178): "x", "dx",
 * "y", "dy", "xu", "yv" (grid) and "body_X", "body_Y" (body points as read, before
 * re-centring). Copies at most n values; returns the array length. */
long cuibm_case_array(cuibm_solver *s, const char *name, double *out, long n);

/* parsed case scalars: out[0] nx, [1] ny, [2] nu, [3] dt, [4] nt, [5] nsave,
 * [6] SolverType (src/types.h:4-11), [7] scaleCV, [8] startStep, [9..10] velocity
 * and Poisson tolerance, [11..12] their maxIterations, [13..14] their preconditioner
 * (src/preconditioner.h:23-29), [15] AMG precision, [16] uInitial, [17] vInitial,
 * [18] uPerturb, [19] vPerturb, [20..27] boundary values u,v for xMinus, xPlus,
 * yMinus, yPlus, [28] mid_h, [29] number of bodies */
int cuibm_case_scalars(cuibm_solver *s, double *out, int n);

/* ---- phase entry points (parity tests) ------------------------------------------- */

/* run one named phase of the step on the solver's stream and synchronise. Names are
 * the reference methods: generateRHS1, solveIntermediateVelocity, generateRHS2,
 * solvePoisson, velocityProjection, generateLHS1, generateLHS2; for the IBM solvers
 * also pre_step, intermediate_velocity, pressure_phase, project_velocity,
 * update_body, post_step, tagPoints, calculateForce, robin and (LUO) weightUhat,
 * weightPressure, rhs1GNInterpolation, rhs1HNInterpolation, preRHS2Interpolation,
 * zeroVelocity. */
int cuibm_phase(cuibm_solver *s, const char *name);

/* ---- field access in reference layouts --------------------------------------------- */

/* element count of a field, or a negative code. Double fields: u, uhat, uold, p,
 * N, Nold, L, bc1, rhs1, rhs2, xm, xp, ym, yp, x, dx, y, dy, xu, yv, LHS1_val,
 * LHS2_val, ustar, pstar, body_intercept_{x,y}, image_point_{x,y},
 * body_intercept_p_{x,y}, image_point_p_{x,y}, distance_from_intersection_to_node,
 * distance_between_nodes_at_IB, distance_from_{u,v}_to_body, uv, body_{x,y}.
 * Int fields: ghostTagsUV, hybridTagsUV, hybridTagsUV2, ghostTagsP, hybridTagsP,
 * LHS{1,2}_rowptr, LHS{1,2}_col. */
long cuibm_field_size(cuibm_solver *s, const char *name);
/* copy a field to host memory (at most n elements); returns the field's element count */
long cuibm_get_field(cuibm_solver *s, const char *name, void *host, long n);
/* overwrite a field from host memory (n must equal its size) */
long cuibm_set_field(cuibm_solver *s, const char *name, const void *host, long n);

/* scalars: out[0] timeStep, [1..2] iterationCount1/2, [3..4] final residual norms,
 * [5..6] forceX/forceY, [7..9] FxX/FxY/FxU, [10..13] startI/startJ/numCellsX/Y,
 * [14..15] midX/midY, [16] AMG levels, [17..18] nx/ny, [19..20] solver tolerances,
 * [21..25] ghost/hybrid velocity, ghost/hybrid pressure and solid node counts,
 * running totals since initialise: [26..27] velocity/Poisson iterations,
 * [28..29] velocity/Poisson solve bytes, [30] Poisson solves, [31] BiCGSTAB
 * breakdown restarts; OSC_CYLINDER only, summed over all assemblies of LHS1 then
 * LHS2: [32|36] 5-point slots a ghost row leaves unwritten, [33|37] hybrid tail slots
 * never written, [34|38] tail slots written twice, [35|39] interpolation corners outside
 * the 3x3 stencil (the reference's result depends on stale slots when any is nonzero);
 * immersed-body solvers: [40] passes of the last step (above 1 only for strongly coupled
 * VIV), [41] transverse velocity of the body centre; [42] largest CFL number so far,
 * [43..45] its cell I, J and time step (NavierStokes/CFL.cu:10-35); OSC_CYLINDER with a
 * moving body: [46] Poisson hierarchies set up again on the device (the others, the first
 * included, were built on the host). n may be up to 47. */
int cuibm_info(cuibm_solver *s, double *out, int n);

/* Logger (src/logger.h:94-232): accumulated host wall-clock seconds per phase, with the
 * reference's event names ("initialise", "Intermediate Velocity Setup", "Intermediate
 * Velocity Solve", "Intermediate Velocity Weight", "Poisson Setup", "Poisson Solve",
 * "Poisson Weight", "Velocity Projection", "Calculate Force", "Move Body", "Update Solver",
 * "CFL", "output"), written as "name<TAB>seconds" lines into buf (at most n bytes, NUL
 * terminated); returns the length of the full text. */
long cuibm_timers(cuibm_solver *s, char *buf, long n);

/* algorithmic HBM bytes of the last step: which = 0 whole step, 1 velocity solve,
 * 2 Poisson solve */
double cuibm_bytes(cuibm_solver *s, int which);

/* ---- timing (HIP events on the solver's stream) ----------------------------------- */

/* run nsteps time steps and return their elapsed device time in *ms */
int cuibm_time_steps(cuibm_solver *s, int nsteps, double *ms);

/* time `reps` back-to-back launches of one kernel of the step on its live inputs and
 * return the mean launch time in *ms and the algorithmic bytes per launch in *bytes.
 * kernel: "spmv_lhs2" (the Poisson operator product y = LHS2 x), "spmv_dot_lhs2"
 * (the same product fused with the two BiCGSTAB dot partials, as launched inside the
 * Poisson solve), "spmv_lhs1", "spmv_dot_lhs1" (the velocity operator product with the
 * same two dot partials, as launched inside the velocity solve), "rhs1" (fused
 * intermediate-velocity right-hand side),
 * "vcycle" (one Poisson preconditioner application), "jacobi0" (one weighted-Jacobi
 * sweep on the finest AMG level), "vc_pre" / "vc_post" (the fused finest-level
 * pre-smoothing + residual and prolongation + post-smoothing kernels of the stencil
 * form, as launched inside the V-cycle), "copy" (a copy of 8 Mi doubles,
 * the calibration kernel for hardware byte counters), "copy:N" (a copy of N doubles,
 * 1 <= N <= 2^28: the rate a plain copy reaches at a given launch size). "live:vc_pre" instead runs
 * `reps` time steps and returns the mean duration of the finest-level pre-smoothing
 * launches inside them (HIP events around each launch); the return value is then the
 * number of launches timed. "live:poisson" likewise returns the mean device time of one
 * whole Poisson solve (events recorded before and after each solve on the solver's
 * stream, host waits inside the solve included) over `reps` time steps, with *bytes the
 * mean algorithmic bytes of one solve as run; the return value is the number of solves. */
int cuibm_time_kernel(cuibm_solver *s, const char *kernel, int reps, double *ms, double *bytes);

/* The algorithmic bytes of each kernel launch of one Poisson BiCGSTAB iteration with the
 * AMG V-cycle (what = "poisson_iteration"), in launch order, into out (up to cap values);
 * returns their number (0 without the AMG preconditioner). With the rate a plain copy
 * reaches at each launch size ("copy:N" above) this gives the time the iteration would
 * take if every launch ran at that rate. No reference counterpart. */
long cuibm_launch_bytes(cuibm_solver *s, const char *what, double *out, long cap);

/* block until all work on the current device has finished (hipDeviceSynchronize);
 * the timing harness brackets its timed region with it. No reference counterpart. */
int cuibm_device_synchronize(void);

/* ---- discrete adjoint (no reference counterpart) ----------------------------------- */

/* Per-point body velocities, a build extension: the reference gives every body point the
 * velocity uB[0] (luo_base/kernels/biLinearInterpolation.cu:87,181). Here ub[k], vb[k]
 * (n = number of body points) are interpolated linearly along the body segment that holds
 * each ghost or hybrid node's body intercept, and the solid nodes take the rigid rotation
 * u = -omega (y - yc), v = omega (x - xc) about the body centre. Static immersed bodies
 * (LUO), after cuibm_initialise; the initial field's solid nodes are updated. */
int cuibm_set_body_velocity(cuibm_solver *s, const double *ub, const double *vb, int n, double omega);

/* Discrete adjoint of J = (1 / (n1 - n0 + 1)) sum_{n = n0..n1} 2 Fx^n (mean C_D over
 * steps n0..n1, 1-based) for a run of nsteps steps from a freshly initialised LUO solver
 * with per-point body velocities: the forward run keeps checkpoints every
 * checkpoint_every steps; with backward != 0 each segment is recomputed from its
 * checkpoint and stepped backwards (transposed LHS1 / LHS2 solves, transposed
 * interpolation, transposed stencils). Writes *J; with backward, grad_ub[k] = dJ/dub[k],
 * grad_vb[k] = dJ/dvb[k] (arrays of n) and *grad_omega = dJ/domega; stats (may be null,
 * 8 values): forward seconds, backward seconds, forward Poisson and velocity iterations,
 * adjoint Poisson and velocity iterations, the backward pass's set-up seconds (transposed
 * operators, their hierarchy, the inverse stencil maps) and its seconds recomputing the
 * segments. The solver's state afterwards is not the
 * state after nsteps steps (the backward pass recomputes earlier segments). */
int cuibm_adjoint_run(cuibm_solver *s, int nsteps, int n0, int n1, int checkpoint_every, int backward, double *J,
                      double *grad_ub, double *grad_vb, double *grad_omega, double *stats);

/* ---- ensembles of independent cases ------------------------------------------------ */

/* No counterpart in the reference, which runs one case per process (src/cuIBM.cu:30-82):
 * run n cases to completion concurrently on the current device. Each case is created from
 * case_folders[i] with the common overrides argc/argv (as cuibm_create), initialised and
 * stepped `steps` times (steps <= 0: its own nt) on its own solver and HIP stream.
 * `workers` host threads (1..64) drive the cases; each case goes to the least loaded
 * thread, largest cells x steps first. Outputs (arrays of n, may be null): cells[i] =
 * nx*ny, cd[i] = 2 Fx and cl[i] = 2 Fy after the last step. *wall_s = seconds from the
 * first create to the last step, *steps_wall_s = seconds from the moment every case is
 * initialised to the last step. On failure returns the error code of the lowest-numbered
 * failing case and writes its message to err (err_len bytes, may be null). */
int cuibm_ensemble_run(int n, const char *const *case_folders, int argc, const char *const *argv, int steps,
                       int workers, long long *cells, double *cd, double *cl, double *wall_s, double *steps_wall_s,
                       char *err, int err_len);

/* No counterpart in the reference: the same run in lockstep. Every phase of the time step
 * is one launch over all n cases (fields in concatenated arrays, a block map giving each
 * block its case); the BiCGSTAB solves keep per-case scalars, stopping tests and (for the
 * Poisson solve, as -poissonRestart sets them) reliable-update restarts, so each case stops
 * at its own tolerance. The cases must be LUO with static bodies at rest,
 * stencil operators (the default), a DIAGONAL velocity and a SMOOTHED_AGGREGATION Poisson
 * preconditioner of one AMG precision, and equal tolerances, iteration limits and step
 * counts; otherwise returns CUIBM_ERR_UNSUPPORTED or CUIBM_ERR_ARGUMENT. `workers` host
 * threads (1..64) run the cases' set-up. argv may also hold -lockstepGraphs 0|1: with 1
 * (the default) each solve's blocks of eight BiCGSTAB iterations run as a HIP graph
 * captured at the block's first use and replayed after; with 0 their kernels are launched
 * one by one. Outputs as cuibm_ensemble_run. */
int cuibm_ensemble_lockstep(int n, const char *const *case_folders, int argc, const char *const *argv, int steps,
                            int workers, long long *cells, double *cd, double *cl, double *wall_s,
                            double *steps_wall_s, char *err, int err_len);

/* ---- linear algebra (replaces cusp::krylov::bicgstab / cusp::precond) ------------- */

/* upload a CSR matrix (n x n, 0-based, sorted columns) */
int cuibm_csr_create(int n, const int *rowptr, const int *col, const double *val, cuibm_csr **out);
void cuibm_csr_destroy(cuibm_csr *A);
/* y = A x (host vectors) */
int cuibm_csr_spmv(cuibm_csr *A, const double *x, double *y);
/* cusp::krylov::bicgstab(A, x, b, cusp::monitor(b, maxit, rtol), M): x (host) is the
 * warm start and the result; precond = 0 NONE, 1 DIAGONAL, 2 SMOOTHED_AGGREGATION
 * (preconditioner.h:23-29); amg_precision = 64 (FP64 hierarchy), 32 (the V-cycle in FP32:
 * matrix values, level vectors and smoother arithmetic; it reads the FP64 residual and
 * returns an FP64 correction) or 3264 (FP32 matrix values with FP64 vectors and
 * arithmetic). Returns 0 when converged,
 * CUIBM_ERR_NOT_CONVERGED otherwise, with *iters and *resid set either way. */
int cuibm_bicgstab(cuibm_csr *A, double *x, const double *b, int precond, int amg_precision, double rtol, int maxit,
                   int *iters, double *resid);
/* No counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same solve as cuibm_bicgstab with precond 0 (NONE) or 1 (DIAGONAL), whose device
 * buffers (the staged x and b, the Jacobi diagonal, the eight BiCGSTAB vectors, the
 * reduction partials and the scalar block) all live in one library-owned device array
 * of 11 * 2^18 + 4112 doubles (23.1 MB), one per GPU, held until the process ends.
 * Allocates nothing (no hipMalloc, hipFree or pinned host memory). n above 262144
 * (2^18) returns CUIBM_ERR_ARGUMENT; precond 2 returns CUIBM_ERR_UNSUPPORTED (an AMG
 * hierarchy is allocated). Every call on one GPU shares that array, so calls that
 * overlap on one GPU race: serialise them. A is made beforehand by cuibm_csr_create,
 * which does allocate. Return values as cuibm_bicgstab. */
int cuibm_bicgstab_noalloc(cuibm_csr *A, double *x, const double *b, int precond, double rtol, int maxit,
                           int *iters, double *resid);
/* smoothed-aggregation hierarchy of A: level sizes and nonzeros; returns the level count */
int cuibm_amg_levels(cuibm_csr *A, int *sizes, int *nnz, int cap);
/* one V-cycle x = M^-1 b from a zero guess (host vectors) */
int cuibm_amg_vcycle(cuibm_csr *A, int amg_precision, const double *b, double *x);

#ifdef __cplusplus
}
#endif

#endif /* CUIBM_AMD_H */
