// extern "C" boundary (include/cuibm_amd.h). Exceptions never cross it: every entry
// point catches cuibm::Error / std::exception and returns a negative code.
#include <chrono>
#include <cstring>
#include <mutex>
#include <thread>
#include <sstream>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cuibm_amd.h"
#include "batch.hpp"
#include "noalloc.hpp"
#include "solver.hpp"
#include "yaml_lite.hpp"

using namespace cuibm;

struct cuibm_solver {
    ParameterDB db;
    Domain dom;
    SolverOptions opt;
    std::unique_ptr<NavierStokesSolver> s;
};

struct cuibm_csr {
    DCsr<double> A;
    HCsr h;
    hipStream_t stream = nullptr;
    std::unique_ptr<AmgPrecond> amg[3];  // FP64, FP32 V-cycle, FP32 values
    ~cuibm_csr() {
        for (auto &a : amg) a.reset();
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {
thread_local std::string g_err;

template <class F>
int guard(F &&f) {
    try {
        return f();
    } catch (const Error &e) {
        g_err = e.what();
        return e.code;
    } catch (const YamlError &e) {  // unreadable or malformed case files
        g_err = e.what();
        return CUIBM_ERR_INPUT;
    } catch (const std::exception &e) {
        g_err = e.what();
        return CUIBM_ERR_GENERIC;
    }
}

template <class F>
long guardl(F &&f) {
    try {
        return f();
    } catch (const Error &e) {
        g_err = e.what();
        return e.code;
    } catch (const YamlError &e) {  // unreadable or malformed case files
        g_err = e.what();
        return CUIBM_ERR_INPUT;
    } catch (const std::exception &e) {
        g_err = e.what();
        return CUIBM_ERR_GENERIC;
    }
}

int need(cuibm_solver *s) {
    if (!s || !s->s) throw Error("null solver handle", ERR_ARGUMENT);
    return 0;
}
}  // namespace

extern "C" {

int cuibm_version(void) { return 10000; }

const char *cuibm_last_error(void) { return g_err.c_str(); }

int cuibm_create(const char *case_folder, int argc, const char *const *argv, cuibm_solver **out) {
    return guard([&] {
        if (!out) throw Error("null output handle", ERR_ARGUMENT);
        *out = nullptr;
        std::unique_ptr<cuibm_solver> h(new cuibm_solver);
        readInputs(case_folder ? case_folder : ".", argc, argv, h->db, h->dom);
        for (int i = 0; i + 1 < argc; i++) {
            if (std::strcmp(argv[i], "-outputFolder") == 0) h->opt.outputFolder = argv[i + 1];
            if (std::strcmp(argv[i], "-output") == 0) h->opt.output = std::atoi(argv[i + 1]) != 0;
            if (std::strcmp(argv[i], "-diagnostics") == 0) h->opt.keepDiagnostics = std::atoi(argv[i + 1]) != 0;
        }
        h->s = createSolver(&h->db, &h->dom, h->opt);
        *out = h.release();
        return 0;
    });
}

int cuibm_initialise(cuibm_solver *s) {
    return guard([&] {
        need(s);
        s->s->initialise();
        if (s->opt.output)
            writeInfoFile(s->db, s->dom, s->opt.outputFolder.empty() ? s->db.caseFolder : s->opt.outputFolder);
        return 0;
    });
}

int cuibm_step_time(cuibm_solver *s) {
    return guard([&] {
        need(s);
        s->s->stepTime();
        return 0;
    });
}

int cuibm_write_data(cuibm_solver *s) {
    return guard([&] {
        need(s);
        s->s->writeData();
        return 0;
    });
}

int cuibm_finished(cuibm_solver *s) {
    return guard([&] {
        need(s);
        return s->s->finished() ? 1 : 0;
    });
}

int cuibm_shut_down(cuibm_solver *s) {
    return guard([&] {
        need(s);
        s->s->shutDown();
        return 0;
    });
}

void cuibm_destroy(cuibm_solver *s) {
    try {
        delete s;
    } catch (...) {
    }
}

int cuibm_name(cuibm_solver *s, char *buf, int len) {
    return guard([&] {
        need(s);
        const std::string n = s->s->name();
        if (buf && len > 0) {
            std::strncpy(buf, n.c_str(), len - 1);
            buf[len - 1] = 0;
        }
        return (int)n.size();
    });
}

long cuibm_case_array(cuibm_solver *s, const char *name, double *out, long n) {
    return guardl([&] {
        if (!s) throw Error("null solver handle", ERR_ARGUMENT);
        const std::string k = name ? name : "";
        std::vector<double> v;
        if (k == "x") v = s->dom.x;
        else if (k == "dx") v = s->dom.dx;
        else if (k == "y") v = s->dom.y;
        else if (k == "dy") v = s->dom.dy;
        else if (k == "xu") v = s->dom.xu;
        else if (k == "yv") v = s->dom.yv;
        else if (k == "body_X" || k == "body_Y") {
            for (auto &b : s->db.bodies) {
                const std::vector<double> &src = k == "body_X" ? b.X : b.Y;
                v.insert(v.end(), src.begin(), src.end());
            }
        } else throw Error("unknown case array '" + k + "'", ERR_ARGUMENT);
        if (out) std::memcpy(out, v.data(), sizeof(double) * std::min<long>(n, (long)v.size()));
        return (long)v.size();
    });
}

int cuibm_case_scalars(cuibm_solver *s, double *out, int n) {
    return guard([&] {
        if (!s) throw Error("null solver handle", ERR_ARGUMENT);
        const ParameterDB &d = s->db;
        std::vector<double> v = {(double)s->dom.nx, (double)s->dom.ny, d.nu, d.dt, (double)d.nt, (double)d.nsave,
                                 (double)d.solverType, d.scaleCV, (double)d.startStep, d.velocitySolve.tolerance,
                                 d.PoissonSolve.tolerance, (double)d.velocitySolve.maxIterations,
                                 (double)d.PoissonSolve.maxIterations, (double)d.velocitySolve.preconditioner,
                                 (double)d.PoissonSolve.preconditioner, (double)d.PoissonSolve.amgPrecision,
                                 d.uInitial, d.vInitial, d.uPerturb, d.vPerturb};
        for (int loc = 0; loc < 4; loc++) {
            v.push_back(d.bc[loc][0].value);
            v.push_back(d.bc[loc][1].value);
        }
        v.push_back(s->dom.mid_h);
        v.push_back((double)d.bodies.size());
        std::memcpy(out, v.data(), sizeof(double) * std::min<size_t>(n, v.size()));
        return (int)v.size();
    });
}

int cuibm_phase(cuibm_solver *s, const char *name) {
    return guard([&] {
        need(s);
        if (s->s->phase(name ? name : "") != 0) throw Error(std::string("unknown phase '") + (name ? name : "") + "'", ERR_ARGUMENT);
        return 0;
    });
}

long cuibm_field_size(cuibm_solver *s, const char *name) {
    return guardl([&] {
        need(s);
        const long n = s->s->fieldCount(name ? name : "");
        if (n < 0) throw Error(std::string("unknown field '") + (name ? name : "") + "'", ERR_ARGUMENT);
        return n;
    });
}

long cuibm_get_field(cuibm_solver *s, const char *name, void *host, long n) {
    return guardl([&] {
        need(s);
        const long c = s->s->getField(name ? name : "", host, n);
        if (c < 0) throw Error(std::string("unknown field '") + (name ? name : "") + "'", ERR_ARGUMENT);
        return c;
    });
}

long cuibm_set_field(cuibm_solver *s, const char *name, const void *host, long n) {
    return guardl([&] {
        need(s);
        const long c = s->s->setField(name ? name : "", host, n);
        if (c < 0) throw Error(std::string("unknown or read-only field '") + (name ? name : "") + "'", ERR_ARGUMENT);
        return c;
    });
}

int cuibm_info(cuibm_solver *s, double *out, int n) {
    return guard([&] {
        need(s);
        s->s->info(out, n);
        return 0;
    });
}

long cuibm_timers(cuibm_solver *s, char *buf, long n) {
    return guardl([&]() -> long {
        need(s);
        std::ostringstream os;
        for (const auto &e : s->s->timers().timers()) os << e.first << '\t' << e.second << '\n';
        const std::string t = os.str();
        if (buf && n > 0) {
            const long c = std::min<long>(n - 1, (long)t.size());
            std::memcpy(buf, t.data(), (size_t)c);
            buf[c] = '\0';
        }
        return (long)t.size();
    });
}

double cuibm_bytes(cuibm_solver *s, int which) {
    if (!s || !s->s) return -1;
    switch (which) {
        case 0: return s->s->stepBytes();
        case 1: return s->s->velocityBytes();
        case 2: return s->s->poissonBytes();
    }
    return -1;
}

int cuibm_time_steps(cuibm_solver *s, int nsteps, double *ms) {
    return guard([&] {
        need(s);
        hipStream_t st = s->s->stream();
        hipEvent_t a, b;
        CUIBM_HIP(hipEventCreate(&a));
        CUIBM_HIP(hipEventCreate(&b));
        CUIBM_HIP(hipEventRecord(a, st));
        for (int k = 0; k < nsteps; k++) s->s->stepTime();
        CUIBM_HIP(hipEventRecord(b, st));
        CUIBM_HIP(hipEventSynchronize(b));
        float t = 0;
        CUIBM_HIP(hipEventElapsedTime(&t, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        if (ms) *ms = t;
        return 0;
    });
}

int cuibm_device_synchronize(void) {
    return guard([&] {
        CUIBM_HIP(hipDeviceSynchronize());
        return 0;
    });
}

int cuibm_time_kernel(cuibm_solver *s, const char *kernel, int reps, double *ms, double *bytes) {
    return guard([&] { need(s);
        return s->s->timeKernel(kernel ? kernel : "", reps, ms, bytes); });
}

long cuibm_launch_bytes(cuibm_solver *s, const char *what, double *out, long cap) {
    return guardl([&] { need(s);
        const std::string w = what ? what : "";
        if (w != "poisson_iteration") throw Error("unknown launch list '" + w + "'", ERR_ARGUMENT);
        const std::vector<double> L = s->s->poissonIterationLaunches();
        if (out)
            for (long i = 0; i < std::min<long>(cap, (long)L.size()); i++) out[i] = L[i];
        return (long)L.size();
    });
}

// ---- standalone linear algebra

int cuibm_csr_create(int n, const int *rowptr, const int *col, const double *val, cuibm_csr **out) {
    return guard([&] {
        if (!out || n <= 0 || !rowptr || !col || !val) throw Error("bad CSR arguments", ERR_ARGUMENT);
        if (rowptr[0] != 0) throw Error("CSR row pointer must start at 0", ERR_ARGUMENT);
        for (int i = 0; i < n; i++)
            if (rowptr[i + 1] < rowptr[i]) throw Error("CSR row pointer decreases at row " + std::to_string(i), ERR_ARGUMENT);
        std::unique_ptr<cuibm_csr> h(new cuibm_csr);
        CUIBM_HIP(hipStreamCreate(&h->stream));
        h->h.n = n;
        h->h.ncols = n;
        h->h.rp.assign(rowptr, rowptr + n + 1);
        const int nnz = rowptr[n];
        h->h.ci.assign(col, col + nnz);
        h->h.v.assign(val, val + nnz);
        for (int c : h->h.ci)
            if (c < 0 || c >= n) throw Error("CSR column index out of range", ERR_ARGUMENT);
        upload_csr(h->h, h->A, h->stream);
        *out = h.release();
        return 0;
    });
}

void cuibm_csr_destroy(cuibm_csr *A) { delete A; }

int cuibm_csr_spmv(cuibm_csr *A, const double *x, double *y) {
    return guard([&] {
        if (!A) throw Error("null matrix", ERR_ARGUMENT);
        DArray<double> dx, dy(A->A.n);
        dx.upload(x, A->A.n, A->stream);
        launch_spmv<double, double>(A->A, dx.p(), dy.p(), A->stream);
        dy.download(y, A->A.n, A->stream);
        return 0;
    });
}

static AmgPrecond &amg_of(cuibm_csr *A, int precision) {
    const int k = precision == 32 ? 1 : precision == 3264 ? 2 : 0;
    if (!A->amg[k]) {
        AmgParams p;
        p.precision = precision == 32 || precision == 3264 ? precision : 64;
        A->amg[k].reset(new AmgPrecond(A->h, p, A->stream));
    }
    return *A->amg[k];
}

int cuibm_bicgstab(cuibm_csr *A, double *x, const double *b, int precond, int amg_precision, double rtol, int maxit,
                   int *iters, double *resid) {
    return guard([&] {
        if (!A) throw Error("null matrix", ERR_ARGUMENT);
        const int n = A->A.n;
        DArray<double> dx, db;
        dx.upload(x, n, A->stream);
        db.upload(b, n, A->stream);
        std::unique_ptr<Preconditioner> own;
        Preconditioner *M = nullptr;
        if (precond == NONE) own.reset(new IdentityPrecond(n));
        else if (precond == DIAGONAL) own.reset(new JacobiPrecond(A->A, A->stream));
        else if (precond == SMOOTHED_AGGREGATION) M = &amg_of(A, amg_precision);
        else throw Error("unsupported preconditioner", ERR_UNSUPPORTED);
        if (!M) M = own.get();
        Bicgstab ks(n);
        const SolveStat st = ks.solve(CsrOp(A->A), dx.p(), db.p(), *M, rtol, maxit, A->stream);
        dx.download(x, n, A->stream);
        if (iters) *iters = st.iterations;
        if (resid) *resid = st.residual;
        return st.converged ? 0 : (int)ERR_NOT_CONVERGED;
    });
}

int cuibm_set_body_velocity(cuibm_solver *s, const double *ub, const double *vb, int n, double omega) {
    return guard([&] { need(s);
        auto *l = dynamic_cast<luo_base *>(s->s.get());
        if (!l) throw Error("per-point body velocities need an immersed-body solver", ERR_UNSUPPORTED);
        if (!ub || !vb) throw Error("null body velocity arrays", ERR_ARGUMENT);
        l->setBodyVelocity(ub, vb, n, omega);
        return 0;
    });
}

int cuibm_adjoint_run(cuibm_solver *s, int nsteps, int n0, int n1, int checkpoint_every, int backward, double *J,
                      double *grad_ub, double *grad_vb, double *grad_omega, double *stats) {
    return guard([&] { need(s);
        auto *l = dynamic_cast<luoIBM *>(s->s.get());
        if (!l) throw Error("the discrete adjoint is built for the LUO solver", ERR_UNSUPPORTED);
        l->adjointRun(nsteps, n0, n1, checkpoint_every, backward != 0, J, grad_ub, grad_vb, grad_omega, stats);
        return 0;
    });
}

int cuibm_bicgstab_noalloc(cuibm_csr *A, double *x, const double *b, int precond, double rtol, int maxit, int *iters,
                           double *resid) {
    return guard([&] {
        if (!A || !x || !b) throw Error("null matrix or vector", ERR_ARGUMENT);
        const SolveStat st = bicgstab_noalloc(A->A, x, b, precond, rtol, maxit, A->stream);
        if (iters) *iters = st.iterations;
        if (resid) *resid = st.residual;
        return st.converged ? 0 : (int)ERR_NOT_CONVERGED;
    });
}

int cuibm_amg_levels(cuibm_csr *A, int *sizes, int *nnz, int cap) {
    return guard([&] {
        if (!A) throw Error("null matrix", ERR_ARGUMENT);
        const AmgHierarchyInfo I = amg_of(A, 64).info();
        for (int l = 0; l < I.levels && l < cap; l++) {
            if (sizes) sizes[l] = I.n[l];
            if (nnz) nnz[l] = I.nnz[l];
        }
        return I.levels;
    });
}

int cuibm_amg_vcycle(cuibm_csr *A, int amg_precision, const double *b, double *x) {
    return guard([&] {
        if (!A) throw Error("null matrix", ERR_ARGUMENT);
        const int n = A->A.n;
        AmgPrecond &M = amg_of(A, amg_precision);
        DArray<double> db, dx(n);
        db.upload(b, n, A->stream);
        M.apply(db.p(), dx.p(), A->stream);
        dx.download(x, n, A->stream);
        return 0;
    });
}

}  // extern "C"

extern "C" int cuibm_ensemble_lockstep(int n, const char *const *case_folders, int argc, const char *const *argv,
                                       int steps, int workers, long long *cells, double *cd, double *cl,
                                       double *wall_s, double *steps_wall_s, char *err, int err_len) {
    auto now = [] { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); };
    std::vector<cuibm_solver *> hs;
    auto fail = [&](int rc, const std::string &m) {
        if (err && err_len > 0) {
            const size_t k = std::min((size_t)err_len - 1, m.size());
            std::copy(m.begin(), m.begin() + k, err);
            err[k] = '\0';
        }
        for (auto *h : hs) cuibm_destroy(h);
        return rc;
    };
    if (n <= 0 || !case_folders) return fail(CUIBM_ERR_ARGUMENT, "no cases");
    if (workers < 1 || workers > 64) return fail(CUIBM_ERR_ARGUMENT, "workers must be 1..64");
    // -lockstepGraphs 0|1 is the batch's own switch; the other overrides go to every case
    bool graphs = true;
    std::vector<const char *> args;
    for (int i = 0; i < argc; i++) {
        if (argv[i] && std::string(argv[i]) == "-lockstepGraphs") {
            const std::string v = i + 1 < argc && argv[i + 1] ? argv[i + 1] : "";
            if (v != "0" && v != "1") return fail(CUIBM_ERR_INPUT, "-lockstepGraphs takes 0 or 1");
            graphs = v == "1";
            i++;
            continue;
        }
        args.push_back(argv[i]);
    }
    const double t0 = now();
    int nsteps = -1;
    for (int i = 0; i < n; i++) {
        cuibm_solver *h = nullptr;
        const int rc = cuibm_create(case_folders[i], (int)args.size(), args.data(), &h);
        if (rc != 0) return fail(rc, "case " + std::to_string(i) + ": " + cuibm_last_error());
        hs.push_back(h);
        const int k = steps > 0 ? steps : h->db.nt;
        if (nsteps >= 0 && k != nsteps) return fail(CUIBM_ERR_ARGUMENT, "lockstep cases need equal step counts");
        nsteps = k;
    }
    // set-up of every case on `workers` threads (tags, operators, AMG hierarchies)
    std::mutex mu;
    std::string msg;
    int rc = 0, next = 0;
    std::vector<std::thread> th;
    for (int w = 0; w < std::min(workers, n); w++)
        th.emplace_back([&] {
            for (;;) {
                int i;
                {
                    std::lock_guard<std::mutex> g(mu);
                    if (next >= n || rc != 0) return;
                    i = next++;
                }
                const int r = cuibm_initialise(hs[i]);
                if (r != 0) {
                    std::lock_guard<std::mutex> g(mu);
                    if (rc == 0) {
                        rc = r;
                        msg = "case " + std::to_string(i) + ": " + cuibm_last_error();
                    }
                    return;
                }
            }
        });
    for (auto &t : th) t.join();
    if (rc != 0) return fail(rc, msg);
    double t1 = 0, t2 = 0;
    const int r2 = guard([&] {
        for (auto *h : hs) CUIBM_HIP(hipStreamSynchronize(h->s->stream()));
        std::vector<NavierStokesSolver *> ss;
        for (auto *h : hs) ss.push_back(h->s.get());
        LockstepBatch batch(ss, hs[0]->s->stream(), graphs);
        t1 = now();
        for (int k = 0; k < nsteps; k++) batch.step();
        std::vector<double> vcd, vcl;
        batch.coefficients(vcd, vcl);  // synchronises
        t2 = now();
        for (int i = 0; i < n; i++) {
            if (cells) cells[i] = (long long)hs[i]->s->info_cells();
            if (cd) cd[i] = vcd[i];
            if (cl) cl[i] = vcl[i];
        }
        return 0;
    });
    if (r2 != 0) return fail(r2, cuibm_last_error());
    if (wall_s) *wall_s = t2 - t0;
    if (steps_wall_s) *steps_wall_s = t2 - t1;
    for (auto *h : hs) cuibm_destroy(h);
    return 0;
}
