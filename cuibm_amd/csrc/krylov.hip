// Preconditioned BiCGSTAB replacing cusp::krylov::bicgstab + cusp::monitor
// (call sites NavierStokesSolver.cu:227-228 and :256-257 under src/solvers/NavierStokes/).
//
// Right-preconditioned BiCGSTAB (van der Vorst 1992) with the monitor rule
// ||r||_2 <= rtol ||b||_2 checked on r before every iteration and on s half-way
// through it; the iteration count counts full iterations. The recurrence scalars
// (rho, alpha, omega, beta) never leave the GPU: reductions finish in a one-block
// kernel that updates them in place, and the host reads back the scalar block twice
// per iteration for the stopping test.
//
// Breakdown (alpha or ||s|| not finite, omega = 0, rho = 0, beta not finite) restarts
// the recurrence from the true residual b - A x of the current iterate, keeping the
// iteration count; x never takes a non-finite update (omega falls back to 0). This
// never triggers in a run whose recurrence stays finite, so such runs follow the
// textbook iteration exactly; it matters for preconditioners that are not exactly
// linear (FP32 storage), where a breakdown would otherwise spin to maxit on NaN.
#include <cmath>

#include "krylov.hpp"

namespace cuibm {

JacobiPrecond::JacobiPrecond(const DCsr<double> &A, hipStream_t s) : n_(A.n) {
    dinv_.alloc(A.n);
    launch_diag_inv(A, dinv_.p(), s);
}

void JacobiPrecond::apply(const double *r, double *z, hipStream_t s) { launch_jacobi_apply(n_, dinv_.p(), r, z, s); }

Bicgstab::Bicgstab(int n) : n_(n) {
    for (auto *a : {&r_, &rs_, &p_, &Mp_, &AMp_, &s_, &Ms_, &AMs_}) a->alloc(n);
    part_.alloc(2 * RED_GRID);
    sc_.alloc(SC_COUNT);
    CUIBM_HIP(hipHostMalloc((void **)&host_sc_, SC_COUNT * sizeof(double), hipHostMallocDefault));
}

Bicgstab::~Bicgstab() {
    if (host_sc_) (void)hipHostFree(host_sc_);
}

SolveStat Bicgstab::solve(const LinOp &A, double *x, const double *b, Preconditioner &M, double rtol,
                          int maxit, hipStream_t s) {
    if (A.n() != n_) throw Error("BiCGSTAB workspace size mismatch", ERR_ARGUMENT);
    const int n = n_;
    double *sc = sc_.p(), *part = part_.p();
    auto read = [&](int slot) {
        CUIBM_HIP(hipMemcpyAsync(host_sc_ + slot, sc + slot, sizeof(double), hipMemcpyDeviceToHost, s));
        CUIBM_HIP(hipStreamSynchronize(s));
        return host_sc_[slot];
    };
    SolveStat st;
    auto fetch = [&]() {  // the whole scalar block: one small copy, one sync
        CUIBM_HIP(hipMemcpyAsync(host_sc_, sc, SC_D0 * sizeof(double), hipMemcpyDeviceToHost, s));
        CUIBM_HIP(hipStreamSynchronize(s));
    };
    auto start = [&]() {  // r = b - A x, rs = p = r
        A.init(x, b, r_.p(), p_.p(), rs_.p(), part, s);
        launch_bicg_finish(part, FIN_INIT, sc, s);
        fetch();
    };
    start();
    const double tol = rtol * std::sqrt(host_sc_[SC_BB]);
    double rn = std::sqrt(host_sc_[SC_RR]);
    int it = 0, restarts = 0;
    st.tolerance = tol;
    while (!(rn <= tol) && it < maxit) {
        M.apply(p_.p(), Mp_.p(), s);
        A.spmv_dot(Mp_.p(), AMp_.p(), rs_.p(), false, part, s);
        launch_bicg_finish(part, FIN_ALPHA, sc, s);
        launch_bicg_s(n, r_.p(), AMp_.p(), s_.p(), sc, part, s);
        launch_bicg_finish(part, FIN_SS, sc, s);
        fetch();
        const double ss = host_sc_[SC_SS];
        if (!std::isfinite(ss) || !std::isfinite(host_sc_[SC_ALPHA])) {
            if (++restarts > kMaxRestarts) break;
            start();  // x is untouched in this half-iteration
            rn = std::sqrt(host_sc_[SC_RR]);
            continue;
        }
        rn = std::sqrt(ss);
        if (rn <= tol) {
            launch_bicg_xhalf(n, x, Mp_.p(), sc, s);
            break;
        }
        M.apply(s_.p(), Ms_.p(), s);
        A.spmv_dot(Ms_.p(), AMs_.p(), s_.p(), true, part, s);
        launch_bicg_finish(part, FIN_OMEGA, sc, s);
        launch_bicg_xr(n, x, Mp_.p(), Ms_.p(), r_.p(), s_.p(), AMs_.p(), rs_.p(), sc, part, s);
        launch_bicg_finish(part, FIN_BETA, sc, s);
        launch_bicg_p(n, p_.p(), r_.p(), AMp_.p(), sc, s);
        it++;
        fetch();
        rn = std::sqrt(host_sc_[SC_RR]);
        if (!(rn <= tol) && (host_sc_[SC_OMEGA] == 0.0 || host_sc_[SC_RHO] == 0.0 ||
                             !std::isfinite(host_sc_[SC_BETA]) || !std::isfinite(rn))) {
            if (++restarts > kMaxRestarts) break;
            start();
            rn = std::sqrt(host_sc_[SC_RR]);
        }
    }
    st.restarts = restarts;
    st.iterations = it;
    st.residual = rn;
    st.converged = rn <= tol;
    return st;
}

}  // namespace cuibm
