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
    own_.alloc(A.n);
    dinv_ = own_.p();
    launch_diag_inv(A, dinv_, s);
}

JacobiPrecond::JacobiPrecond(const DCsr<double> &A, double *dinv, hipStream_t s) : n_(A.n), dinv_(dinv) {
    launch_diag_inv(A, dinv_, s);
}

void JacobiPrecond::refresh(const DCsr<double> &A, hipStream_t s) {
    if (A.n != n_) throw Error("Jacobi refresh: matrix size changed", ERR_ARGUMENT);
    launch_diag_inv(A, dinv_, s);
}

void JacobiPrecond::apply(const double *r, double *z, hipStream_t s) {
    launch_jacobi_apply(n_, dinv_, r, z, s, stop_);
}

Bicgstab::Bicgstab(int n) : n_(n) {
    own_.alloc(workspace_doubles(n));
    double *ws = own_.p();
    const size_t st = stride(n);
    for (double **a : {&r_, &rs_, &p_, &Mp_, &AMp_, &s_, &Ms_, &AMs_}) {
        *a = ws;
        ws += st;
    }
    part_ = ws;
    sc_ = ws + 2 * RED_GRID;
    CUIBM_HIP(hipHostMalloc((void **)&host_sc_, SC_COUNT * sizeof(double), hipHostMallocDefault));
    pinned_ = true;
}

Bicgstab::Bicgstab(int n, double *ws, double *host_sc) : n_(n), host_sc_(host_sc) {
    const size_t st = stride(n);
    for (double **a : {&r_, &rs_, &p_, &Mp_, &AMp_, &s_, &Ms_, &AMs_}) {
        *a = ws;
        ws += st;
    }
    part_ = ws;
    sc_ = ws + 2 * RED_GRID;
}

Bicgstab::~Bicgstab() {
    if (pinned_ && host_sc_) (void)hipHostFree(host_sc_);
}

SolveStat Bicgstab::solve(const LinOp &A, double *x, const double *b, Preconditioner &M, double rtol,
                          int maxit, hipStream_t s) {
    if (A.n() != n_) throw Error("BiCGSTAB workspace size mismatch", ERR_ARGUMENT);
    const int n = n_;
    double *sc = sc_, *part = part_;
    const double *stop = sc + SC_STOP;
    A.set_stop(stop);
    M.set_stop(stop);
    SolveStat st;
    auto fetch = [&]() {  // the scalar block: one small copy, one sync
        CUIBM_HIP(hipMemcpyAsync(host_sc_, sc, SC_D0 * sizeof(double), hipMemcpyDeviceToHost, s));
        CUIBM_HIP(hipStreamSynchronize(s));
    };
    auto start = [&]() {  // r = b - A x, rs = p = r; clears the stop code
        A.init(x, b, r_, p_, rs_, part, s);
        launch_bicg_finish(part, FIN_INIT, sc, s);
        fetch();
    };
    launch_fill(sc + SC_ITS, 0.0, 1, s);
    start();
    const double tol = rtol * std::sqrt(host_sc_[SC_BB]);
    double rn = std::sqrt(host_sc_[SC_RR]);
    const double rn0 = rn;
    int it = 0, restarts = 0;
    st.tolerance = tol;
    // the first half of iteration k (through the test on ||s||), then the second half.
    // A preconditioner that forms weighted sums of its input gets them from the kernels
    // that produce s and p (for the p of a (re)start, it forms them itself).
    const InputWeights iw = M.input_weights();
    const InputWeights *wp = iw.part ? &iw : nullptr;
    bool p_weighted = false;
    auto half1 = [&]() {
        if (wp && p_weighted) M.apply_weighted(p_, Mp_, s);
        else M.apply(p_, Mp_, s);
        A.spmv_dot(Mp_, AMp_, rs_, false, part, s);
        launch_bicg_finish(part, FIN_ALPHA, sc, s);
        launch_bicg_s(n, r_, AMp_, s_, sc, part, s, wp);
        launch_bicg_finish(part, FIN_SS, sc, s, tol, maxit);
    };
    auto half2 = [&]() {
        if (wp) M.apply_weighted(s_, Ms_, s);
        else M.apply(s_, Ms_, s);
        A.spmv_dot(Ms_, AMs_, s_, true, part, s);
        launch_bicg_finish(part, FIN_OMEGA, sc, s);
        launch_bicg_xr(n, x, Mp_, Ms_, r_, s_, AMs_, rs_, sc, part, s);
        launch_bicg_finish(part, FIN_BETA, sc, s, tol, maxit);
        launch_bicg_p(n, p_, r_, AMp_, sc, s, wp);
        p_weighted = wp != nullptr;
    };
    // The recurrence residual r can drift from b - A x through rounding when a solve runs
    // many iterations with large updates (seen with the FP32 V-cycle at tolerances near
    // 1e-12: after thousands of iterations r claimed 5e-11 of ||b|| while b - A x was
    // larger than b). A solve that converged after kVerifyIts iterations or a restart
    // therefore forms b - A x once; more than kDriftFactor times the tolerance restarts
    // the iteration from it (counted as a restart), at most kMaxChecks times.
    int checks = 0;
    const double reliable = M.restart_reduction();
    double rn_ref = rn;  // the true residual at the last (re)start
    for (;;) {
        while (!(rn <= tol) && it < maxit) {
            // iterations to queue before the next read-back: a fraction of the estimate from
            // the residual reduction so far (or of the last solve), so that few launches
            // fall past the end; one in exact mode
            int batch = 1;
            if (!exact_) {
                double rate = rate_;
                if (it > 0 && rn < rn0) rate = std::pow(rn / rn0, 1.0 / it);
                if (rate > 0 && rate < 1) {
                    const double left = std::log(tol / rn) / std::log(rate);
                    batch = (int)std::max(1.0, std::min(16.0, std::floor(0.7 * left)));
                }
                batch = std::min(batch, maxit - it);
            }
            for (int j = 0; j < batch; j++) {
                half1();
                if (exact_) {  // no launch past a half-step exit or breakdown
                    fetch();
                    if (host_sc_[SC_STOP] == STOP_HALF_PENDING) {  // the half-step update alone
                        launch_bicg_xr(n, x, Mp_, Ms_, r_, s_, AMs_, rs_, sc, part, s);
                        launch_bicg_finish(part, FIN_BETA, sc, s, tol, maxit);
                    }
                    if (host_sc_[SC_STOP] != 0.0) break;
                }
                half2();
            }
            fetch();
            it = (int)host_sc_[SC_ITS];
            const int code = (int)host_sc_[SC_STOP];
            if (code == STOP_HALF) {
                rn = std::sqrt(host_sc_[SC_SS]);
                break;
            }
            rn = std::sqrt(host_sc_[SC_RR]);
            if (code == STOP_BREAK_HALF || code == STOP_BREAK_FULL) {
                if (++restarts > kMaxRestarts) break;
                start();  // x is untouched by a half-step breakdown
                p_weighted = false;
                rn = rn_ref = std::sqrt(host_sc_[SC_RR]);
            } else if (reliable > 0 && code == 0 && rn < reliable * rn_ref) {
                if (++restarts > kMaxRestarts) break;
                start();  // reliable update: the true residual, new search directions
                p_weighted = false;
                rn = rn_ref = std::sqrt(host_sc_[SC_RR]);
            }
        }
        if (!(rn <= tol) || (it < kVerifyIts && restarts == 0) || checks >= kMaxChecks) break;
        checks++;
        start();
        p_weighted = false;
        const double tr = std::sqrt(host_sc_[SC_RR]);
        if (tr <= kDriftFactor * tol) break;
        rn = rn_ref = tr;
        if (++restarts > kMaxRestarts) break;
    }
    if (it > 0 && rn > 0 && rn0 > 0 && rn < rn0) rate_ = std::pow(rn / rn0, 1.0 / it);
    A.set_stop(nullptr);
    M.set_stop(nullptr);
    st.restarts = restarts;
    st.iterations = it;
    st.residual = rn;
    st.converged = rn <= tol;
    return st;
}

}  // namespace cuibm
