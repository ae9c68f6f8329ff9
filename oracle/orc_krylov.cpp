// ORACLE — test infrastructure only (see orc.hpp).
// Linear solvers replacing CUSP 0.5.1 (not vendored; pinned by README.md:6,73 and
// src/Makefile:28). Call sites: NavierStokesSolver.cu:227-228 (velocity) and
// :256-257 (Poisson); preconditioner construction src/preconditioner.h:84-195.
//
// * BiCGSTAB: right-preconditioned BiCGSTAB (van der Vorst 1992) with the
//   cusp::monitor stopping rule ||r||_2 <= rtol * ||b||_2, checked on r before each
//   iteration and on s half-way through it; the iteration count counts full
//   iterations, the warm start is the caller's x.
// * DIAGONAL: Jacobi, z = r / diag(A).
// * SMOOTHED_AGGREGATION: the smoothed-aggregation AMG of Vanek, Mandel & Brezina
//   (1996), fixed as follows (CUSP's internal defaults cannot be read offline):
//     strength: every nonzero off-diagonal entry of the row is a strong connection;
//     aggregation: greedy three-phase standard aggregation in index order;
//     tentative P: piecewise constant, columns scaled to unit 2-norm;
//     smoothed P = (I - 4/(3 rho) D^-1 A) T, rho = max_i sum_j |a_ij / a_ii|;
//     R = P^T, A_c = R A P;  coarsening stops at <= 256 unknowns (dense inverse);
//     V-cycle: 2 pre + 2 post weighted-Jacobi sweeps, omega = 0.8, zero start.
// Dot products use a fixed blocked summation so the result does not depend on
// the number of threads.
#include <algorithm>
#include <cmath>
#include <cstring>
#include "orc.hpp"

namespace orc {

static const int DOT_BLOCK = 4096;

static double dot(const double *a, const double *b, int n) {
    const int nb = (n + DOT_BLOCK - 1) / DOT_BLOCK;
    std::vector<double> part(nb);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nb; k++) {
        double s = 0;
        const int e = std::min(n, (k + 1) * DOT_BLOCK);
        for (int i = k * DOT_BLOCK; i < e; i++) s += a[i] * b[i];
        part[k] = s;
    }
    double s = 0;
    for (int k = 0; k < nb; k++) s += part[k];
    return s;
}

void spmv(const Csr &A, const double *x, double *y) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < A.n; i++) {
        double s = 0;
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++) s += A.v[k] * x[A.ci[k]];
        y[i] = s;
    }
}

static std::vector<double> diag_inv(const Csr &A) {
    std::vector<double> d(A.n, 0.0);
    for (int i = 0; i < A.n; i++) {
        double s = 0;
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++)
            if (A.ci[k] == i) s += A.v[k];
        d[i] = 1.0 / s;
    }
    return d;
}

static Csr transpose(const Csr &A, int ncols) {
    Csr T;
    T.n = ncols;
    T.rp.assign(ncols + 1, 0);
    for (int k = 0; k < A.nnz(); k++) T.rp[A.ci[k] + 1]++;
    for (int i = 0; i < ncols; i++) T.rp[i + 1] += T.rp[i];
    T.ci.resize(A.nnz());
    T.v.resize(A.nnz());
    std::vector<int> pos(T.rp.begin(), T.rp.end() - 1);
    for (int i = 0; i < A.n; i++)
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++) {
            const int c = A.ci[k];
            T.ci[pos[c]] = i;
            T.v[pos[c]] = A.v[k];
            pos[c]++;
        }
    return T;
}

// C = A * B (row-wise, dense accumulator, columns sorted ascending)
static Csr spgemm(const Csr &A, const Csr &B, int bcols) {
    Csr C;
    C.n = A.n;
    C.rp.assign(A.n + 1, 0);
    std::vector<double> acc(bcols, 0.0);
    std::vector<int> mark(bcols, -1), cols;
    for (int i = 0; i < A.n; i++) {
        cols.clear();
        for (int ka = A.rp[i]; ka < A.rp[i + 1]; ka++) {
            const int k = A.ci[ka];
            const double a = A.v[ka];
            for (int kb = B.rp[k]; kb < B.rp[k + 1]; kb++) {
                const int j = B.ci[kb];
                if (mark[j] != i) { mark[j] = i; acc[j] = 0.0; cols.push_back(j); }
                acc[j] += a * B.v[kb];
            }
        }
        std::sort(cols.begin(), cols.end());
        for (int j : cols) { C.ci.push_back(j); C.v.push_back(acc[j]); }
        C.rp[i + 1] = (int)C.ci.size();
    }
    return C;
}

// greedy three-phase standard aggregation over the nonzero off-diagonal graph
static int aggregate(const Csr &A, std::vector<int> &agg) {
    const int n = A.n;
    agg.assign(n, -1);
    int nagg = 0;
    for (int i = 0; i < n; i++) {
        if (agg[i] != -1) continue;
        bool ok = true;
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++) {
            const int j = A.ci[k];
            if (j != i && A.v[k] != 0.0 && agg[j] != -1) { ok = false; break; }
        }
        if (!ok) continue;
        agg[i] = nagg;
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++) {
            const int j = A.ci[k];
            if (j != i && A.v[k] != 0.0) agg[j] = nagg;
        }
        nagg++;
    }
    std::vector<int> agg1(agg);
    for (int i = 0; i < n; i++) {
        if (agg[i] != -1) continue;
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++) {
            const int j = A.ci[k];
            if (j != i && A.v[k] != 0.0 && agg1[j] != -1) { agg[i] = agg1[j]; break; }
        }
    }
    for (int i = 0; i < n; i++) {
        if (agg[i] != -1) continue;
        agg[i] = nagg;
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++) {
            const int j = A.ci[k];
            if (j != i && A.v[k] != 0.0 && agg[j] == -1) agg[j] = nagg;
        }
        nagg++;
    }
    return nagg;
}

static bool dense_inverse(int n, std::vector<double> &M, std::vector<double> &Inv) {
    Inv.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) Inv[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int r = c + 1; r < n; r++)
            if (std::fabs(M[(size_t)r * n + c]) > std::fabs(M[(size_t)piv * n + c])) piv = r;
        if (M[(size_t)piv * n + c] == 0.0) return false;
        if (piv != c)
            for (int k = 0; k < n; k++) {
                std::swap(M[(size_t)c * n + k], M[(size_t)piv * n + k]);
                std::swap(Inv[(size_t)c * n + k], Inv[(size_t)piv * n + k]);
            }
        const double d = 1.0 / M[(size_t)c * n + c];
        for (int k = 0; k < n; k++) { M[(size_t)c * n + k] *= d; Inv[(size_t)c * n + k] *= d; }
        // Gauss-Jordan elimination of column c from every other row; rows are
        // independent, and the columns of M left of c are already zero
#pragma omp parallel for schedule(static)
        for (int r = 0; r < n; r++) {
            if (r == c) continue;
            const double f = M[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int k = c; k < n; k++) M[(size_t)r * n + k] -= f * M[(size_t)c * n + k];
            for (int k = 0; k < n; k++) Inv[(size_t)r * n + k] -= f * Inv[(size_t)c * n + k];
        }
    }
    return true;
}

int amg_default_coarse_max = 2500;

void Amg::setup(const Csr &A0, int coarse_max, int max_levels) {
    if (coarse_max <= 0) coarse_max = amg_default_coarse_max;
    lv.clear();
    Csr A = A0;
    while (A.n > coarse_max && (int)lv.size() < max_levels - 1) {
        AmgLevel L;
        L.dinv = diag_inv(A);
        std::vector<int> agg;
        const int nagg = aggregate(A, agg);
        if (nagg >= A.n || nagg == 0) break;
        std::vector<int> cnt(nagg, 0);
        for (int i = 0; i < A.n; i++) cnt[agg[i]]++;
        Csr T;
        T.n = A.n;
        T.rp.resize(A.n + 1);
        T.ci.resize(A.n);
        T.v.resize(A.n);
        for (int i = 0; i < A.n; i++) {
            T.rp[i] = i;
            T.ci[i] = agg[i];
            T.v[i] = 1.0 / std::sqrt((double)cnt[agg[i]]);
        }
        T.rp[A.n] = A.n;
        double rho = 0;
        for (int i = 0; i < A.n; i++) {
            double s = 0;
            for (int k = A.rp[i]; k < A.rp[i + 1]; k++) s += std::fabs(A.v[k] * L.dinv[i]);
            rho = std::max(rho, s);
        }
        const double w = 4.0 / (3.0 * rho);
        Csr AT = spgemm(A, T, nagg);
        // P = T - w D^-1 A T  (merge rows of T and AT)
        Csr P;
        P.n = A.n;
        P.rp.assign(A.n + 1, 0);
        for (int i = 0; i < A.n; i++) {
            bool tdone = false;
            for (int k = AT.rp[i]; k < AT.rp[i + 1]; k++) {
                const int j = AT.ci[k];
                if (!tdone && T.ci[i] < j) {
                    P.ci.push_back(T.ci[i]);
                    P.v.push_back(T.v[i]);
                    tdone = true;
                }
                double val = -w * L.dinv[i] * AT.v[k];
                if (!tdone && T.ci[i] == j) { val = T.v[i] + val; tdone = true; }
                P.ci.push_back(j);
                P.v.push_back(val);
            }
            if (!tdone) { P.ci.push_back(T.ci[i]); P.v.push_back(T.v[i]); }
            P.rp[i + 1] = (int)P.ci.size();
        }
        L.R = transpose(P, nagg);
        if (scale0 && lv.empty())  // R = P^T D^-1: column j of P^T scaled by 1/a_jj
            for (size_t k = 0; k < L.R.v.size(); k++) L.R.v[k] = L.R.v[k] * L.dinv[L.R.ci[k]];
        Csr AP = spgemm(A, P, nagg);
        Csr Ac = spgemm(L.R, AP, nagg);
        L.A = A;
        L.P = P;
        lv.push_back(L);
        A = Ac;
    }
    nc = A.n;
    std::vector<double> M((size_t)nc * nc, 0.0);
    for (int i = 0; i < nc; i++)
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++) M[(size_t)i * nc + A.ci[k]] += A.v[k];
    dense_inverse(nc, M, coarse_inv);
    AmgLevel Lc;
    Lc.A = A;
    Lc.dinv = diag_inv(A);
    lv.push_back(Lc);
}

static void vcycle(Amg &G, int l, const double *b, double *x) {
    AmgLevel &L = G.lv[l];
    const int n = L.A.n;
    if (l == (int)G.lv.size() - 1) {
        for (int i = 0; i < n; i++) {
            double s = 0;
            for (int j = 0; j < n; j++) s += G.coarse_inv[(size_t)i * n + j] * b[j];
            x[i] = s;
        }
        return;
    }
    std::vector<double> r(n);
    // first pre-smoothing sweep from x = 0
    for (int i = 0; i < n; i++) x[i] = G.omega * L.dinv[i] * b[i];
    for (int s = 1; s < G.npre; s++) {
        spmv(L.A, x, r.data());
        for (int i = 0; i < n; i++) x[i] += G.omega * L.dinv[i] * (b[i] - r[i]);
    }
    spmv(L.A, x, r.data());
    for (int i = 0; i < n; i++) r[i] = b[i] - r[i];
    const int nc = G.lv[l + 1].A.n;
    std::vector<double> bc(nc), xc(nc);
    spmv(L.R, r.data(), bc.data());
    vcycle(G, l + 1, bc.data(), xc.data());
    spmv(L.P, xc.data(), r.data());
    for (int i = 0; i < n; i++) x[i] += r[i];
    for (int s = 0; s < G.npost; s++) {
        spmv(L.A, x, r.data());
        for (int i = 0; i < n; i++) x[i] += G.omega * L.dinv[i] * (b[i] - r[i]);
    }
}

void Amg::apply(const double *b, double *x) { vcycle(*this, 0, b, x); }

void Precond::build(const Csr &A, int t, bool amg_scale0, int amg_coarse_max) {
    type = t;
    if (t == P_DIAGONAL) dinv = diag_inv(A);
    else if (t == P_SMOOTHED_AGGREGATION) {
        amg.scale0 = amg_scale0;
        amg.setup(A, amg_coarse_max);
    }
}

void Precond::apply(const double *r, double *z) {
    const int n = (int)(type == P_DIAGONAL ? dinv.size() : 0);
    if (type == P_DIAGONAL) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++) z[i] = dinv[i] * r[i];
    } else if (type == P_SMOOTHED_AGGREGATION) {
        amg.apply(r, z);
    }
}

SolveStat bicgstab(const Csr &A, double *x, const double *b, Precond &M, double rtol, int maxit) {
    const int n = A.n;
    std::vector<double> r(n), rs(n), p(n), Mp(n), AMp(n), s(n), Ms(n), AMs(n);
    auto prec = [&](const double *in, double *out) {
        if (M.type == P_NONE) std::memcpy(out, in, sizeof(double) * n);
        else M.apply(in, out);
    };
    SolveStat st;
    const double tol = rtol * std::sqrt(dot(b, b, n));
    spmv(A, x, r.data());
    for (int i = 0; i < n; i++) r[i] = b[i] - r[i];
    p = r;
    rs = r;
    double rho = dot(rs.data(), r.data(), n);
    int it = 0;
    double rn = std::sqrt(dot(r.data(), r.data(), n));
    while (!(rn <= tol) && it < maxit) {
        prec(p.data(), Mp.data());
        spmv(A, Mp.data(), AMp.data());
        const double alpha = rho / dot(rs.data(), AMp.data(), n);
        for (int i = 0; i < n; i++) s[i] = r[i] - alpha * AMp[i];
        rn = std::sqrt(dot(s.data(), s.data(), n));
        if (rn <= tol) {
            for (int i = 0; i < n; i++) x[i] = x[i] + alpha * Mp[i];
            break;
        }
        prec(s.data(), Ms.data());
        spmv(A, Ms.data(), AMs.data());
        const double omega = dot(AMs.data(), s.data(), n) / dot(AMs.data(), AMs.data(), n);
        for (int i = 0; i < n; i++) x[i] = x[i] + alpha * Mp[i] + omega * Ms[i];
        for (int i = 0; i < n; i++) r[i] = s[i] - omega * AMs[i];
        const double rho_new = dot(rs.data(), r.data(), n);
        const double beta = (rho_new / rho) * (alpha / omega);
        rho = rho_new;
        for (int i = 0; i < n; i++) p[i] = r[i] + beta * p[i] - beta * omega * AMp[i];
        it++;
        rn = std::sqrt(dot(r.data(), r.data(), n));
    }
    st.iters = it;
    st.resid = rn;
    st.converged = (rn <= tol);
    return st;
}

}  // namespace orc
