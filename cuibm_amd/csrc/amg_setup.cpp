// Host set-up of the smoothed-aggregation hierarchy that replaces
// cusp::precond::aggregation::smoothed_aggregation (src/preconditioner.h:101-104,
// 142-145). CUSP 0.5.1's internal defaults are not available offline, so the method
// is fixed here (Vanek, Mandel & Brezina 1996):
//   strong connections: every nonzero off-diagonal entry of a row;
//   aggregation: greedy three-phase standard aggregation in index order
//     (1) a node whose neighbours are all free forms an aggregate with them,
//     (2) a free node joins the aggregate of its first phase-1 neighbour,
//     (3) a node still free forms an aggregate with its free neighbours;
//   tentative prolongator: piecewise constant, each column scaled to unit norm;
//   smoothed prolongator P = (I - 4/(3 rho) D^-1 A) T, rho = max_i sum_j |a_ij/a_ii|;
//   restriction R = P^T, coarse operator A_c = R A P; with scale0 the finest level
//     restricts with R = P^T D^-1 instead (the hierarchy of the row-scaled D^-1 A, whose
//     smoothing and prolongator are those of A): a matrix that mixes unit-diagonal
//     interpolation rows with Poisson rows of diagonal 4 dt/h^2 otherwise gets coarse
//     diagonals of either sign where the two kinds of row share an aggregate;
//   coarsening stops at <= coarse_max unknowns (2500), solved with a dense inverse.
// Everything is sequential and in index order, so the hierarchy is deterministic.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <utility>

#include "krylov.hpp"

namespace cuibm {

namespace {

std::vector<double> diag_of(const HCsr &A) {
    std::vector<double> d(A.n, 0.0);
    for (int i = 0; i < A.n; i++)
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++)
            if (A.ci[k] == i) d[i] += A.v[k];
    return d;
}

HCsr transpose(const HCsr &A) {
    HCsr T;
    T.n = A.ncols;
    T.ncols = A.n;
    T.rp.assign(T.n + 1, 0);
    for (int c : A.ci) T.rp[c + 1]++;
    for (int i = 0; i < T.n; i++) T.rp[i + 1] += T.rp[i];
    T.ci.resize(A.ci.size());
    T.v.resize(A.ci.size());
    std::vector<int> next(T.rp.begin(), T.rp.end() - 1);
    for (int i = 0; i < A.n; i++)
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++) {
            const int p = next[A.ci[k]]++;
            T.ci[p] = i;
            T.v[p] = A.v[k];
        }
    return T;
}

// C = A B with sorted columns; sums in the order (row of A) x (row of B)
HCsr multiply(const HCsr &A, const HCsr &B) {
    HCsr C;
    C.n = A.n;
    C.ncols = B.ncols;
    C.rp.assign(A.n + 1, 0);
    std::vector<double> acc(B.ncols, 0.0);
    std::vector<int> seen(B.ncols, -1), cols;
    for (int i = 0; i < A.n; i++) {
        cols.clear();
        for (int ka = A.rp[i]; ka < A.rp[i + 1]; ka++) {
            const int k = A.ci[ka];
            const double a = A.v[ka];
            for (int kb = B.rp[k]; kb < B.rp[k + 1]; kb++) {
                const int j = B.ci[kb];
                if (seen[j] != i) {
                    seen[j] = i;
                    acc[j] = 0.0;
                    cols.push_back(j);
                }
                acc[j] += a * B.v[kb];
            }
        }
        std::sort(cols.begin(), cols.end());
        for (int j : cols) {
            C.ci.push_back(j);
            C.v.push_back(acc[j]);
        }
        C.rp[i + 1] = (int)C.ci.size();
    }
    return C;
}

int aggregate(const HCsr &A, std::vector<int> &agg) {
    const int n = A.n;
    agg.assign(n, -1);
    auto strong = [&](int i, int k) { return A.ci[k] != i && A.v[k] != 0.0; };
    int count = 0;
    for (int i = 0; i < n; i++) {  // phase 1
        if (agg[i] != -1) continue;
        bool free_nbhd = true;
        for (int k = A.rp[i]; k < A.rp[i + 1] && free_nbhd; k++)
            if (strong(i, k) && agg[A.ci[k]] != -1) free_nbhd = false;
        if (!free_nbhd) continue;
        agg[i] = count;
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++)
            if (strong(i, k)) agg[A.ci[k]] = count;
        count++;
    }
    const std::vector<int> first(agg);
    for (int i = 0; i < n; i++) {  // phase 2
        if (agg[i] != -1) continue;
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++)
            if (strong(i, k) && first[A.ci[k]] != -1) {
                agg[i] = first[A.ci[k]];
                break;
            }
    }
    for (int i = 0; i < n; i++) {  // phase 3
        if (agg[i] != -1) continue;
        agg[i] = count;
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++)
            if (strong(i, k) && agg[A.ci[k]] == -1) agg[A.ci[k]] = count;
        count++;
    }
    return count;
}

}  // namespace

std::vector<HostAmgLevel> amg_build(const HCsr &A0, int coarse_max, int max_levels, bool scale0, bool build_g) {
    std::vector<HostAmgLevel> lv;
    HCsr A = A0;
    A.ncols = A.n;
    while (A.n > coarse_max && (int)lv.size() < max_levels - 1) {
        std::vector<int> agg;
        const int nc = aggregate(A, agg);
        if (nc >= A.n || nc == 0) break;
        std::vector<int> size(nc, 0);
        for (int a : agg) size[a]++;
        HCsr T;
        T.n = A.n;
        T.ncols = nc;
        T.rp.resize(A.n + 1);
        T.ci = agg;
        T.v.resize(A.n);
        for (int i = 0; i <= A.n; i++) T.rp[i] = i;
        for (int i = 0; i < A.n; i++) T.v[i] = 1.0 / std::sqrt((double)size[agg[i]]);
        const std::vector<double> d = diag_of(A);
        double rho = 0;
        for (int i = 0; i < A.n; i++) {
            double s = 0;
            for (int k = A.rp[i]; k < A.rp[i + 1]; k++) s += std::fabs(A.v[k] * (1.0 / d[i]));
            rho = std::max(rho, s);
        }
        const double w = 4.0 / (3.0 * rho);
        const HCsr AT = multiply(A, T);
        HCsr P;
        P.n = A.n;
        P.ncols = nc;
        P.rp.assign(A.n + 1, 0);
        for (int i = 0; i < A.n; i++) {
            const double di = 1.0 / d[i];
            const int tc = T.ci[i];
            const double tv = T.v[i];
            bool placed = false;
            for (int k = AT.rp[i]; k < AT.rp[i + 1]; k++) {
                const int j = AT.ci[k];
                if (!placed && tc < j) {
                    P.ci.push_back(tc);
                    P.v.push_back(tv);
                    placed = true;
                }
                double val = -w * di * AT.v[k];
                if (!placed && tc == j) {
                    val = tv + val;
                    placed = true;
                }
                P.ci.push_back(j);
                P.v.push_back(val);
            }
            if (!placed) {
                P.ci.push_back(tc);
                P.v.push_back(tv);
            }
            P.rp[i + 1] = (int)P.ci.size();
        }
        HostAmgLevel L;
        L.R = transpose(P);
        if (scale0 && lv.empty())  // column j of P^T scaled by 1/a_jj
            for (size_t k = 0; k < L.R.v.size(); k++) L.R.v[k] = L.R.v[k] * (1.0 / d[L.R.ci[k]]);
        const HCsr AP = multiply(A, P);
        HCsr Ac = multiply(L.R, AP);
        if (build_g) {  // G = [R | -R A]
            const HCsr RA = multiply(L.R, A);
            HCsr &G = L.G;
            G.n = L.R.n;
            G.ncols = 2 * A.n;
            G.rp.assign(G.n + 1, 0);
            for (int c = 0; c < G.n; c++) {
                for (int k = L.R.rp[c]; k < L.R.rp[c + 1]; k++) {
                    G.ci.push_back(L.R.ci[k]);
                    G.v.push_back(L.R.v[k]);
                }
                for (int k = RA.rp[c]; k < RA.rp[c + 1]; k++) {
                    G.ci.push_back(A.n + RA.ci[k]);
                    G.v.push_back(-RA.v[k]);
                }
                G.rp[c + 1] = (int)G.ci.size();
            }
        }
        L.A = A;
        L.P = P;
        L.agg = agg;
        lv.push_back(std::move(L));
        A = std::move(Ac);
    }
    HostAmgLevel last;
    last.A = A;
    lv.push_back(std::move(last));
    return lv;
}

void multiply_pattern(const HCsr &A, const HCsr &B, std::vector<int> &rp, std::vector<int> &ci) {
    rp.assign(A.n + 1, 0);
    ci.clear();
    std::vector<int> seen(B.ncols, -1), cols;
    for (int i = 0; i < A.n; i++) {
        cols.clear();
        for (int ka = A.rp[i]; ka < A.rp[i + 1]; ka++) {
            const int k = A.ci[ka];
            for (int kb = B.rp[k]; kb < B.rp[k + 1]; kb++) {
                const int j = B.ci[kb];
                if (seen[j] != i) {
                    seen[j] = i;
                    cols.push_back(j);
                }
            }
        }
        std::sort(cols.begin(), cols.end());
        ci.insert(ci.end(), cols.begin(), cols.end());
        rp[i + 1] = (int)ci.size();
    }
}

HCsr pad_pattern(const HCsr &A, int nx, int margin) {
    const int n = A.n;
    if (nx <= 0 || n % nx != 0 || margin < 0) return A;
    const int ny = n / nx;
    // rows that are not 5-point rows, and the largest offset of their columns
    std::vector<char> band(n, 0);
    int reach = 0;
    for (int i = 0; i < n; i++) {
        const int ix = i % nx, iy = i / nx;
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++) {
            const int dx = std::abs(A.ci[k] % nx - ix), dy = std::abs(A.ci[k] / nx - iy);
            if (dx + dy <= 1) continue;
            band[i] = 1;
            reach = std::max(reach, std::max(dx, dy));
        }
    }
    if (reach == 0 || reach > 4) return A;  // no such rows, or not a local stencil
    // dilate by `margin` cells along x, then along y
    std::vector<char> tmp(n, 0);
    for (int iy = 0; iy < ny; iy++) {
        int last = -(margin + 1);
        for (int ix = 0; ix < nx; ix++) {
            if (band[iy * nx + ix]) last = ix;
            if (ix - last <= margin) tmp[iy * nx + ix] = 1;
        }
        last = nx + margin + 1;
        for (int ix = nx - 1; ix >= 0; ix--) {
            if (band[iy * nx + ix]) last = ix;
            if (last - ix <= margin) tmp[iy * nx + ix] = 1;
        }
    }
    std::fill(band.begin(), band.end(), 0);
    for (int ix = 0; ix < nx; ix++) {
        int last = -(margin + 1);
        for (int iy = 0; iy < ny; iy++) {
            if (tmp[iy * nx + ix]) last = iy;
            if (iy - last <= margin) band[iy * nx + ix] = 1;
        }
        last = ny + margin + 1;
        for (int iy = ny - 1; iy >= 0; iy--) {
            if (tmp[iy * nx + ix]) last = iy;
            if (last - iy <= margin) band[iy * nx + ix] = 1;
        }
    }
    HCsr S;
    S.n = n;
    S.ncols = A.ncols;
    S.rp.assign(n + 1, 0);
    std::vector<std::pair<int, double>> row;
    for (int i = 0; i < n; i++) {
        row.clear();
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++) row.push_back({A.ci[k], A.v[k]});
        if (band[i]) {
            const int ix = i % nx, iy = i / nx;
            for (int y = std::max(0, iy - reach); y <= std::min(ny - 1, iy + reach); y++)
                for (int x = std::max(0, ix - reach); x <= std::min(nx - 1, ix + reach); x++)
                    row.push_back({y * nx + x, 0.0});
        }
        // stable: a column's own entries (in row order) come before the added zero
        std::stable_sort(row.begin(), row.end(),
                         [](const std::pair<int, double> &a, const std::pair<int, double> &b) { return a.first < b.first; });
        for (size_t q = 0; q < row.size(); q++) {
            if (q > 0 && row[q].first == row[q - 1].first) {
                S.v.back() += row[q].second;
                continue;
            }
            S.ci.push_back(row[q].first);
            S.v.push_back(row[q].second);
        }
        S.rp[i + 1] = (int)S.ci.size();
    }
    return S;
}

std::vector<double> dense_inverse(const HCsr &A) {
    const int n = A.n;
    std::vector<double> M((size_t)n * n, 0.0), X((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) {
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++) M[(size_t)i * n + A.ci[k]] += A.v[k];
        X[(size_t)i * n + i] = 1.0;
    }
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int r = c + 1; r < n; r++)
            if (std::fabs(M[(size_t)r * n + c]) > std::fabs(M[(size_t)piv * n + c])) piv = r;
        if (M[(size_t)piv * n + c] == 0.0) throw Error("singular coarsest AMG operator", ERR_GENERIC);
        if (piv != c)
            for (int k = 0; k < n; k++) {
                std::swap(M[(size_t)c * n + k], M[(size_t)piv * n + k]);
                std::swap(X[(size_t)c * n + k], X[(size_t)piv * n + k]);
            }
        // columns left of c are already eliminated in M (zero off the diagonal)
        const double inv = 1.0 / M[(size_t)c * n + c];
        for (int k = c; k < n; k++) M[(size_t)c * n + k] *= inv;
        for (int k = 0; k < n; k++) X[(size_t)c * n + k] *= inv;
        for (int r = 0; r < n; r++) {
            if (r == c) continue;
            const double f = M[(size_t)r * n + c];
            if (f == 0.0) continue;
            double *mr = &M[(size_t)r * n], *xr = &X[(size_t)r * n];
            const double *mc = &M[(size_t)c * n], *xc = &X[(size_t)c * n];
            for (int k = c; k < n; k++) mr[k] -= f * mc[k];
            for (int k = 0; k < n; k++) xr[k] -= f * xc[k];
        }
    }
    return X;
}

}  // namespace cuibm
