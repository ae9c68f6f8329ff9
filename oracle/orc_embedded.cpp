// ORACLE — test infrastructure only (see orc.hpp).
// The Luo embedded method (SolverType OSC_CYLINDER, class luo_iter), restating
// (paths under src/solvers/NavierStokes/):
//   luo_iter.cu:37-47, luo_base.cu:208-285           step order
//   luo_iter/intermediateVelocity.cu:17-198          velocity set-up (alpha, interpolation,
//                                                    count, LHS1, preconditioner, rhs1 update)
//   luo_iter/poisson.cu:15-152                       the same for the Poisson system
//   luo_iter/kernels/weight.cu:11-125                alpha_u, alpha_v, alpha_p
//   luo_iter/kernels/LHS1.cu:12-408                  LHS1_mid_iter_X/Y
//   luo_iter/kernels/LHS2.cu:12-224                  LHS2_mid_iter
//   luo_iter/kernels/intermediateVelocity.cu:18-73   size_LHS1, update_rhs1_x/y
//   luo_iter/kernels/pressure.cu:18-46               size_LHS2, update_rhs2
//
// The interpolation is folded into the matrices. A hybrid node's row blends the 5-point
// row, divided by its centre coefficient, with the bilinear interpolation at the node:
// corners on the 5-point stencil merge into it, a corner on a diagonal takes the row's
// one extra ("tail") slot, and the node's own corner (the body intercept) moves to the
// right-hand side. A ghost node's row is its interpolation: implicit in the fluid
// corners, explicit (current values) in the ghost corners. Every other interior
// velocity row is divided by its diagonal; interior pressure rows stay unscaled.
//
// Slots: the reference writes COO entries at fixed positions (5 per row, one tail slot
// per hybrid node at 5-point-nnz + count - 1) and sorts them. It does not clear LHS1
// between steps, and it zero-fills only LHS2's values, so a slot that a step leaves
// unwritten keeps an entry from the previous step's sorted arrays. This restatement
// keeps exactly the entries a step writes, including the explicit zeros, and counts
// every slot anomaly in emb_stats so that tests can assert that none occurs.
#include <cmath>
#include "orc.hpp"

namespace orc {

namespace {

// a row under construction, as in orc_ns.cpp: (column, value) pairs sorted later
struct ERow {
    int c[8];
    double v[8];
    int n = 0;
    void add(int col, double val) { c[n] = col; v[n] = val; n++; }
};

inline double sq(double x) { return x * x; }

void csr_from(Csr &A, int n, const std::vector<ERow> &rows) {
    A.n = n;
    A.rp.assign(n + 1, 0);
    for (int i = 0; i < n; i++) A.rp[i + 1] = A.rp[i] + rows[i].n;
    A.ci.resize(A.rp[n]);
    A.v.resize(A.rp[n]);
    for (int i = 0; i < n; i++) {
        const ERow &r = rows[i];
        int order[8];
        for (int k = 0; k < r.n; k++) order[k] = k;
        for (int a = 1; a < r.n; a++)  // insertion sort by column (columns are distinct)
            for (int b = a; b > 0 && r.c[order[b]] < r.c[order[b - 1]]; b--) std::swap(order[b], order[b - 1]);
        for (int k = 0; k < r.n; k++) {
            A.ci[A.rp[i] + k] = r.c[order[k]];
            A.v[A.rp[i] + k] = r.v[order[k]];
        }
    }
}

void copy_row(const Csr &A, int i, ERow &r) {
    for (int k = A.rp[i]; k < A.rp[i + 1]; k++) r.add(A.ci[k], A.v[k]);
}

// alpha of luo_iter/kernels/weight.cu:11-47 for one hybrid node h at (I, J) of a
// component with coordinates xs, ys and row stride st; tags and intercepts of the
// same system (velocity: ghostTagsUV, bi; pressure: ghostTagsP, bip)
double alpha_at(int h, int I, int J, int st, const double *xs, const double *ys, const int *ghost,
                const double *bix, const double *biy) {
    double d1 = 0, d2 = 0;
    const double dxl = xs[I] - xs[I - 1], dyl = ys[J] - ys[J - 1];
    if (ghost[h - 1] > 0)
        d1 = std::sqrt(sq(bix[h - 1] - xs[I - 1]) + sq(biy[h - 1] - ys[J]));
    else if (ghost[h + 1] > 0)
        d1 = std::sqrt(sq(bix[h + 1] - xs[I + 1]) + sq(biy[h + 1] - ys[J]));
    if (ghost[h - st] > 0)
        d2 = std::sqrt(sq(bix[h - st] - xs[I]) + sq(biy[h - st] - ys[J - 1]));
    if (ghost[h + st] > 0)
        d2 = std::sqrt(sq(bix[h + st] - xs[I]) + sq(biy[h + st] - ys[J + 1]));
    return std::sqrt(sq(d1 / dxl) + sq(d2 / dyl));
}

// Row of a hybrid node (LHS1.cu:33-100, LHS2.cu:29-103). cns: N, E, S, W already blended,
// i.e. Cns * (1 - alpha) / Cns_P; cin: alpha * interpolation weights. Returns the tail
// entry count written (for the slot accounting) and adds the self-corner term to rhs.
struct Tail { int writes = 0, dropped = 0; int col = -1; double val = 0; };

Tail hybrid_row(int h, int st, const double cns[4], const double cin[4], const int idx[4], const double q[4],
                ERow &r, double &interp_rhs) {
    //   0  1  2     NW N  NE
    //   3  4  5     W  P  E
    //   6  7  8     SW S  SE
    const int sidx[9] = {h + st - 1, h + st, h + st + 1, h - 1, h, h + 1, h - st - 1, h - st, h - st + 1};
    double sten[9] = {0, cns[0], 0, cns[3], 1, cns[1], 0, cns[2], 0};
    const bool used[9] = {false, true, false, true, true, true, false, true, false};
    for (int n = 0; n < 4; n++)
        for (int m = 0; m < 9; m++)
            if (sidx[m] == idx[n] && m != 4) sten[m] -= cin[n];
    for (int m = 0; m < 9; m++)
        if (used[m]) r.add(sidx[m], sten[m]);
    interp_rhs = 0;
    Tail t;
    for (int n = 0; n < 4; n++) {
        bool inside = false;
        for (int m = 0; m < 9; m++) {
            if (sidx[m] != idx[n]) continue;
            inside = true;
            if (!used[m]) {
                t.writes++;
                t.col = idx[n];
                t.val = -cin[n];
            } else if (idx[n] == h) {
                interp_rhs += cin[n] * q[n];
            }
        }
        if (!inside) t.dropped++;
    }
    if (t.writes > 0) r.add(t.col, t.val);
    return t;
}

// Row of a ghost node (LHS1.cu:101-160, LHS2.cu:104-170). sign = +1 for velocity rows
// (u_g + sum C u = 2 uB - sum C q), -1 for pressure rows (p_g - sum C p = dpdn + sum C q).
// Returns the number of 5-point slots the row leaves unwritten.
int ghost_row(int g, int st, const int *ghost, const double cin[4], const int idx[4], const double q[4],
              double sign, ERow &r, double &temp) {
    bool in[4] = {false, false, false, false};
    const int nsi[5] = {g + st, g + 1, g - st, g - 1, g};  // N E S W P
    bool overlap[5] = {false, false, false, false, true};
    int counter = 0;
    temp = 0;
    for (int l = 0; l < 4; l++) {
        if (ghost[idx[l]] > 0) {
            counter++;
            in[l] = true;
        }
        for (int n = 0; n < 5; n++)
            if (idx[l] == nsi[n]) overlap[n] = true;
    }
    r.add(g, 1.0);
    int written = 1;
    for (int i = 0; i < 4; i++) {
        if (!in[i] && idx[i] != g) {
            r.add(idx[i], sign * cin[i]);
            written++;
        } else {
            temp -= sign * cin[i] * q[i];
        }
    }
    int counter2 = 0;
    for (int i = 0; i < 5; i++) {
        if (counter2 >= counter) break;
        if (!overlap[i]) {
            r.add(nsi[i], 0.0);
            written++;
            counter2++;
        }
    }
    return 5 - written;
}

}  // namespace

// alpha_u, alpha_v (luo_iter/kernels/weight.cu:11-86) over the hybrid nodes of the box
void Solver::alpha_uv() {
    std::fill(alpha1.begin(), alpha1.end(), 0.0);
    for (int J = 1; J < ny - 1; J++)
        for (int I = 1; I < nx - 2; I++) {
            const int h = J * (nx - 1) + I;
            if (hybridTagsUV[h] > 0)
                alpha1[h] = alpha_at(h, I, J, nx - 1, G.xu.data(), G.y.data(), ghostTagsUV.data(), bi_x.data(), bi_y.data());
        }
    for (int J = 1; J < ny - 2; J++)
        for (int I = 1; I < nx - 1; I++) {
            const int h = J * nx + I + numU;
            if (hybridTagsUV[h] > 0)
                alpha1[h] = alpha_at(h, I, J, nx, G.x.data(), G.yv.data(), ghostTagsUV.data(), bi_x.data(), bi_y.data());
        }
}

// alpha_p (luo_iter/kernels/weight.cu:89-125)
void Solver::alpha_pressure() {
    std::fill(alpha2.begin(), alpha2.end(), 0.0);
    for (int J = 1; J < ny - 1; J++)
        for (int I = 1; I < nx - 1; I++) {
            const int h = J * nx + I;
            if (hybridTagsP[h] > 0)
                alpha2[h] = alpha_at(h, I, J, nx, G.x.data(), G.y.data(), ghostTagsP.data(), bip_x.data(), bip_y.data());
        }
}

// LHS1_mid_iter_X/Y + LHS_BC_X/Y (luo_iter/intermediateVelocity.cu:141-176)
void Solver::buildLHS1_embedded() {
    const double *dx = G.dx.data(), *dy = G.dy.data();
    const double dt = P.dt, nu = P.nu;
    buildLHS1();  // boundary rows (LHS_BC_X/Y) are the plain ones
    const Csr plain = LHS1;
    std::vector<ERow> rows(numUV);
    std::fill(ns_rhs1.begin(), ns_rhs1.end(), 0.0);
    std::fill(interp_rhs1.begin(), interp_rhs1.end(), 0.0);
    const double ubody[2] = {B.uB[0], B.vB[0]};
    // size_LHS1 (luo_iter/kernels/intermediateVelocity.cu:18-45): the tail slots are
    // numbered over the box, u nodes first; only their existence matters here
    for (int c = 0; c < 2; c++) {
        const int st = c == 0 ? nx - 1 : nx, off = c == 0 ? 0 : numU;
        const int ni = c == 0 ? nx - 1 : nx, nj = c == 0 ? ny : ny - 1;
        for (int J = 0; J < nj; J++)
            for (int I = 0; I < ni; I++) {
                const int i = J * st + I + off;
                ERow &r = rows[i];
                if (I == 0 || I == ni - 1 || J == 0 || J == nj - 1) {
                    copy_row(plain, i, r);
                    continue;
                }
                const int idx[4] = {qidx[0][i], qidx[1][i], qidx[2][i], qidx[3][i]};
                const double q[4] = {qval[0][i], qval[1][i], qval[2][i], qval[3][i]};
                if (hybridTagsUV[i] > 0) {
                    double cns[5];
                    cns[0] = -dt * nu / (dy[J + 1] * (dy[J] + dy[J + 1]));
                    cns[1] = -dt * nu / (dx[I] * (dx[I] + dx[I + 1]));
                    cns[2] = -dt * nu / (dy[J] * (dy[J] + dy[J + 1]));
                    cns[3] = -dt * nu / (dx[I] * (dx[I] + dx[I - 1]));
                    cns[4] = 1 - cns[0] - cns[1] - cns[2] - cns[3];
                    const double a = alpha1[i];
                    double cin[4];
                    for (int l = 0; l < 4; l++) {
                        cns[l] = cns[l] * (1 - a) / cns[4];
                        cin[l] = qcoef[l][i] * a;
                    }
                    double irhs;
                    const Tail t = hybrid_row(i, st, cns, cin, idx, q, r, irhs);
                    ns_rhs1[i] = (1 - a) / cns[4];
                    interp_rhs1[i] = irhs;
                    if (t.writes == 0) emb_stats[1]++;
                    if (t.writes > 1) emb_stats[2]++;
                    emb_stats[3] += t.dropped;
                } else if (ghostTagsUV[i] > 0) {
                    const double cin[4] = {qcoef[0][i], qcoef[1][i], qcoef[2][i], qcoef[3][i]};
                    double temp;
                    emb_stats[0] += ghost_row(i, st, ghostTagsUV.data(), cin, idx, q, 1.0, r, temp);
                    ns_rhs1[i] = 0;
                    interp_rhs1[i] = 2 * ubody[c] + temp;
                } else {
                    double temp, e, w, n, s;
                    if (c == 0) {
                        temp = 1 + 0.5 * dt * nu * (1 / (dx[I + 1] * (dx[I + 1] + dx[I]) * 0.5)) +
                               0.5 * dt * nu * (1 / (dx[I] * (dx[I + 1] + dx[I]) * 0.5)) +
                               0.5 * dt * nu * (1 / (dy[J] * (dy[J + 1] + dy[J]) * 0.5)) +
                               0.5 * dt * nu * (1 / (dy[J] * (dy[J - 1] + dy[J]) * 0.5));
                        e = -0.5 * dt * nu * (1 / (dx[I + 1] * (dx[I + 1] + dx[I]) * 0.5)) / temp;
                        w = -0.5 * dt * nu * (1 / (dx[I] * (dx[I + 1] + dx[I]) * 0.5)) / temp;
                        n = -0.5 * dt * nu * (1 / (dy[J] * (dy[J + 1] + dy[J]) * 0.5)) / temp;
                        s = -0.5 * dt * nu * (1 / (dy[J] * (dy[J - 1] + dy[J]) * 0.5)) / temp;
                    } else {
                        temp = 1 + 0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I + 1]) * 0.5)) +
                               0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I - 1]) * 0.5)) +
                               0.5 * dt * nu * (1 / (dy[J + 1] * (dy[J] + dy[J + 1]) * 0.5)) +
                               0.5 * dt * nu * (1 / (dy[J] * (dy[J] + dy[J + 1]) * 0.5));
                        e = -0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I + 1]) * 0.5)) / temp;
                        w = -0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I - 1]) * 0.5)) / temp;
                        n = -0.5 * dt * nu * (1 / (dy[J + 1] * (dy[J] + dy[J + 1]) * 0.5)) / temp;
                        s = -0.5 * dt * nu * (1 / (dy[J] * (dy[J] + dy[J + 1]) * 0.5)) / temp;
                    }
                    r.add(i + 1, e);
                    r.add(i - 1, w);
                    r.add(i + st, n);
                    r.add(i - st, s);
                    r.add(i, 1.0);
                    ns_rhs1[i] = 1 / temp;
                    interp_rhs1[i] = 0;
                }
            }
    }
    csr_from(LHS1, numUV, rows);
}

// LHS2_mid_iter + LHS2_BC (luo_iter/poisson.cu:120-136)
void Solver::buildLHS2_embedded() {
    const double *dx = G.dx.data(), *dy = G.dy.data();
    const double dt = P.dt;
    buildLHS2();  // boundary rows (LHS2_BC, diagonal of row 0 doubled)
    const Csr plain = LHS2;
    std::vector<ERow> rows(numP);
    std::fill(ns_rhs2.begin(), ns_rhs2.end(), 0.0);
    std::fill(interp_rhs2.begin(), interp_rhs2.end(), 0.0);
    for (int J = 0; J < ny; J++)
        for (int I = 0; I < nx; I++) {
            const int ip = J * nx + I;
            ERow &r = rows[ip];
            if (I == 0 || I == nx - 1 || J == 0 || J == ny - 1) {
                copy_row(plain, ip, r);
                continue;
            }
            const int idx[4] = {pqidx[0][ip], pqidx[1][ip], pqidx[2][ip], pqidx[3][ip]};
            const double q[4] = {pqval[0][ip], pqval[1][ip], pqval[2][ip], pqval[3][ip]};
            if (hybridTagsP[ip] > 0) {
                double cns[5];
                cns[0] = -dt / (dy[J] * (dy[J] + dy[J + 1]) * 0.5);
                cns[1] = -dt / (dx[I] * (dx[I] + dx[I + 1]) * 0.5);
                cns[2] = -dt / (dy[J] * (dy[J] + dy[J - 1]) * 0.5);
                cns[3] = -dt / (dx[I] * (dx[I] + dx[I - 1]) * 0.5);
                cns[4] = -cns[0] - cns[1] - cns[2] - cns[3];
                const double a = alpha2[ip];
                double cin[4];
                for (int l = 0; l < 4; l++) {
                    cns[l] = (1 - a) * cns[l] / cns[4];
                    cin[l] = a * pqcoef[l][ip];
                }
                double irhs;
                const Tail t = hybrid_row(ip, nx, cns, cin, idx, q, r, irhs);
                ns_rhs2[ip] = (1 - a) / cns[4];
                interp_rhs2[ip] = irhs;
                if (t.writes == 0) emb_stats[5]++;
                if (t.writes > 1) emb_stats[6]++;
                emb_stats[7] += t.dropped;
            } else if (ghostTagsP[ip] > 0) {
                const double cin[4] = {pqcoef[0][ip], pqcoef[1][ip], pqcoef[2][ip], pqcoef[3][ip]};
                double temp;
                emb_stats[4] += ghost_row(ip, nx, ghostTagsP.data(), cin, idx, q, -1.0, r, temp);
                ns_rhs2[ip] = 0;
                interp_rhs2[ip] = dpdn[ip] + temp;
            } else {
                double temp = 0;
                const double e = -dt / (dx[I] * (dx[I] + dx[I + 1]) * 0.5);
                temp += dt / (dx[I] * (dx[I] + dx[I + 1]) * 0.5);
                const double w = -dt / (dx[I] * (dx[I] + dx[I - 1]) * 0.5);
                temp += dt / (dx[I] * (dx[I] + dx[I - 1]) * 0.5);
                const double n = -dt / (dy[J] * (dy[J] + dy[J + 1]) * 0.5);
                temp += dt / (dy[J] * (dy[J] + dy[J + 1]) * 0.5);
                const double s = -dt / (dy[J] * (dy[J] + dy[J - 1]) * 0.5);
                temp += dt / (dy[J] * (dy[J] + dy[J - 1]) * 0.5);
                r.add(ip + 1, e);
                r.add(ip - 1, w);
                r.add(ip + nx, n);
                r.add(ip - nx, s);
                r.add(ip, temp);
                ns_rhs2[ip] = 1;
                interp_rhs2[ip] = 0;
            }
        }
    csr_from(LHS2, numP, rows);
}

// update_rhs1_x / update_rhs1_y (luo_iter/kernels/intermediateVelocity.cu:48-73). The y
// kernel is launched with the u grid (luo_iter/intermediateVelocity.cu:197), so only v
// nodes below 256 * ceil((nx-1) ny / 256) are updated.
void Solver::update_rhs1_embedded() {
    for (int iu = 0; iu < numU; iu++) {
        const int I = iu % (nx - 1), J = iu / (nx - 1);
        if (I == 0 || I == nx - 2 || J == 0 || J == ny - 1) continue;
        rhs1[iu] = rhs1[iu] * ns_rhs1[iu] + interp_rhs1[iu];
    }
    const long threads = 256L * ((long)((numU - 0.5) / 256) + 1);
    for (int ip = 0; ip < numV && ip < threads; ip++) {
        const int I = ip % nx, J = ip / nx, iv = ip + numU;
        if (I == 0 || I == nx - 1 || J == 0 || J == ny - 2) continue;
        rhs1[iv] = rhs1[iv] * ns_rhs1[iv] + interp_rhs1[iv];
    }
}

// update_rhs2 (luo_iter/kernels/pressure.cu:36-46): its guard stops at (nx-1) ny
void Solver::update_rhs2_embedded() {
    for (int ip = 0; ip < numP && ip < (nx - 1) * ny; ip++) {
        const int I = ip % nx, J = ip / nx;
        if (I == 0 || I == nx - 1 || J == 0 || J == ny - 1) continue;
        rhs2[ip] = rhs2[ip] * ns_rhs2[ip] + interp_rhs2[ip];
    }
}

// one pass of luo_base::stepTime for the embedded method (luo_iter.cu:37-47)
void Solver::embedded_pass() {
    // intermediate_velocity_setup (luo_iter/intermediateVelocity.cu:17-77)
    u = uold;
    robin();
    set_inside(u);
    generateN();
    generateL();
    generateBC1();
    generateRHS();
    alpha_uv();
    gn_velocity(u, false);
    hn_velocity(u);
    buildLHS1_embedded();
    PC1.build(LHS1, P.precond1);
    update_rhs1_embedded();
    solve1();
    // poisson_setup (luo_iter/poisson.cu:15-52)
    rhs2_nobody();
    alpha_pressure();
    gn_pressure(false);
    hn_pressure();
    buildLHS2_embedded();
    // interpolation rows (unit diagonal) beside Poisson rows; a moving body rebuilds the
    // hierarchy every step on the product side, which then stops coarsening at 60 unknowns
    PC2.build(LHS2, P.precond2, true, B.moving ? 60 : 0);
    update_rhs2_embedded();
    solve2();
    // luo_base/projectVelocity.cu:11-53
    project_nobody();
    gn_velocity(u, true);
    set_inside(u);
    update_body();
}

void Solver::step_embedded() { step_luo(); }

}  // namespace orc

extern "C" {
// slot accounting of the embedded assembly (orc::Solver::emb_stats)
void orc_embedded_stats(void *h, long *out) {
    auto *S = (orc::Solver *)h;
    for (int i = 0; i < 8; i++) out[i] = S->emb_stats[i];
}
}
