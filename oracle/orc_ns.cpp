// ORACLE — test infrastructure only (see orc.hpp).
// Plain Navier-Stokes pieces of one time step, restating
//   NavierStokes/kernels/N.cu:29-248            explicit convection N(u)
//   NavierStokes/kernels/L.cu:24-193            explicit diffusion L(u)
//   NavierStokes/kernels/intermediateVelocity.cu:29-143   bc1 and rhs1
//   NavierStokes/kernels/LHS1.cu:24-351         LHS1 = I - 0.5 dt nu Lap (assembled as CSR)
//   NavierStokes/kernels/LHS2.cu:23-173         LHS2 = -dt Lap, Neumann walls, diagonal pins
//   luoIBM/kernels/LHS2.cu:12-62                LHS2 without the centre pin (LUO)
//   NavierStokes/kernels/intermediatePressure.cu:30-71    rhs2 = -div(uhat)
//   NavierStokes/kernels/projectVelocity.cu:25-70         projection
// Paths are relative to src/solvers/NavierStokes/ unless they start with src/.
// Arithmetic expressions keep the reference's operation order.
#include <algorithm>
#include <cmath>
#include "orc.hpp"

namespace orc {

// N.cu:29-128 (Nmidx, Nbcx), N.cu:140-248 (Nmidy, Nbcy)
void Solver::generateN() {
    const double *dx = G.dx.data(), *dy = G.dy.data();
    const double *xm = bc[XMINUS].data(), *xp = bc[XPLUS].data();
    const double *ym = bc[YMINUS].data(), *yp = bc[YPLUS].data();
    const double *uu = u.data();
#pragma omp parallel for schedule(static)
    for (int i = 0; i < numU; i++) {
        const int I = i % (nx - 1), J = i / (nx - 1), iv = (nx - 1) * ny + nx * J + I;
        double t = 0;
        // x part
        const double ue = (I == nx - 2) ? xp[J] : uu[i + 1];
        const double uw = (I == 0) ? xm[J] : uu[i - 1];
        t += uu[i] * (ue * dx[I] / (dx[I + 1] * (dx[I] + dx[I + 1]))
                      + uw * (dx[I] / (dx[I] * (dx[I] + dx[I + 1])) - 1 / dx[I])
                      + uu[i] * (-dx[I] / (dx[I + 1] * (dx[I] + dx[I + 1])) -
                                 (dx[I] / (dx[I] * (dx[I] + dx[I + 1])) - 1 / dx[I])));
        // y part
        if (J == ny - 1) {
            t += (((yp[(nx - 1) + I + 1] - yp[(nx - 1) + I]) * dx[I] / (dx[I] + dx[I + 1]) + yp[(nx - 1) + I]) / 2 +
                  ((uu[iv - nx + 1] - uu[iv - nx]) * dx[I] / (dx[I] + dx[I + 1]) + uu[iv - nx]) / 2) *
                 ((2 * yp[I] - uu[i]) * dy[J] / (dy[J] * (dy[J] + dy[J]))
                  + uu[i - nx + 1] * (dy[J - 1] / (dy[J] * (dy[J] + dy[J - 1])) - 1 / dy[J])
                  + uu[i] * (-dy[J] / (dy[J] * (dy[J] + dy[J])) - (dy[J - 1] / (dy[J] * (dy[J] + dy[J - 1])) - 1 / dy[J])));
        } else if (J == 0) {
            t += (((uu[iv + 1] - uu[iv]) * dx[I] / (dx[I] + dx[I + 1]) + uu[iv]) / 2 +
                  ((ym[(nx - 1) + I + 1] - ym[(nx - 1) + I]) * dx[I] / (dx[I] + dx[I + 1]) + ym[(nx - 1) + I]) / 2) *
                 (uu[i + nx - 1] * dy[J] / (dy[J] * (dy[J + 1] + dy[J]))
                  + (2 * ym[I] - uu[i]) * (dy[J] / (dy[J] * (dy[J] + dy[J])) - 1 / dy[J])
                  + uu[i] * (-dy[J] / (dy[J] * (dy[J + 1] + dy[J])) - (dy[J] / (dy[J] * (dy[J] + dy[J])) - 1 / dy[J])));
        } else {
            t += (((uu[iv + 1] - uu[iv]) * dx[I] / (dx[I] + dx[I + 1]) + uu[iv]) / 2 +
                  ((uu[iv - nx + 1] - uu[iv - nx]) * dx[I] / (dx[I] + dx[I + 1]) + uu[iv - nx]) / 2) *
                 (uu[i + nx - 1] * dy[J] / (dy[J] * (dy[J + 1] + dy[J]))
                  + uu[i - nx + 1] * (dy[J - 1] / (dy[J] * (dy[J] + dy[J - 1])) - 1 / dy[J])
                  + uu[i] * (-dy[J] / (dy[J] * (dy[J + 1] + dy[J])) - (dy[J - 1] / (dy[J] * (dy[J] + dy[J - 1])) - 1 / dy[J])));
        }
        N[i] = t;
    }
#pragma omp parallel for schedule(static)
    for (int ip = 0; ip < numV; ip++) {
        const int I = ip % nx, J = ip / nx, iu = (nx - 1) * J + I, iv = ip + (nx - 1) * ny;
        double t = 0;
        if (I == nx - 1) {
            t += (((xp[J + 1] - xp[J]) * dy[J] / (dy[J] + dy[J + 1]) + xp[J]) / 2 +
                  ((uu[iu + (nx - 1) - 1] - uu[iu - 1]) * dy[J] / (dy[J] + dy[J + 1]) + uu[iu - 1]) / 2) *
                 ((2 * xp[ny + J] - uu[iv]) * dx[I] / (dx[I] * (dx[I] + dx[I]))
                  + uu[iv - 1] * (dx[I - 1] / (dx[I] * (dx[I] + dx[I - 1])) - 1 / dx[I])
                  + uu[iv] * (-dx[I] / (dx[I] * (dx[I] + dx[I])) - (dx[I - 1] / (dx[I] * (dx[I] + dx[I - 1])) - 1 / dx[I])));
        } else if (I == 0) {
            t += (((uu[iu + (nx - 1)] - uu[iu]) * dy[J] / (dy[J] + dy[J + 1]) + uu[iu]) / 2 +
                  ((xm[J + 1] - xm[J]) * dy[J] / (dy[J] + dy[J + 1]) + xm[J]) / 2) *
                 (uu[iv + 1] * dx[I] / (dx[I] * (dx[I] + dx[I + 1]))
                  + (2 * xm[ny + J] - uu[iv]) * (dx[I] / (dx[I] * (dx[I] + dx[I])) - 1 / dx[I])
                  + uu[iv] * (-dx[I] / (dx[I] * (dx[I] + dx[I + 1])) - (dx[I] / (dx[I] * (dx[I] + dx[I])) - 1 / dx[I])));
        } else {
            t += (((uu[iu + (nx - 1)] - uu[iu]) * dy[J] / (dy[J] + dy[J + 1]) + uu[iu]) / 2 +
                  ((uu[iu + (nx - 1) - 1] - uu[iu - 1]) * dy[J] / (dy[J] + dy[J + 1]) + uu[iu - 1]) / 2) *
                 (uu[iv + 1] * dx[I] / (dx[I] * (dx[I] + dx[I + 1]))
                  + uu[iv - 1] * (dx[I - 1] / (dx[I] * (dx[I] + dx[I - 1])) - 1 / dx[I])
                  + uu[iv] * (-dx[I] / (dx[I] * (dx[I] + dx[I + 1])) - (dx[I - 1] / (dx[I] * (dx[I] + dx[I - 1])) - 1 / dx[I])));
        }
        const double vn = (J == ny - 2) ? yp[(nx - 1) + I] : uu[iv + nx];
        const double vs = (J == 0) ? ym[(nx - 1) + I] : uu[iv - nx];
        t += uu[iv] * (vn * dy[J] / (dy[J + 1] * (dy[J + 1] + dy[J]))
                       + vs * (dy[J] / (dy[J] * (dy[J + 1] + dy[J])) - 1 / dy[J])
                       + uu[iv] * (-dy[J] / (dy[J + 1] * (dy[J + 1] + dy[J])) - (dy[J] / (dy[J] * (dy[J + 1] + dy[J])) - 1 / dy[J])));
        N[iv] = t;
    }
}

// L.cu:24-193 (interior and wall rows give the same expressions where both apply)
void Solver::generateL() {
    const double *dx = G.dx.data(), *dy = G.dy.data();
    const double *xm = bc[XMINUS].data(), *xp = bc[XPLUS].data();
    const double *ym = bc[YMINUS].data(), *yp = bc[YPLUS].data();
    const double *uu = u.data();
    const double nu = P.nu;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < numU; i++) {
        const int I = i % (nx - 1), J = i / (nx - 1);
        double t = 0;
        if (I != nx - 2) t += nu * (uu[i + 1] - uu[i]) / (dx[I + 1] * (dx[I + 1] + dx[I]) * 0.5);
        else t += nu * (xp[J] - uu[i]) / (dx[I + 1] * (dx[I + 1] + dx[I]) * 0.5);
        if (I != 0) t += nu * (uu[i - 1] - uu[i]) / (dx[I] * (dx[I + 1] + dx[I]) * 0.5);
        else t += nu * (xm[J] - uu[i]) / (dx[I] * (dx[I + 1] + dx[I]) * 0.5);
        if (J != ny - 1) t += nu * (uu[i + (nx - 1)] - uu[i]) / (dy[J] * (dy[J + 1] + dy[J]) * 0.5);
        else t += nu * (2 * yp[I] - 2 * uu[i]) / (dy[J] * dy[J]);
        if (J != 0) t += nu * (uu[i - (nx - 1)] - uu[i]) / (dy[J] * (dy[J - 1] + dy[J]) * 0.5);
        else t += nu * (2 * ym[I] - 2 * uu[i]) / (dy[J] * dy[J]);
        L[i] = t;
    }
#pragma omp parallel for schedule(static)
    for (int ip = 0; ip < numV; ip++) {
        const int I = ip % nx, J = ip / nx, iv = ip + (nx - 1) * ny;
        double t = 0;
        if (I != nx - 1) t += nu * (uu[iv + 1] - uu[iv]) / (dx[I] * (dx[I] + dx[I + 1]) * 0.5);
        else t += nu * (2 * xp[ny + J] - 2 * uu[iv]) / (dx[I] * dx[I]);
        if (I != 0) t += nu * (uu[iv - 1] - uu[iv]) / (dx[I] * (dx[I] + dx[I - 1]) * 0.5);
        else t += nu * (2 * xm[ny + J] - 2 * uu[iv]) / (dx[I] * dx[I]);
        if (J != ny - 2) t += nu * (uu[iv + nx] - uu[iv]) / (dy[J + 1] * (dy[J] + dy[J + 1]) * 0.5);
        else t += nu * (yp[nx - 1 + I] - uu[iv]) / (dy[J + 1] * (dy[J] + dy[J + 1]) * 0.5);
        if (J != 0) t += nu * (uu[iv - nx] - uu[iv]) / (dy[J] * (dy[J] + dy[J + 1]) * 0.5);
        else t += nu * (ym[nx - 1 + I] - uu[iv]) / (dy[J] * (dy[J] + dy[J + 1]) * 0.5);
        L[iv] = t;
    }
}

// kernels/intermediateVelocity.cu:53-143 (bc1X, bc1Y)
void Solver::generateBC1() {
    const double *dx = G.dx.data(), *dy = G.dy.data();
    const double *xm = bc[XMINUS].data(), *xp = bc[XPLUS].data();
    const double *ym = bc[YMINUS].data(), *yp = bc[YPLUS].data();
    const double nu = P.nu, dt = P.dt;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < numU; i++) {
        const int I = i % (nx - 1), J = i / (nx - 1);
        double t = 0;
        if (I == nx - 2) t += xp[J] * 0.5 * dt * nu * (1 / (dx[I + 1] * (dx[I + 1] + dx[I]) * 0.5));
        if (I == 0) t += xm[J] * 0.5 * dt * nu * (1 / (dx[I] * (dx[I + 1] + dx[I]) * 0.5));
        if (J == ny - 1) t += (2 * yp[I] - u[i]) * nu * dt * 0.5 / dy[J] / dy[J];
        if (J == 0) t += (2 * ym[I] - u[i]) * nu * dt * 0.5 / dy[J] / dy[J];
        bc1[i] = t;
    }
#pragma omp parallel for schedule(static)
    for (int ip = 0; ip < numV; ip++) {
        const int I = ip % nx, J = ip / nx, iv = ip + (nx - 1) * ny;
        double t = 0;
        if (I == nx - 1) t += (2 * xp[ny + J] - u[iv]) * nu * dt * 0.5 / dx[I] / dx[I];
        if (I == 0) t += (2 * xm[ny + J] - u[iv]) * nu * dt * 0.5 / dx[I] / dx[I];
        if (J == ny - 2) t += yp[(nx - 1) + I] * 0.5 * dt * nu * (1 / (dy[J + 1] * (dy[J] + dy[J + 1]) * 0.5));
        if (J == 0) t += ym[(nx - 1) + I] * 0.5 * dt * nu * (1 / (dy[J] * (dy[J] + dy[J + 1]) * 0.5));
        bc1[iv] = t;
    }
}

// kernels/intermediateVelocity.cu:29-35
void Solver::generateRHS() {
    const double dt = P.dt;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < numUV; i++)
        rhs1[i] = u[i] + dt * (0.5 * Nold[i] - 1.5 * N[i] + 0.5 * L[i]) + bc1[i];
}

// a row under construction: (column, value) pairs, later sorted by column
struct RowBuf {
    int c[8];
    double v[8];
    int n = 0;
    void add(int col, double val) { c[n] = col; v[n] = val; n++; }
};

static void csr_from_rows(Csr &A, int n, const std::vector<RowBuf> &rows) {
    A.n = n;
    A.rp.assign(n + 1, 0);
    for (int i = 0; i < n; i++) A.rp[i + 1] = A.rp[i] + rows[i].n;
    A.ci.resize(A.rp[n]);
    A.v.resize(A.rp[n]);
    for (int i = 0; i < n; i++) {
        const RowBuf &r = rows[i];
        int order[8];
        for (int k = 0; k < r.n; k++) order[k] = k;
        std::sort(order, order + r.n, [&](int a, int b) { return r.c[a] < r.c[b]; });
        for (int k = 0; k < r.n; k++) {
            A.ci[A.rp[i] + k] = r.c[order[k]];
            A.v[A.rp[i] + k] = r.v[order[k]];
        }
    }
}

// LHS1.cu:24-351. Fadlun rows (FadlunModified/kernels/LHS1.cu:12-484) keep the 5-point
// pattern: the stencil neighbours that the reference fills with explicit zeros stay as
// zero entries, the tagged coupling takes its neighbour's slot.
void Solver::buildLHS1() {
    const double *dx = G.dx.data(), *dy = G.dy.data();
    const double nu = P.nu, dt = P.dt;
    std::vector<RowBuf> rows(numUV);
    const bool fad = (P.solver_type == FADLUN);
    for (int i = 0; i < numU; i++) {
        const int I = i % (nx - 1), J = i / (nx - 1);
        RowBuf &r = rows[i];
        const bool interior = !(I == 0 || I == nx - 2 || J == 0 || J == ny - 1);
        if (fad && interior && !((hybridTagsUV[i] == -1 && ghostTagsUV[i] == -1) || ghostTagsUV[i] == 0)) {
            // tagged Fadlun row: u_i - c u_k = rhs
            int k = -1;
            double c = 0;
            if (hybridTagsUV[i] != -1) {
                const int h2 = hybridTagsUV2[i];
                if (hybridTagsUV[i] == h2 - 1 || hybridTagsUV[i] == h2 + 1 ||
                    hybridTagsUV[i] == h2 + (nx - 1) || hybridTagsUV[i] == h2 - (nx - 1)) {
                    k = h2;
                    c = -dist_int[i] / (dist_int[i] + dist_between[i]);
                }
            } else if (ghostTagsUV[i] > 0) {
                int h = -1;
                if (hybridTagsUV[i + 1] != -1) { h = i + 1; k = i - 1; }
                else if (hybridTagsUV[i - 1] != -1) { h = i - 1; k = i + 1; }
                else if (hybridTagsUV[i + (nx - 1)] != -1) { h = i + (nx - 1); k = i - (nx - 1); }
                else if (hybridTagsUV[i - (nx - 1)] != -1) { h = i - (nx - 1); k = i + (nx - 1); }
                if (h >= 0) c = -(dx[I] - dist_int[h]) / (dx[I] - dist_int[h] + dist_between[h]);
            }
            const int nb[4] = {i + 1, i - 1, i + (nx - 1), i - (nx - 1)};
            for (int q = 0; q < 4; q++) r.add(nb[q], nb[q] == k ? c : 0.0);
            r.add(i, 1.0);
            continue;
        }
        double temp = 1;
        if (I != nx - 2) {
            r.add(i + 1, -0.5 * dt * nu * (1 / (dx[I + 1] * (dx[I + 1] + dx[I]) * 0.5)));
            temp += 0.5 * dt * nu * (1 / (dx[I + 1] * (dx[I + 1] + dx[I]) * 0.5));
        } else temp += 0.5 * dt * nu * (1 / (dx[I + 1] * (dx[I + 1] + dx[I]) * 0.5));
        if (I != 0) {
            r.add(i - 1, -0.5 * dt * nu * (1 / (dx[I] * (dx[I + 1] + dx[I]) * 0.5)));
            temp += 0.5 * dt * nu * (1 / (dx[I] * (dx[I + 1] + dx[I]) * 0.5));
        } else temp += 0.5 * dt * nu * (1 / (dx[I] * (dx[I + 1] + dx[I]) * 0.5));
        if (J != ny - 1) {
            r.add(i + (nx - 1), -0.5 * dt * nu * (1 / (dy[J] * (dy[J + 1] + dy[J]) * 0.5)));
            temp += 0.5 * dt * nu * (1 / (dy[J] * (dy[J + 1] + dy[J]) * 0.5));
        } else temp += 0.5 * dt * nu * (1 / (dy[J] * (dy[J] + dy[J]) * 0.5));
        if (J != 0) {
            r.add(i - (nx - 1), -0.5 * dt * nu * (1 / (dy[J] * (dy[J - 1] + dy[J]) * 0.5)));
            temp += 0.5 * dt * nu * (1 / (dy[J] * (dy[J - 1] + dy[J]) * 0.5));
        } else temp += 0.5 * dt * nu * (1 / (dy[J] * (dy[J] + dy[J]) * 0.5));
        r.add(i, temp);
    }
    for (int ip = 0; ip < numV; ip++) {
        const int I = ip % nx, J = ip / nx, i = ip + numU;
        RowBuf &r = rows[i];
        const bool interior = !(I == 0 || I == nx - 1 || J == 0 || J == ny - 2);
        if (fad && interior && !((hybridTagsUV[i] == -1 && ghostTagsUV[i] == -1) || ghostTagsUV[i] == 0)) {
            int k = -1;
            double c = 0;
            if (hybridTagsUV[i] > 0) {
                const int h2 = hybridTagsUV2[i];
                if (hybridTagsUV[i] == h2 - 1 || hybridTagsUV[i] == h2 + 1 ||
                    hybridTagsUV[i] == h2 + nx || hybridTagsUV[i] == h2 - nx) {
                    k = h2;
                    c = -dist_int[i] / (dist_between[i] + dist_int[i]);
                }
            } else if (ghostTagsUV[i] != -1) {
                int h = -1;
                if (hybridTagsUV[i + 1] != -1) { h = i + 1; k = i - 1; }
                else if (hybridTagsUV[i - 1] != -1) { h = i - 1; k = i + 1; }
                else if (hybridTagsUV[i + nx] != -1) { h = i + nx; k = i - nx; }
                else if (hybridTagsUV[i - nx] != -1) { h = i - nx; k = i + nx; }
                if (h >= 0) c = -(dx[I] - dist_int[h]) / (dx[I] - dist_int[h] + dist_between[h]);
            }
            const int nb[4] = {i + 1, i - 1, i + nx, i - nx};
            for (int q = 0; q < 4; q++) r.add(nb[q], nb[q] == k ? c : 0.0);
            r.add(i, 1.0);
            continue;
        }
        double temp = 1;
        if (I != nx - 1) {
            r.add(i + 1, -0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I + 1]) * 0.5)));
            temp += 0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I + 1]) * 0.5));
        } else temp += 0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I]) * 0.5));
        if (I != 0) {
            r.add(i - 1, -0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I - 1]) * 0.5)));
            temp += 0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I - 1]) * 0.5));
        } else temp += 0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I]) * 0.5));
        if (J != ny - 2) {
            r.add(i + nx, -0.5 * dt * nu * (1 / (dy[J + 1] * (dy[J] + dy[J + 1]) * 0.5)));
            temp += 0.5 * dt * nu * (1 / (dy[J + 1] * (dy[J] + dy[J + 1]) * 0.5));
        } else temp += 0.5 * dt * nu * (1 / (dy[J + 1] * (dy[J] + dy[J + 1]) * 0.5));
        if (J != 0) {
            r.add(i - nx, -0.5 * dt * nu * (1 / (dy[J] * (dy[J] + dy[J + 1]) * 0.5)));
            temp += 0.5 * dt * nu * (1 / (dy[J] * (dy[J] + dy[J + 1]) * 0.5));
        } else temp += 0.5 * dt * nu * (1 / (dy[J] * (dy[J] + dy[J + 1]) * 0.5));
        r.add(i, temp);
    }
    csr_from_rows(LHS1, numUV, rows);
}

// LHS2.cu:23-173; luoIBM/kernels/LHS2.cu:12-62; FadlunModified/kernels/LHS2.cu:12-248
void Solver::buildLHS2() {
    const double *dx = G.dx.data(), *dy = G.dy.data();
    const double dt = P.dt;
    const bool fad = (P.solver_type == FADLUN);
    const bool centre_pin = (P.solver_type == NAVIERSTOKES || P.solver_type == FADLUN);
    const int centre = (ny / 2) * nx + nx / 2;
    std::vector<RowBuf> rows(numP);
    for (int ip = 0; ip < numP; ip++) {
        const int I = ip % nx, J = ip / nx;
        RowBuf &r = rows[ip];
        const bool interior = !(I == 0 || I == nx - 1 || J == 0 || J == ny - 1);
        double temp = 0;
        if (fad && interior && hybridTagsP[ip] != -1) {
            const double du = dist_u_body[ip], dv = dist_v_body[ip];
            const bool cu = du > dx[I] / 2 && du < dx[I], cv = dv > dy[J] / 2 && dv < dy[J];
            if (ghostTagsP[ip + 1] == -1) {
                if (cu) { r.add(ip + 1, -dt / dx[I] / du); temp += dt / dx[I] / du; }
                else { r.add(ip + 1, -dt / dx[I] / dx[I]); temp += dt / dx[I] / dx[I]; }
            } else r.add(ip + 1, 0.0);
            if (ghostTagsP[ip - 1] == -1) {
                if (cu) { r.add(ip - 1, -dt / dx[I] / du); temp += dt / dx[I] / du; }
                else { r.add(ip - 1, -dt / dx[I] / dx[I]); temp += dt / dx[I] / dx[I]; }
            } else r.add(ip - 1, 0.0);
            if (ghostTagsP[ip + nx] == -1) {
                if (cv) { r.add(ip + nx, -dt / dy[J] / dv); temp += dt / dy[J] / dv; }
                else { r.add(ip + nx, -dt / dy[J] / dy[J]); temp += dt / dy[J] / dy[J]; }
            } else r.add(ip + nx, 0.0);
            if (ghostTagsP[ip - nx] == -1) {
                if (cv) { r.add(ip - nx, -dt / dy[J] / dv); temp += dt / dy[J] / dv; }
                else { r.add(ip - nx, -dt / dy[J] / dy[J]); temp += dt / dy[J] / dy[J]; }
            } else r.add(ip - nx, 0.0);
        } else if (fad && interior && ghostTagsP[ip] > 0) {
            const int nb[4] = {ip + 1, ip - 1, ip + nx, ip - nx};
            for (int q = 0; q < 4; q++) {
                const double h = (q < 2) ? dx[I] : dy[J];
                if (ghostTagsP[nb[q]] == 0) { r.add(nb[q], -dt / h / h); temp += dt / h / h; }
                else r.add(nb[q], 0.0);
            }
        } else {
            if (I != nx - 1) {
                r.add(ip + 1, -dt / (dx[I] * (dx[I] + dx[I + 1]) * 0.5));
                temp += dt / (dx[I] * (dx[I] + dx[I + 1]) * 0.5);
            }
            if (I != 0) {
                r.add(ip - 1, -dt / (dx[I] * (dx[I] + dx[I - 1]) * 0.5));
                temp += dt / (dx[I] * (dx[I] + dx[I - 1]) * 0.5);
            }
            if (J != ny - 1) {
                r.add(ip + nx, -dt / (dy[J] * (dy[J] + dy[J + 1]) * 0.5));
                temp += dt / (dy[J] * (dy[J] + dy[J + 1]) * 0.5);
            }
            if (J != 0) {
                r.add(ip - nx, -dt / (dy[J] * (dy[J] + dy[J - 1]) * 0.5));
                temp += dt / (dy[J] * (dy[J] + dy[J - 1]) * 0.5);
            }
        }
        if (ip == 0) temp += temp;                                  // LHS2.cu:104-105
        if (centre_pin && interior && ip == centre) temp += temp;   // LHS2.cu:169-172
        r.add(ip, temp);
    }
    csr_from_rows(LHS2, numP, rows);
}

// NavierStokes/kernels/intermediatePressure.cu:30-71 (== luoIBM/kernels/intermediatePressure.cu:17-58)
void Solver::rhs2_nobody() {
    const double *dx = G.dx.data(), *dy = G.dy.data();
    const double *xm = bc[XMINUS].data(), *xp = bc[XPLUS].data();
    const double *ym = bc[YMINUS].data(), *yp = bc[YPLUS].data();
#pragma omp parallel for schedule(static)
    for (int ip = 0; ip < numP; ip++) {
        const int I = ip % nx, J = ip / nx, iu = (nx - 1) * J + I, iv = (nx - 1) * ny + nx * J + I;
        double t = 0;
        if (I != nx - 1) t -= uhat[iu] / dx[I];
        else t -= xp[J] / dx[I];
        if (I != 0) t += uhat[iu - 1] / dx[I];
        else t += xm[J] / dx[I];
        if (J != ny - 1) t -= uhat[iv] / dy[J];
        else t -= yp[(nx - 1) + I] / dy[J];
        if (J != 0) t += uhat[iv - nx] / dy[J];
        else t += ym[(nx - 1) + I] / dy[J];
        rhs2[ip] = t;
    }
}

// NavierStokes/kernels/projectVelocity.cu:25-70 (== luo_base/kernels/projectVelocity.cu:14-48)
void Solver::project_nobody() {
    const double *dx = G.dx.data(), *dy = G.dy.data();
    const double dt = P.dt;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < numU; i++) {
        const int I = i % (nx - 1), J = i / (nx - 1), ip = J * nx + I;
        uold[i] = u[i];
        u[i] = uhat[i] - dt * (p[ip + 1] - p[ip]) / (0.5 * dx[I + 1] + 0.5 * dx[I]);
    }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < numV; k++) {
        const int I = k % nx, J = k / nx, ip = J * nx + I, i = k + numU;
        uold[i] = u[i];
        u[i] = uhat[i] - dt * (p[ip + nx] - p[ip]) / (0.5 * dy[J + 1] + 0.5 * dy[J]);
    }
}

}  // namespace orc
