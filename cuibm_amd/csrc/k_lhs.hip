// CSR assembly of the two system matrices, one thread per row.
//
// LHS1 = I - 0.5 dt nu Lap on the packed (u, v) unknowns with Dirichlet ghost
// diagonals at the walls (NavierStokes/kernels/LHS1.cu:24-351); LHS2 = -dt Lap on
// pressure with homogeneous Neumann walls, row 0 pinned by a doubled diagonal and,
// for the plain and Fadlun solvers, the centre row too (NavierStokes/kernels/LHS2.cu:23-173,
// luoIBM/kernels/LHS2.cu:12-62). The modified-Fadlun rows
// (FadlunModified/kernels/LHS1.cu:12-484, FadlunModified/kernels/LHS2.cu:12-248)
// keep the 5-point pattern: stencil neighbours the reference fills with explicit
// zeros stay as zero entries. Columns are stored in ascending order
// (S, W, centre, E, N), so the row pointer is the closed-form count of present
// neighbours and no sort is needed.
#include "kernels.hpp"

namespace cuibm {

void lhs1_rowptr(int nx, int ny, std::vector<int> &rp) {
    const int numU = (nx - 1) * ny, numV = nx * (ny - 1);
    rp.assign(numU + numV + 1, 0);
    int k = 0;
    for (int J = 0; J < ny; J++)
        for (int I = 0; I < nx - 1; I++, k++)
            rp[k + 1] = rp[k] + 1 + (J != 0) + (I != 0) + (I != nx - 2) + (J != ny - 1);
    for (int J = 0; J < ny - 1; J++)
        for (int I = 0; I < nx; I++, k++)
            rp[k + 1] = rp[k] + 1 + (J != 0) + (I != 0) + (I != nx - 1) + (J != ny - 2);
}

void lhs2_rowptr(int nx, int ny, std::vector<int> &rp) {
    rp.assign(nx * ny + 1, 0);
    int k = 0;
    for (int J = 0; J < ny; J++)
        for (int I = 0; I < nx; I++, k++)
            rp[k + 1] = rp[k] + 1 + (J != 0) + (I != 0) + (I != nx - 1) + (J != ny - 1);
}

__global__ void __launch_bounds__(BS) k_lhs1(const int *__restrict__ rp, int *__restrict__ ci, double *__restrict__ v,
                                             GridDev g, double nu, double dt, FadlunLhsDev f) {
    const int nx = g.nx, ny = g.ny, numU = (nx - 1) * ny, numUV = numU + nx * (ny - 1);
    const double *dx = g.dx, *dy = g.dy;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < numUV; i += gridDim.x * blockDim.x) {
        const bool isu = i < numU;
        const int st = isu ? nx - 1 : nx;
        const int k = isu ? i : i - numU;
        const int I = k % st, J = k / st;
        const int lastI = isu ? nx - 2 : nx - 1, lastJ = isu ? ny - 1 : ny - 2;
        int pos = rp[i];
        const bool interior = !(I == 0 || I == lastI || J == 0 || J == lastJ);
        if (f.ghostUV && interior &&
            !((f.hybridUV[i] == -1 && f.ghostUV[i] == -1) || f.ghostUV[i] == 0)) {
            // Fadlun tagged row: u_i + c u_k = rhs_i
            int kk = -1;
            double c = 0;
            if (f.hybridUV[i] > 0 || (isu && f.hybridUV[i] != -1)) {
                const int h = f.hybridUV[i], h2 = f.hybridUV2[i];
                if (h == h2 - 1 || h == h2 + 1 || h == h2 + st || h == h2 - st) {
                    kk = h2;
                    c = isu ? -f.dist_int[i] / (f.dist_int[i] + f.dist_between[i])
                            : -f.dist_int[i] / (f.dist_between[i] + f.dist_int[i]);
                }
            } else if (f.ghostUV[i] > 0) {
                int hn = -1;
                if (f.hybridUV[i + 1] != -1) { hn = i + 1; kk = i - 1; }
                else if (f.hybridUV[i - 1] != -1) { hn = i - 1; kk = i + 1; }
                else if (f.hybridUV[i + st] != -1) { hn = i + st; kk = i - st; }
                else if (f.hybridUV[i - st] != -1) { hn = i - st; kk = i + st; }
                if (hn >= 0) c = -(dx[I] - f.dist_int[hn]) / (dx[I] - f.dist_int[hn] + f.dist_between[hn]);
            }
            const int cols[5] = {i - st, i - 1, i, i + 1, i + st};
            for (int q = 0; q < 5; q++) {
                ci[pos + q] = cols[q];
                v[pos + q] = (q == 2) ? 1.0 : (cols[q] == kk ? c : 0.0);
            }
            continue;
        }
        double eS = 0, eW = 0, eE = 0, eN = 0, temp = 1;
        if (isu) {
            const double ce = 0.5 * dt * nu * (1 / (dx[I + 1] * (dx[I + 1] + dx[I]) * 0.5));
            const double cw = 0.5 * dt * nu * (1 / (dx[I] * (dx[I + 1] + dx[I]) * 0.5));
            if (I != nx - 2) { eE = -0.5 * dt * nu * (1 / (dx[I + 1] * (dx[I + 1] + dx[I]) * 0.5)); temp += ce; }
            else temp += ce;
            if (I != 0) { eW = -0.5 * dt * nu * (1 / (dx[I] * (dx[I + 1] + dx[I]) * 0.5)); temp += cw; }
            else temp += cw;
            if (J != ny - 1) {
                eN = -0.5 * dt * nu * (1 / (dy[J] * (dy[J + 1] + dy[J]) * 0.5));
                temp += 0.5 * dt * nu * (1 / (dy[J] * (dy[J + 1] + dy[J]) * 0.5));
            } else temp += 0.5 * dt * nu * (1 / (dy[J] * (dy[J] + dy[J]) * 0.5));
            if (J != 0) {
                eS = -0.5 * dt * nu * (1 / (dy[J] * (dy[J - 1] + dy[J]) * 0.5));
                temp += 0.5 * dt * nu * (1 / (dy[J] * (dy[J - 1] + dy[J]) * 0.5));
            } else temp += 0.5 * dt * nu * (1 / (dy[J] * (dy[J] + dy[J]) * 0.5));
        } else {
            if (I != nx - 1) {
                eE = -0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I + 1]) * 0.5));
                temp += 0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I + 1]) * 0.5));
            } else temp += 0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I]) * 0.5));
            if (I != 0) {
                eW = -0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I - 1]) * 0.5));
                temp += 0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I - 1]) * 0.5));
            } else temp += 0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I]) * 0.5));
            const double cn = 0.5 * dt * nu * (1 / (dy[J + 1] * (dy[J] + dy[J + 1]) * 0.5));
            const double cs = 0.5 * dt * nu * (1 / (dy[J] * (dy[J] + dy[J + 1]) * 0.5));
            if (J != ny - 2) { eN = -0.5 * dt * nu * (1 / (dy[J + 1] * (dy[J] + dy[J + 1]) * 0.5)); temp += cn; }
            else temp += cn;
            if (J != 0) { eS = -0.5 * dt * nu * (1 / (dy[J] * (dy[J] + dy[J + 1]) * 0.5)); temp += cs; }
            else temp += cs;
        }
        if (J != 0) { ci[pos] = i - st; v[pos++] = eS; }
        if (I != 0) { ci[pos] = i - 1; v[pos++] = eW; }
        ci[pos] = i; v[pos++] = temp;
        if (I != lastI) { ci[pos] = i + 1; v[pos++] = eE; }
        if (J != lastJ) { ci[pos] = i + st; v[pos++] = eN; }
    }
}

__global__ void __launch_bounds__(BS) k_lhs2(const int *__restrict__ rp, int *__restrict__ ci, double *__restrict__ v,
                                             GridDev g, double dt, int centre_pin, FadlunLhsDev f) {
    const int nx = g.nx, ny = g.ny, numP = nx * ny;
    const double *dx = g.dx, *dy = g.dy;
    const int centre = (ny / 2) * nx + nx / 2;
    for (int ip = blockIdx.x * blockDim.x + threadIdx.x; ip < numP; ip += gridDim.x * blockDim.x) {
        const int I = ip % nx, J = ip / nx;
        const bool interior = !(I == 0 || I == nx - 1 || J == 0 || J == ny - 1);
        double eS = 0, eW = 0, eE = 0, eN = 0, temp = 0;
        if (f.ghostP && interior && f.hybridP[ip] != -1) {
            const double du = f.du[ip], dv = f.dv[ip];
            const bool cu = du > dx[I] / 2 && du < dx[I], cv = dv > dy[J] / 2 && dv < dy[J];
            if (f.ghostP[ip + 1] == -1) {
                if (cu) { eE = -dt / dx[I] / du; temp += dt / dx[I] / du; }
                else { eE = -dt / dx[I] / dx[I]; temp += dt / dx[I] / dx[I]; }
            }
            if (f.ghostP[ip - 1] == -1) {
                if (cu) { eW = -dt / dx[I] / du; temp += dt / dx[I] / du; }
                else { eW = -dt / dx[I] / dx[I]; temp += dt / dx[I] / dx[I]; }
            }
            if (f.ghostP[ip + nx] == -1) {
                if (cv) { eN = -dt / dy[J] / dv; temp += dt / dy[J] / dv; }
                else { eN = -dt / dy[J] / dy[J]; temp += dt / dy[J] / dy[J]; }
            }
            if (f.ghostP[ip - nx] == -1) {
                if (cv) { eS = -dt / dy[J] / dv; temp += dt / dy[J] / dv; }
                else { eS = -dt / dy[J] / dy[J]; temp += dt / dy[J] / dy[J]; }
            }
        } else if (f.ghostP && interior && f.ghostP[ip] > 0) {
            if (f.ghostP[ip + 1] == 0) { eE = -dt / dx[I] / dx[I]; temp += dt / dx[I] / dx[I]; }
            if (f.ghostP[ip - 1] == 0) { eW = -dt / dx[I] / dx[I]; temp += dt / dx[I] / dx[I]; }
            if (f.ghostP[ip + nx] == 0) { eN = -dt / dy[J] / dy[J]; temp += dt / dy[J] / dy[J]; }
            if (f.ghostP[ip - nx] == 0) { eS = -dt / dy[J] / dy[J]; temp += dt / dy[J] / dy[J]; }
        } else {
            if (I != nx - 1) { eE = -dt / (dx[I] * (dx[I] + dx[I + 1]) * 0.5); temp += dt / (dx[I] * (dx[I] + dx[I + 1]) * 0.5); }
            if (I != 0) { eW = -dt / (dx[I] * (dx[I] + dx[I - 1]) * 0.5); temp += dt / (dx[I] * (dx[I] + dx[I - 1]) * 0.5); }
            if (J != ny - 1) { eN = -dt / (dy[J] * (dy[J] + dy[J + 1]) * 0.5); temp += dt / (dy[J] * (dy[J] + dy[J + 1]) * 0.5); }
            if (J != 0) { eS = -dt / (dy[J] * (dy[J] + dy[J - 1]) * 0.5); temp += dt / (dy[J] * (dy[J] + dy[J - 1]) * 0.5); }
        }
        if (ip == 0) temp += temp;
        if (centre_pin && interior && ip == centre) temp += temp;
        int pos = rp[ip];
        if (J != 0) { ci[pos] = ip - nx; v[pos++] = eS; }
        if (I != 0) { ci[pos] = ip - 1; v[pos++] = eW; }
        ci[pos] = ip; v[pos++] = temp;
        if (I != nx - 1) { ci[pos] = ip + 1; v[pos++] = eE; }
        if (J != ny - 1) { ci[pos] = ip + nx; v[pos++] = eN; }
    }
}

static int grid_for(long n) {
    const long b = (n + BS - 1) / BS;
    return (int)(b < 65536 ? b : 65536);
}

void launch_lhs1(const int *rp, int *ci, double *v, const GridDev &g, double nu, double dt, const FadlunLhsDev &fad,
                 hipStream_t s) {
    const long n = (long)(g.nx - 1) * g.ny + (long)g.nx * (g.ny - 1);
    hipLaunchKernelGGL(k_lhs1, dim3(grid_for(n)), dim3(BS), 0, s, rp, ci, v, g, nu, dt, fad);
    CUIBM_LAUNCH_CHECK();
}

void launch_lhs2(const int *rp, int *ci, double *v, const GridDev &g, double dt, bool centre_pin,
                 const FadlunLhsDev &fad, hipStream_t s) {
    hipLaunchKernelGGL(k_lhs2, dim3(grid_for((long)g.nx * g.ny)), dim3(BS), 0, s, rp, ci, v, g, dt, (int)centre_pin, fad);
    CUIBM_LAUNCH_CHECK();
}

}  // namespace cuibm
