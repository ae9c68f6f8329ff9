// Staggered-grid stencil kernels of one plain Navier-Stokes step (FP64).
//
// One fused sweep replaces the reference's ten launches for the intermediate
// velocity right-hand side (N.cu:29-248, L.cu:24-193, intermediateVelocity.cu:29-143
// under src/solvers/NavierStokes/NavierStokes/kernels/): each thread owns one u or v
// node, evaluates explicit convection N, diffusion L and the Crank-Nicolson wall
// terms bc1 with the reference's operation order, and writes N and
// rhs1 = u + dt(0.5 Nold - 1.5 N + 0.5 L) + bc1 in one pass (L and bc1 are stored
// only when asked for). The modified-Fadlun row overrides
// (FadlunModified/kernels/intermediateVelocity.cu:44-63) are fused as an epilogue.
#include "batch.hpp"
#include "kernels.hpp"

namespace cuibm {

// N, L and bc1 for u node i (I,J); N.cu:29-128, L.cu:24-101, intermediateVelocity.cu:53-89
__device__ __forceinline__ void u_terms(const double *__restrict__ u, const GridDev &g, const BcDev &b, double nu,
                                        double dt, int i, double &Nv, double &Lv, double &Bv) {
    const int nx = g.nx, ny = g.ny;
    const int I = i % (nx - 1), J = i / (nx - 1), iv = (nx - 1) * ny + nx * J + I;
    const double *dx = g.dx, *dy = g.dy;
    const double ui = u[i];
    const double ue = (I == nx - 2) ? b.xp[J] : u[i + 1];
    const double uw = (I == 0) ? b.xm[J] : u[i - 1];
    double t = ui * (ue * dx[I] / (dx[I + 1] * (dx[I] + dx[I + 1])) + uw * (dx[I] / (dx[I] * (dx[I] + dx[I + 1])) - 1 / dx[I]) +
                     ui * (-dx[I] / (dx[I + 1] * (dx[I] + dx[I + 1])) - (dx[I] / (dx[I] * (dx[I] + dx[I + 1])) - 1 / dx[I])));
    if (J == ny - 1) {
        t += (((b.yp[(nx - 1) + I + 1] - b.yp[(nx - 1) + I]) * dx[I] / (dx[I] + dx[I + 1]) + b.yp[(nx - 1) + I]) / 2 +
              ((u[iv - nx + 1] - u[iv - nx]) * dx[I] / (dx[I] + dx[I + 1]) + u[iv - nx]) / 2) *
             ((2 * b.yp[I] - ui) * dy[J] / (dy[J] * (dy[J] + dy[J])) + u[i - nx + 1] * (dy[J - 1] / (dy[J] * (dy[J] + dy[J - 1])) - 1 / dy[J]) +
              ui * (-dy[J] / (dy[J] * (dy[J] + dy[J])) - (dy[J - 1] / (dy[J] * (dy[J] + dy[J - 1])) - 1 / dy[J])));
    } else if (J == 0) {
        t += (((u[iv + 1] - u[iv]) * dx[I] / (dx[I] + dx[I + 1]) + u[iv]) / 2 +
              ((b.ym[(nx - 1) + I + 1] - b.ym[(nx - 1) + I]) * dx[I] / (dx[I] + dx[I + 1]) + b.ym[(nx - 1) + I]) / 2) *
             (u[i + nx - 1] * dy[J] / (dy[J] * (dy[J + 1] + dy[J])) + (2 * b.ym[I] - ui) * (dy[J] / (dy[J] * (dy[J] + dy[J])) - 1 / dy[J]) +
              ui * (-dy[J] / (dy[J] * (dy[J + 1] + dy[J])) - (dy[J] / (dy[J] * (dy[J] + dy[J])) - 1 / dy[J])));
    } else {
        t += (((u[iv + 1] - u[iv]) * dx[I] / (dx[I] + dx[I + 1]) + u[iv]) / 2 +
              ((u[iv - nx + 1] - u[iv - nx]) * dx[I] / (dx[I] + dx[I + 1]) + u[iv - nx]) / 2) *
             (u[i + nx - 1] * dy[J] / (dy[J] * (dy[J + 1] + dy[J])) + u[i - nx + 1] * (dy[J - 1] / (dy[J] * (dy[J] + dy[J - 1])) - 1 / dy[J]) +
              ui * (-dy[J] / (dy[J] * (dy[J + 1] + dy[J])) - (dy[J - 1] / (dy[J] * (dy[J] + dy[J - 1])) - 1 / dy[J])));
    }
    Nv = t;
    double l = 0;
    if (I != nx - 2) l += nu * (u[i + 1] - ui) / (dx[I + 1] * (dx[I + 1] + dx[I]) * 0.5);
    else l += nu * (b.xp[J] - ui) / (dx[I + 1] * (dx[I + 1] + dx[I]) * 0.5);
    if (I != 0) l += nu * (u[i - 1] - ui) / (dx[I] * (dx[I + 1] + dx[I]) * 0.5);
    else l += nu * (b.xm[J] - ui) / (dx[I] * (dx[I + 1] + dx[I]) * 0.5);
    if (J != ny - 1) l += nu * (u[i + (nx - 1)] - ui) / (dy[J] * (dy[J + 1] + dy[J]) * 0.5);
    else l += nu * (2 * b.yp[I] - 2 * ui) / (dy[J] * dy[J]);
    if (J != 0) l += nu * (u[i - (nx - 1)] - ui) / (dy[J] * (dy[J - 1] + dy[J]) * 0.5);
    else l += nu * (2 * b.ym[I] - 2 * ui) / (dy[J] * dy[J]);
    Lv = l;
    double c = 0;
    if (I == nx - 2) c += b.xp[J] * 0.5 * dt * nu * (1 / (dx[I + 1] * (dx[I + 1] + dx[I]) * 0.5));
    if (I == 0) c += b.xm[J] * 0.5 * dt * nu * (1 / (dx[I] * (dx[I + 1] + dx[I]) * 0.5));
    if (J == ny - 1) c += (2 * b.yp[I] - ui) * nu * dt * 0.5 / dy[J] / dy[J];
    if (J == 0) c += (2 * b.ym[I] - ui) * nu * dt * 0.5 / dy[J] / dy[J];
    Bv = c;
}

// same for v node k = ip (I,J), stored at numU + k; N.cu:140-248, L.cu:113-193,
// intermediateVelocity.cu:107-143
__device__ __forceinline__ void v_terms(const double *__restrict__ u, const GridDev &g, const BcDev &b, double nu,
                                        double dt, int ip, double &Nv, double &Lv, double &Bv) {
    const int nx = g.nx, ny = g.ny;
    const int I = ip % nx, J = ip / nx, iu = (nx - 1) * J + I, iv = ip + (nx - 1) * ny;
    const double *dx = g.dx, *dy = g.dy;
    const double vi = u[iv];
    double t;
    if (I == nx - 1) {
        t = (((b.xp[J + 1] - b.xp[J]) * dy[J] / (dy[J] + dy[J + 1]) + b.xp[J]) / 2 +
             ((u[iu + (nx - 1) - 1] - u[iu - 1]) * dy[J] / (dy[J] + dy[J + 1]) + u[iu - 1]) / 2) *
            ((2 * b.xp[ny + J] - vi) * dx[I] / (dx[I] * (dx[I] + dx[I])) + u[iv - 1] * (dx[I - 1] / (dx[I] * (dx[I] + dx[I - 1])) - 1 / dx[I]) +
             vi * (-dx[I] / (dx[I] * (dx[I] + dx[I])) - (dx[I - 1] / (dx[I] * (dx[I] + dx[I - 1])) - 1 / dx[I])));
    } else if (I == 0) {
        t = (((u[iu + (nx - 1)] - u[iu]) * dy[J] / (dy[J] + dy[J + 1]) + u[iu]) / 2 +
             ((b.xm[J + 1] - b.xm[J]) * dy[J] / (dy[J] + dy[J + 1]) + b.xm[J]) / 2) *
            (u[iv + 1] * dx[I] / (dx[I] * (dx[I] + dx[I + 1])) + (2 * b.xm[ny + J] - vi) * (dx[I] / (dx[I] * (dx[I] + dx[I])) - 1 / dx[I]) +
             vi * (-dx[I] / (dx[I] * (dx[I] + dx[I + 1])) - (dx[I] / (dx[I] * (dx[I] + dx[I])) - 1 / dx[I])));
    } else {
        t = (((u[iu + (nx - 1)] - u[iu]) * dy[J] / (dy[J] + dy[J + 1]) + u[iu]) / 2 +
             ((u[iu + (nx - 1) - 1] - u[iu - 1]) * dy[J] / (dy[J] + dy[J + 1]) + u[iu - 1]) / 2) *
            (u[iv + 1] * dx[I] / (dx[I] * (dx[I] + dx[I + 1])) + u[iv - 1] * (dx[I - 1] / (dx[I] * (dx[I] + dx[I - 1])) - 1 / dx[I]) +
             vi * (-dx[I] / (dx[I] * (dx[I] + dx[I + 1])) - (dx[I - 1] / (dx[I] * (dx[I] + dx[I - 1])) - 1 / dx[I])));
    }
    const double vn = (J == ny - 2) ? b.yp[(nx - 1) + I] : u[iv + nx];
    const double vs = (J == 0) ? b.ym[(nx - 1) + I] : u[iv - nx];
    t += vi * (vn * dy[J] / (dy[J + 1] * (dy[J + 1] + dy[J])) + vs * (dy[J] / (dy[J] * (dy[J + 1] + dy[J])) - 1 / dy[J]) +
               vi * (-dy[J] / (dy[J + 1] * (dy[J + 1] + dy[J])) - (dy[J] / (dy[J] * (dy[J + 1] + dy[J])) - 1 / dy[J])));
    Nv = t;
    double l = 0;
    if (I != nx - 1) l += nu * (u[iv + 1] - vi) / (dx[I] * (dx[I] + dx[I + 1]) * 0.5);
    else l += nu * (2 * b.xp[ny + J] - 2 * vi) / (dx[I] * dx[I]);
    if (I != 0) l += nu * (u[iv - 1] - vi) / (dx[I] * (dx[I] + dx[I - 1]) * 0.5);
    else l += nu * (2 * b.xm[ny + J] - 2 * vi) / (dx[I] * dx[I]);
    if (J != ny - 2) l += nu * (u[iv + nx] - vi) / (dy[J + 1] * (dy[J] + dy[J + 1]) * 0.5);
    else l += nu * (b.yp[nx - 1 + I] - vi) / (dy[J + 1] * (dy[J] + dy[J + 1]) * 0.5);
    if (J != 0) l += nu * (u[iv - nx] - vi) / (dy[J] * (dy[J] + dy[J + 1]) * 0.5);
    else l += nu * (b.ym[nx - 1 + I] - vi) / (dy[J] * (dy[J] + dy[J + 1]) * 0.5);
    Lv = l;
    double c = 0;
    if (I == nx - 1) c += (2 * b.xp[ny + J] - vi) * nu * dt * 0.5 / dx[I] / dx[I];
    if (I == 0) c += (2 * b.xm[ny + J] - vi) * nu * dt * 0.5 / dx[I] / dx[I];
    if (J == ny - 2) c += b.yp[(nx - 1) + I] * 0.5 * dt * nu * (1 / (dy[J + 1] * (dy[J] + dy[J + 1]) * 0.5));
    if (J == 0) c += b.ym[(nx - 1) + I] * 0.5 * dt * nu * (1 / (dy[J] * (dy[J] + dy[J + 1]) * 0.5));
    Bv = c;
}

__global__ void __launch_bounds__(BS) k_rhs1(const double *__restrict__ u, const double *__restrict__ Nold,
                                             double *__restrict__ N, double *__restrict__ rhs, double *__restrict__ Lout,
                                             double *__restrict__ Bout, GridDev g, BcDev b, double nu, double dt,
                                             FadlunRhsDev fad) {
    const int numU = (g.nx - 1) * g.ny, numUV = numU + g.nx * (g.ny - 1);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < numUV; i += gridDim.x * blockDim.x) {
        double Nv, Lv, Bv;
        if (i < numU) u_terms(u, g, b, nu, dt, i, Nv, Lv, Bv);
        else v_terms(u, g, b, nu, dt, i - numU, Nv, Lv, Bv);
        N[i] = Nv;
        double r = u[i] + dt * (0.5 * Nold[i] - 1.5 * Nv + 0.5 * Lv) + Bv;
        if (fad.hybrid) {
            const int h = fad.hybrid[i], gt = fad.ghost[i];
            r = (h == -1) * (gt <= 0) * (r) + (h != -1) * fad.dist_between[i] / (fad.dist_int[i] + fad.dist_between[i]) * fad.uv[i];
        }
        rhs[i] = r;
        if (Lout) Lout[i] = Lv;
        if (Bout) Bout[i] = Bv;
    }
}

// -div(uhat) at cell ip of the plain solver (NavierStokes/kernels/intermediatePressure.cu:30-71)
__device__ __forceinline__ double rhs2_plain(const double *__restrict__ uhat, const GridDev &g, const BcDev &b, int ip) {
    const int nx = g.nx, ny = g.ny;
    const double *dx = g.dx, *dy = g.dy;
    const int I = ip % nx, J = ip / nx, iu = (nx - 1) * J + I, iv = (nx - 1) * ny + nx * J + I;
    double t = 0;
    if (I != nx - 1) t -= uhat[iu] / dx[I];
    else t -= b.xp[J] / dx[I];
    if (I != 0) t += uhat[iu - 1] / dx[I];
    else t += b.xm[J] / dx[I];
    if (J != ny - 1) t -= uhat[iv] / dy[J];
    else t -= b.yp[(nx - 1) + I] / dy[J];
    if (J != 0) t += uhat[iv - nx] / dy[J];
    else t += b.ym[(nx - 1) + I] / dy[J];
    return t;
}

// rhs2 = -div(uhat) (NavierStokes/kernels/intermediatePressure.cu:30-71); with
// `fad` set, the modified-Fadlun cut-cell variant (FadlunModified/kernels/intermediatePressure.cu:17-138)
__global__ void __launch_bounds__(BS) k_rhs2(const double *__restrict__ uhat, double *__restrict__ rhs2, GridDev g,
                                             BcDev b, FadlunP2Dev fad) {
    const int nx = g.nx, ny = g.ny, numP = nx * ny;
    const double *dx = g.dx, *dy = g.dy;
    for (int ip = blockIdx.x * blockDim.x + threadIdx.x; ip < numP; ip += gridDim.x * blockDim.x) {
        const int I = ip % nx, J = ip / nx, iu = (nx - 1) * J + I, iv = (nx - 1) * ny + nx * J + I;
        double t = 0;
        const int hp = fad.hybridP ? fad.hybridP[ip] : -1;
        const int gp = fad.ghostP ? fad.ghostP[ip] : -1;
        if (hp != -1) {
            const double du = fad.du[ip], dv = fad.dv[ip];
            const int *G = fad.ghostP;
            if (G[ip + 1] == -1) {
                if (du > dx[I] / 2 && du < dx[I]) t -= uhat[iu] / du;
                else t -= uhat[iu] / dx[I];
            }
            if (G[ip - 1] == -1) {
                if (du > dx[I] / 2 && du < dx[I]) t += uhat[iu - 1] / du;
                else t += uhat[iu - 1] / dx[I];
            }
            if (G[ip + nx] == -1) {
                if (dv > dy[J] / 2 && dv < dy[J]) t -= uhat[iv] / dv;
                else t -= uhat[iv] / dy[J];
            }
            if (G[ip - nx] == -1) {
                if (dv > dy[J] / 2 && dv < dy[J]) t += uhat[iv - nx] / dv;
                else t += uhat[iv - nx] / dy[J];
            }
        } else if (gp > 0) {
            const int *G = fad.ghostP;
            if (G[ip + 1] == 0) t -= uhat[iu] / dx[I];
            if (G[ip - 1] == 0) t += uhat[iu - 1] / dx[I];
            if (G[ip + nx] == 0) t -= uhat[iv] / dy[J];
            if (G[ip - nx] == 0) t += uhat[iv - nx] / dy[J];
        } else {
            t = rhs2_plain(uhat, g, b, ip);
        }
        rhs2[ip] = t;
    }
}

// uhat - dt grad p at velocity node i of the plain solver (NavierStokes/kernels/projectVelocity.cu:25-70)
__device__ __forceinline__ double project_plain(const double *__restrict__ uhat, const double *__restrict__ p,
                                                const GridDev &g, double dt, int i) {
    const int nx = g.nx, ny = g.ny, numU = (nx - 1) * ny;
    if (i < numU) {
        const int I = i % (nx - 1), J = i / (nx - 1), ip = J * nx + I;
        return uhat[i] - dt * (p[ip + 1] - p[ip]) / (0.5 * g.dx[I + 1] + 0.5 * g.dx[I]);
    }
    const int k = i - numU, I = k % nx, J = k / nx, ip = J * nx + I;
    return uhat[i] - dt * (p[ip + nx] - p[ip]) / (0.5 * g.dy[J + 1] + 0.5 * g.dy[J]);
}

// u = uhat - dt grad p (NavierStokes/kernels/projectVelocity.cu:25-70, writes uold);
// with ghostP set, the masked Fadlun projection (FadlunModified/kernels/projectVelocity.cu:14-44,
// no uold update)
__global__ void __launch_bounds__(BS) k_project(double *__restrict__ u, const double *__restrict__ uhat,
                                                double *__restrict__ uold, const double *__restrict__ p, GridDev g,
                                                double dt, const int *__restrict__ ghostP) {
    const int nx = g.nx, ny = g.ny, numU = (nx - 1) * ny, numUV = numU + nx * (ny - 1);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < numUV; i += gridDim.x * blockDim.x) {
        double nu_;
        if (!ghostP) {
            nu_ = project_plain(uhat, p, g, dt, i);
        } else if (i < numU) {
            const int I = i % (nx - 1), J = i / (nx - 1), ip = J * nx + I;
            nu_ = uhat[i] - (ghostP[ip + 1] == -1 && ghostP[ip] == -1) * dt * (p[ip + 1] - p[ip]) / (0.5 * g.dx[I + 1] + 0.5 * g.dx[I]);
        } else {
            const int k = i - numU, I = k % nx, J = k / nx, ip = J * nx + I;
            nu_ = uhat[i] - (ghostP[ip + nx] == -1 && ghostP[ip] == -1) * dt * (p[ip + nx] - p[ip]) / (0.5 * g.dy[J + 1] + 0.5 * g.dy[J]);
        }
        if (!ghostP) uold[i] = u[i];
        u[i] = nu_;
    }
}

// convective outflow at x+ (FadlunModified/kernels/intermediateVelocity.cu:19-41),
// Uinf = Vinf = 1 as hard-coded in luo_base/intermediateVelocity.cu:13-14
__device__ __forceinline__ void robin_one(const double *__restrict__ u, double *__restrict__ xp, const GridDev &g,
                                          double dt, int t) {
    const int nx = g.nx, ny = g.ny, numU = (nx - 1) * ny;
    if (t < ny) {
        const int J = t, i = J * (nx - 1) + nx - 1;
        const double beta = 1.0 * dt / g.dx[nx - 1];
        xp[J] = xp[J] * (1 - beta) + beta * u[i - 1];
    } else if (t < 2 * ny - 1) {
        const int J = t - ny, i = J * nx + nx;
        const double beta = 1.0 * dt / g.dx[nx - 1];
        xp[J + ny] = xp[J + ny] * (1 - beta) + beta * u[i + numU - 1];
    }
}

__global__ void k_robin(const double *__restrict__ u, double *__restrict__ xp, GridDev g, double dt) {
    robin_one(u, xp, g, dt, blockIdx.x * blockDim.x + threadIdx.x);
}

static int grid_for(long n) {
    const long b = (n + BS - 1) / BS;
    return (int)(b < 65536 ? b : 65536);
}

void launch_rhs1(const double *u, const double *Nold, double *N, double *rhs, double *L, double *bc1, const GridDev &g,
                 const BcDev &b, double nu, double dt, const FadlunRhsDev &fad, hipStream_t s) {
    const long n = (long)(g.nx - 1) * g.ny + (long)g.nx * (g.ny - 1);
    hipLaunchKernelGGL(k_rhs1, dim3(grid_for(n)), dim3(BS), 0, s, u, Nold, N, rhs, L, bc1, g, b, nu, dt, fad);
    CUIBM_LAUNCH_CHECK();
}

void launch_rhs2(const double *uhat, double *rhs2, const GridDev &g, const BcDev &b, const FadlunP2Dev &fad,
                 hipStream_t s) {
    hipLaunchKernelGGL(k_rhs2, dim3(grid_for((long)g.nx * g.ny)), dim3(BS), 0, s, uhat, rhs2, g, b, fad);
    CUIBM_LAUNCH_CHECK();
}

void launch_project(double *u, const double *uhat, double *uold, const double *p, const GridDev &g, double dt,
                    const int *ghostP, hipStream_t s) {
    const long n = (long)(g.nx - 1) * g.ny + (long)g.nx * (g.ny - 1);
    hipLaunchKernelGGL(k_project, dim3(grid_for(n)), dim3(BS), 0, s, u, uhat, uold, p, g, dt, ghostP);
    CUIBM_LAUNCH_CHECK();
}

// ---- CFL monitor (NavierStokes/kernels/CFL.cu:10-23 + the thrust::max_element of
// NavierStokes/CFL.cu:19-21): per-cell dt (|u|/dx + |v|/dy), its maximum and the first
// index attaining it; the reference never writes the last column and row of its array,
// which stay zero
__device__ __forceinline__ void argmax_merge(double &v, int &i, double v2, int i2) {
    if (v2 > v || (v2 == v && i2 < i)) {
        v = v2;
        i = i2;
    }
}

__device__ void block_argmax(double &best, int &bi) {
    for (int o = 32; o > 0; o >>= 1) {
        const double v2 = __shfl_down(best, o, 64);
        const int i2 = __shfl_down(bi, o, 64);
        argmax_merge(best, bi, v2, i2);
    }
    __shared__ double sv[BS / 64];
    __shared__ int si[BS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        sv[wave] = best;
        si[wave] = bi;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < BS / 64; w++) argmax_merge(best, bi, sv[w], si[w]);
}

__global__ void __launch_bounds__(BS) k_cfl_part(const double *__restrict__ u, GridDev g, double dt,
                                                 double *__restrict__ pv, int *__restrict__ pi) {
    const int nx = g.nx, ny = g.ny, n = nx * ny, numU = (nx - 1) * ny;
    double best = -1.0;
    int bi = n;
    for (int ip = blockIdx.x * blockDim.x + threadIdx.x; ip < n; ip += gridDim.x * blockDim.x) {
        const int I = ip % nx, J = ip / nx;
        double c = 0;
        if (I != nx - 1 && J != ny - 1)
            c = dt * (fabs(u[(nx - 1) * J + I]) / g.dx[I] + fabs(u[numU + nx * J + I]) / g.dy[J]);
        argmax_merge(best, bi, c, ip);
    }
    block_argmax(best, bi);
    if (threadIdx.x == 0) {
        pv[blockIdx.x] = best;
        pi[blockIdx.x] = bi;
    }
}

__global__ void __launch_bounds__(BS) k_cfl_final(const double *__restrict__ pv, const int *__restrict__ pi, int nb,
                                                  double *__restrict__ out) {
    double best = -1.0;
    int bi = 0x7fffffff;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) argmax_merge(best, bi, pv[k], pi[k]);
    block_argmax(best, bi);
    if (threadIdx.x == 0) {
        out[0] = best;
        out[1] = bi;
    }
}

void launch_cfl(const double *u, const GridDev &g, double dt, double *part, double *out, hipStream_t s) {
    int *pi = reinterpret_cast<int *>(part + CFL_BLOCKS);
    hipLaunchKernelGGL(k_cfl_part, dim3(CFL_BLOCKS), dim3(BS), 0, s, u, g, dt, part, pi);
    CUIBM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(BS), 0, s, part, pi, CFL_BLOCKS, out);
    CUIBM_LAUNCH_CHECK();
}

void launch_robin(const double *u, double *xp, const GridDev &g, double dt, hipStream_t s) {
    hipLaunchKernelGGL(k_robin, dim3(nblk(2 * g.ny - 1)), dim3(BS), 0, s, u, xp, g, dt);
    CUIBM_LAUNCH_CHECK();
}

// ---- lockstep batches: the same sweeps over every case of a batch, a block map giving
// each block its case and its block within the case ----

__global__ void __launch_bounds__(BS) k_b_robin(const BCase *__restrict__ tab, BMapDev m, const double *__restrict__ u,
                                                double *__restrict__ xp) {
    const BCase &C = tab[m.cs[blockIdx.x]];
    const int t = m.loc[blockIdx.x] * BS + threadIdx.x;
    if (t < 2 * C.g.ny - 1) robin_one(u + C.offUV, xp + C.offXP, C.g, C.dt, t);
}

__global__ void __launch_bounds__(BS) k_b_rhs1(const BCase *__restrict__ tab, BMapDev m, const double *__restrict__ u,
                                               const double *__restrict__ Nold, double *__restrict__ N,
                                               double *__restrict__ rhs) {
    const BCase &C = tab[m.cs[blockIdx.x]];
    const int i = m.loc[blockIdx.x] * BS + threadIdx.x;
    if (i >= C.numUV) return;
    const double *uc = u + C.offUV;
    const int numU = (C.g.nx - 1) * C.g.ny;
    double Nv, Lv, Bv;
    if (i < numU) u_terms(uc, C.g, C.bc, C.nu, C.dt, i, Nv, Lv, Bv);
    else v_terms(uc, C.g, C.bc, C.nu, C.dt, i - numU, Nv, Lv, Bv);
    N[C.offUV + i] = Nv;
    rhs[C.offUV + i] = uc[i] + C.dt * (0.5 * Nold[C.offUV + i] - 1.5 * Nv + 0.5 * Lv) + Bv;
}

__global__ void __launch_bounds__(BS) k_b_rhs2(const BCase *__restrict__ tab, BMapDev m,
                                               const double *__restrict__ uhat, double *__restrict__ rhs2) {
    const BCase &C = tab[m.cs[blockIdx.x]];
    const int ip = m.loc[blockIdx.x] * BS + threadIdx.x;
    if (ip < C.numP) rhs2[C.offP + ip] = rhs2_plain(uhat + C.offUV, C.g, C.bc, ip);
}

__global__ void __launch_bounds__(BS) k_b_project(const BCase *__restrict__ tab, BMapDev m, double *__restrict__ u,
                                                  const double *__restrict__ uhat, double *__restrict__ uold,
                                                  const double *__restrict__ p) {
    const BCase &C = tab[m.cs[blockIdx.x]];
    const int i = m.loc[blockIdx.x] * BS + threadIdx.x;
    if (i >= C.numUV) return;
    const double v = project_plain(uhat + C.offUV, p + C.offP, C.g, C.dt, i);
    uold[C.offUV + i] = u[C.offUV + i];
    u[C.offUV + i] = v;
}

void launch_b_robin(const BCase *tab, const BMapDev &m, const double *u, double *xp, hipStream_t s) {
    hipLaunchKernelGGL(k_b_robin, dim3(m.nblk), dim3(BS), 0, s, tab, m, u, xp);
    CUIBM_LAUNCH_CHECK();
}

void launch_b_rhs1(const BCase *tab, const BMapDev &m, const double *u, const double *Nold, double *N, double *rhs,
                   hipStream_t s) {
    hipLaunchKernelGGL(k_b_rhs1, dim3(m.nblk), dim3(BS), 0, s, tab, m, u, Nold, N, rhs);
    CUIBM_LAUNCH_CHECK();
}

void launch_b_rhs2(const BCase *tab, const BMapDev &m, const double *uhat, double *rhs2, hipStream_t s) {
    hipLaunchKernelGGL(k_b_rhs2, dim3(m.nblk), dim3(BS), 0, s, tab, m, uhat, rhs2);
    CUIBM_LAUNCH_CHECK();
}

void launch_b_project(const BCase *tab, const BMapDev &m, double *u, const double *uhat, double *uold,
                      const double *p, hipStream_t s) {
    hipLaunchKernelGGL(k_b_project, dim3(m.nblk), dim3(BS), 0, s, tab, m, u, uhat, uold, p);
    CUIBM_LAUNCH_CHECK();
}

}  // namespace cuibm
