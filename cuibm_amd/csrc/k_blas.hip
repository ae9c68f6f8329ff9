// Vector, reduction and sparse-product kernels for the Krylov and AMG solvers.
//
// Reductions are deterministic: a fixed grid of RED_GRID blocks strides over the
// vector, each block reduces its threads in a fixed tree (wavefront shuffles, then
// LDS across the 4 wavefronts of a 256-thread block) and writes one partial; a
// single-block kernel then sums the partials in a fixed order. The result depends
// only on n, never on scheduling.
#include "batch.hpp"
#include "krylov.hpp"
#include "kernels.hpp"

namespace cuibm {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// reduce `k` values per thread across a 256-thread block; thread 0 gets the sums
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K]) {
    __shared__ double sh[K][BS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; k++) sh[k][wid] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = (sh[k][0] + sh[k][1]) + (sh[k][2] + sh[k][3]);
}

template <class T>
__global__ void k_fill(T *x, T v, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] = v;
}
__global__ void k_copy(double *__restrict__ y, const double *__restrict__ x, long n, const double *__restrict__ stop) {
    const double stopv = stop ? *stop : 0.0;  // read at once, tested after the first load
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double v = x[i];
        if (stopv != 0.0) return;
        y[i] = v;
    }
}
__global__ void k_to_float(float *__restrict__ y, const double *__restrict__ x, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = (float)x[i];
}
__global__ void k_to_double(double *__restrict__ y, const float *__restrict__ x, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = (double)x[i];
}

static int grid_for(long n) {
    const long b = (n + BS - 1) / BS;
    return (int)(b < 65536 ? b : 65536);
}

void launch_fill(double *x, double v, long n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(n)), dim3(BS), 0, s, x, v, n);
    CUIBM_LAUNCH_CHECK();
}
void launch_fill_f(float *x, float v, long n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill<float>, dim3(grid_for(n)), dim3(BS), 0, s, x, v, n);
    CUIBM_LAUNCH_CHECK();
}
void launch_fill_i(int *x, int v, long n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill<int>, dim3(grid_for(n)), dim3(BS), 0, s, x, v, n);
    CUIBM_LAUNCH_CHECK();
}
void launch_copy(double *y, const double *x, long n, hipStream_t s, const double *stop) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(BS), 0, s, y, x, n, stop);
    CUIBM_LAUNCH_CHECK();
}
void launch_to_float(float *y, const double *x, long n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_to_float, dim3(grid_for(n)), dim3(BS), 0, s, y, x, n);
    CUIBM_LAUNCH_CHECK();
}
void launch_to_double(double *y, const float *x, long n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_to_double, dim3(grid_for(n)), dim3(BS), 0, s, y, x, n);
    CUIBM_LAUNCH_CHECK();
}

// ---- sparse matrix-vector products --------------------------------------------------

template <class V, class X>
__device__ __forceinline__ X row_dot(const int *__restrict__ rp, const int *__restrict__ ci, const V *__restrict__ v,
                                     const X *__restrict__ x, int i) {
    return (X)csr_row_sum<double>(ci, v, x, rp[i], rp[i + 1]);  // FP64 accumulation also for float operands
}

template <class V, class X>
__global__ void __launch_bounds__(BS) k_spmv(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                             const V *__restrict__ v, const X *__restrict__ x, X *__restrict__ y) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = row_dot(rp, ci, v, x, i);
}

template <class V, class X>
void launch_spmv(const DCsr<V> &A, const X *x, X *y, hipStream_t s) {
    if (A.n <= 0) return;
    hipLaunchKernelGGL((k_spmv<V, X>), dim3(grid_for(A.n)), dim3(BS), 0, s, A.n, A.rp.p(), A.ci.p(), A.v.p(), x, y);
    CUIBM_LAUNCH_CHECK();
}
template void launch_spmv<double, double>(const DCsr<double> &, const double *, double *, hipStream_t);
template void launch_spmv<float, double>(const DCsr<float> &, const double *, double *, hipStream_t);

// y = A x with block partials of (y . a) and, when b2 is set, (y . y)
__global__ void __launch_bounds__(BS) k_spmv_dot(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                                 const double *__restrict__ v, const double *__restrict__ x,
                                                 double *__restrict__ y, const double *__restrict__ a, int want_yy,
                                                 double *__restrict__ part, const double *__restrict__ stop) {
    const double stopv = stop ? *stop : 0.0;  // read at once, tested after the first row
    double acc[2] = {0, 0};
    bool tested = false;
    // The first nb blocks (a multiple of 8, so spread evenly over the XCDs) each take a
    // contiguous chunk of rows, and the chunks of one XCD are contiguous too, so that the
    // gathers of x from neighbouring rows stay in that XCD's L2; the other blocks write
    // zero partials.
    const int nb = min(RED_GRID, (int)((n / BS + 8) / 8 * 8));
    const int chunk = (n + nb - 1) / nb;
    const int r0 = (int)blockIdx.x < nb ? xcd_block(nb) * chunk : n, r1 = min(n, r0 + chunk);
    for (int i = r0 + threadIdx.x; i < r1; i += BS) {
        const double yi = row_dot(rp, ci, v, x, i);
        if (!tested) {
            if (stopv != 0.0) return;  // every thread returns here or below: no barrier is reached
            tested = true;
        }
        y[i] = yi;
        acc[0] += yi * a[i];
        acc[1] += yi * yi;
    }
    if (stopv != 0.0) return;
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = acc[0];
        if (want_yy) part[RED_GRID + blockIdx.x] = acc[1];
    }
}

void launch_spmv_dot(const DCsr<double> &A, const double *x, double *y, const double *a, bool want_yy, double *part,
                     hipStream_t s, const double *stop) {
    hipLaunchKernelGGL(k_spmv_dot, dim3(RED_GRID), dim3(BS), 0, s, A.n, A.rp.p(), A.ci.p(), A.v.p(), x, y, a,
                       (int)want_yy, part, stop);
    CUIBM_LAUNCH_CHECK();
}

// r = b - A x; also p = r and rs = r when given; partials of (r.r) and (b.b)
__global__ void __launch_bounds__(BS) k_bicg_init(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                                  const double *__restrict__ v, const double *__restrict__ x,
                                                  const double *__restrict__ b, double *__restrict__ r,
                                                  double *__restrict__ p, double *__restrict__ rs,
                                                  double *__restrict__ part) {
    double acc[2] = {0, 0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double bi = b[i];
        const double ri = bi - row_dot(rp, ci, v, x, i);
        r[i] = ri;
        p[i] = ri;
        rs[i] = ri;
        acc[0] += ri * ri;
        acc[1] += bi * bi;
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = acc[0];
        part[RED_GRID + blockIdx.x] = acc[1];
    }
}

void launch_bicg_init(const DCsr<double> &A, const double *x, const double *b, double *r, double *p, double *rs,
                      double *part, hipStream_t s) {
    hipLaunchKernelGGL(k_bicg_init, dim3(RED_GRID), dim3(BS), 0, s, A.n, A.rp.p(), A.ci.p(), A.v.p(), x, b, r, p, rs,
                       part);
    CUIBM_LAUNCH_CHECK();
}

// s = r - alpha AMp with partials of (s.s)
template <bool W>
__global__ void __launch_bounds__(BS) k_bicg_s(int n, const double *__restrict__ r, const double *__restrict__ AMp,
                                               double *__restrict__ s, const double *__restrict__ sc,
                                               double *__restrict__ part, InputWeights w) {
    const double stopv = sc[SC_STOP], alpha = sc[SC_ALPHA];
    double acc[2] = {0, 0};
    bool tested = false;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double si = r[i] - alpha * AMp[i];
        if (!tested) {
            if (stopv != 0.0) return;  // every thread returns here or below: no barrier is reached
            tested = true;
        }
        s[i] = si;
        acc[0] += si * si;
        if (W) acc[1] += (w.wx[i % w.nx] * w.wy[i / w.nx]) * si;
    }
    if (stopv != 0.0) return;
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = acc[0];
        if (W) w.part[blockIdx.x] = acc[1];
    }
}

void launch_bicg_s(int n, const double *r, const double *AMp, double *s_, const double *sc, double *part,
                   hipStream_t s, const InputWeights *w) {
    if (w) hipLaunchKernelGGL(k_bicg_s<true>, dim3(RED_GRID), dim3(BS), 0, s, n, r, AMp, s_, sc, part, *w);
    else hipLaunchKernelGGL(k_bicg_s<false>, dim3(RED_GRID), dim3(BS), 0, s, n, r, AMp, s_, sc, part, InputWeights{});
    CUIBM_LAUNCH_CHECK();
}

// x += alpha Mp + omega Ms; r = s - omega AMs; partials of (rs.r) and (r.r). In the
// iteration that ended at its half step (SC_STOP = STOP_HALF_PENDING) it applies the
// half-step update x += alpha Mp alone.
__global__ void __launch_bounds__(BS) k_bicg_xr(int n, double *__restrict__ x, const double *__restrict__ Mp,
                                                const double *__restrict__ Ms, double *__restrict__ r,
                                                const double *__restrict__ s, const double *__restrict__ AMs,
                                                const double *__restrict__ rs, const double *__restrict__ sc,
                                                double *__restrict__ part) {
    const double stopv = sc[SC_STOP], alpha = sc[SC_ALPHA], omega = sc[SC_OMEGA];
    const bool half = stopv == STOP_HALF_PENDING;
    double acc[2] = {0, 0};
    bool tested = false;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double xh = x[i] + alpha * Mp[i];
        const double xi = xh + omega * Ms[i];
        const double ri = s[i] - omega * AMs[i];
        if (!tested) {
            if (stopv != 0.0 && !half) return;  // every thread returns here or below: no barrier is reached
            tested = true;
        }
        if (half) {
            x[i] = xh;
            continue;
        }
        x[i] = xi;
        r[i] = ri;
        acc[0] += rs[i] * ri;
        acc[1] += ri * ri;
    }
    if (stopv != 0.0) return;  // includes the half-step update
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = acc[0];
        part[RED_GRID + blockIdx.x] = acc[1];
    }
}

void launch_bicg_xr(int n, double *x, const double *Mp, const double *Ms, double *r, const double *s_,
                    const double *AMs, const double *rs, const double *sc, double *part, hipStream_t s) {
    hipLaunchKernelGGL(k_bicg_xr, dim3(RED_GRID), dim3(BS), 0, s, n, x, Mp, Ms, r, s_, AMs, rs, sc, part);
    CUIBM_LAUNCH_CHECK();
}

// p = r + beta p - beta omega AMp
__global__ void k_bicg_p(int n, double *__restrict__ p, const double *__restrict__ r, const double *__restrict__ AMp,
                         const double *__restrict__ sc) {
    const double stopv = sc[SC_STOP], beta = sc[SC_BETA], omega = sc[SC_OMEGA];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double v = r[i] + beta * p[i] - beta * omega * AMp[i];
        if (stopv != 0.0) return;
        p[i] = v;
    }
}

// the same with the weighted partial sums of the new p (a fixed grid of RED_GRID blocks)
__global__ void __launch_bounds__(BS) k_bicg_pw(int n, double *__restrict__ p, const double *__restrict__ r,
                                                const double *__restrict__ AMp, const double *__restrict__ sc,
                                                InputWeights w) {
    const double stopv = sc[SC_STOP], beta = sc[SC_BETA], omega = sc[SC_OMEGA];
    double acc[1] = {0};
    bool tested = false;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double v = r[i] + beta * p[i] - beta * omega * AMp[i];
        if (!tested) {
            if (stopv != 0.0) return;  // every thread returns here or below: no barrier is reached
            tested = true;
        }
        p[i] = v;
        acc[0] += (w.wx[i % w.nx] * w.wy[i / w.nx]) * v;
    }
    if (stopv != 0.0) return;
    block_sum<1>(acc);
    if (threadIdx.x == 0) w.part[blockIdx.x] = acc[0];
}

void launch_bicg_p(int n, double *p, const double *r, const double *AMp, const double *sc, hipStream_t s,
                   const InputWeights *w) {
    if (w) hipLaunchKernelGGL(k_bicg_pw, dim3(RED_GRID), dim3(BS), 0, s, n, p, r, AMp, sc, *w);
    else hipLaunchKernelGGL(k_bicg_p, dim3(grid_for(n)), dim3(BS), 0, s, n, p, r, AMp, sc);
    CUIBM_LAUNCH_CHECK();
}

// the scalar updates and stopping tests of one finish, from the summed partials tot
__device__ __forceinline__ void finish_update(int op, const double (&tot)[2], double *__restrict__ sc, double tol,
                                              int maxit) {
    switch (op) {
        case FIN_INIT:  // (r.r, b.b); a (re)start clears the stop code, the count goes on
            sc[SC_RR] = tot[0];
            sc[SC_RHO] = tot[0];
            sc[SC_BB] = tot[1];
            sc[SC_STOP] = 0.0;
            break;
        case FIN_ALPHA:  // rs.AMp
            sc[SC_ALPHA] = sc[SC_RHO] / tot[0];
            break;
        case FIN_SS:
            sc[SC_SS] = tot[0];
            if (!isfinite(tot[0]) || !isfinite(sc[SC_ALPHA])) {
                sc[SC_STOP] = STOP_BREAK_HALF;
            } else if (sqrt(tot[0]) <= tol) {
                sc[SC_STOP] = STOP_HALF_PENDING;
                sc[SC_HALFK] = sc[SC_ITS];
            }
            break;
        case FIN_OMEGA: {  // (AMs.s, AMs.AMs); a breakdown leaves omega = 0 so that x stays finite
            const double w = tot[0] / tot[1];
            sc[SC_OMEGA] = isfinite(w) ? w : 0.0;
            break;
        }
        case FIN_BETA: {  // (rs.r, r.r)
            sc[SC_BETA] = (tot[0] / sc[SC_RHO]) * (sc[SC_ALPHA] / sc[SC_OMEGA]);
            sc[SC_RHO] = tot[0];
            sc[SC_RR] = tot[1];
            sc[SC_ITS] = sc[SC_ITS] + 1.0;
            const double rn = sqrt(tot[1]);
            if (rn <= tol)
                sc[SC_STOP] = STOP_CONVERGED;
            else if (sc[SC_OMEGA] == 0.0 || sc[SC_RHO] == 0.0 || !isfinite(sc[SC_BETA]) || !isfinite(rn))
                sc[SC_STOP] = STOP_BREAK_FULL;
            else if (sc[SC_ITS] >= (double)maxit)
                sc[SC_STOP] = STOP_MAXIT;
            break;
        }
        case FIN_DOT2:
            sc[SC_D0] = tot[0];
            sc[SC_D1] = tot[1];
            break;
    }
}

// sum the partial groups and update the recurrence scalars (one block); FIN_SS and
// FIN_BETA then apply the host's stopping tests (cusp::monitor on ||s|| and ||r||, the
// breakdown tests, the iteration limit) and record the outcome in SC_STOP
__global__ void __launch_bounds__(1024) k_bicg_finish(const double *__restrict__ part, int op, double *__restrict__ sc,
                                                      double tol, int maxit) {
    const double stopv = (op != FIN_INIT && op != FIN_DOT2) ? sc[SC_STOP] : 0.0;  // tested after the loads
    if (op == FIN_BETA && stopv == STOP_HALF_PENDING) {  // the half-step update is done
        if (threadIdx.x == 0) sc[SC_STOP] = STOP_HALF;
        return;
    }
    __shared__ double sh[2][16];
    double v[2] = {0, 0};
    const int ng = (op == FIN_ALPHA || op == FIN_SS) ? 1 : 2;
    for (int g = 0; g < ng; g++) {
        double t = 0;
        for (int k = threadIdx.x; k < RED_GRID; k += blockDim.x) t += part[g * RED_GRID + k];
        v[g] = t;
    }
    if (stopv != 0.0) return;  // uniform across the block, before the barrier
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int g = 0; g < ng; g++) {
        double t = wave_sum(v[g]);
        if (lane == 0) sh[g][wid] = t;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double tot[2] = {0, 0};
    for (int g = 0; g < ng; g++) {
        double t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += sh[g][w];
        tot[g] = t;
    }
    finish_update(op, tot, sc, tol, maxit);
}

void launch_bicg_finish(const double *part, int op, double *sc, hipStream_t s, double tol, int maxit) {
    hipLaunchKernelGGL(k_bicg_finish, dim3(1), dim3(1024), 0, s, part, op, sc, tol, maxit);
    CUIBM_LAUNCH_CHECK();
}

// partial dot products (x.y) and (x.x)
__global__ void __launch_bounds__(BS) k_dot2(long n, const double *__restrict__ x, const double *__restrict__ y,
                                             double *__restrict__ part) {
    double acc[2] = {0, 0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        acc[0] += x[i] * y[i];
        acc[1] += x[i] * x[i];
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = acc[0];
        part[RED_GRID + blockIdx.x] = acc[1];
    }
}

void launch_dot2(long n, const double *x, const double *y, double *part, hipStream_t s) {
    hipLaunchKernelGGL(k_dot2, dim3(RED_GRID), dim3(BS), 0, s, n, x, y, part);
    CUIBM_LAUNCH_CHECK();
}

// z = dinv * r (Jacobi)
__global__ void k_jacobi_apply(int n, const double *__restrict__ dinv, const double *__restrict__ r,
                               double *__restrict__ z, const double *__restrict__ stop) {
    const double stopv = stop ? *stop : 0.0;  // read at once, tested after the first loads
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double v = dinv[i] * r[i];
        if (stopv != 0.0) return;
        z[i] = v;
    }
}

void launch_jacobi_apply(int n, const double *dinv, const double *r, double *z, hipStream_t s, const double *stop) {
    hipLaunchKernelGGL(k_jacobi_apply, dim3(grid_for(n)), dim3(BS), 0, s, n, dinv, r, z, stop);
    CUIBM_LAUNCH_CHECK();
}

// dinv = 1 / (sum of the diagonal entries of each row)
__global__ void k_diag_inv(int n, const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ v,
                           double *__restrict__ dinv) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double d = 0;
        for (int k = rp[i]; k < rp[i + 1]; k++)
            if (ci[k] == i) d += v[k];
        dinv[i] = 1.0 / d;
    }
}

void launch_diag_inv(const DCsr<double> &A, double *dinv, hipStream_t s) {
    hipLaunchKernelGGL(k_diag_inv, dim3(grid_for(A.n)), dim3(BS), 0, s, A.n, A.rp.p(), A.ci.p(), A.v.p(), dinv);
    CUIBM_LAUNCH_CHECK();
}

// ---- lockstep batches: BiCGSTAB vector updates with per-case scalars ----
// Each block belongs to one case, so a stopped case's blocks return before any barrier.

struct BSeg {
    double *sc;
    long off;
    int n;
};
__device__ __forceinline__ BSeg bseg(int op, const BCase &C) {
    return op == BOP_VEL ? BSeg{C.scV, C.offUV, C.numUV} : BSeg{C.scP, C.offP, C.numP};
}

__global__ void __launch_bounds__(BS) k_b_bicg_s(int op, const BCase *__restrict__ tab, BMapDev m,
                                                 const double *__restrict__ r, const double *__restrict__ AMp,
                                                 double *__restrict__ s, double *__restrict__ part) {
    const BSeg g = bseg(op, tab[m.cs[blockIdx.x]]);
    if (g.sc[SC_STOP] != 0.0) return;
    const double alpha = g.sc[SC_ALPHA];
    const int i = m.loc[blockIdx.x] * BS + threadIdx.x;
    double acc[1] = {0};
    if (i < g.n) {
        const double si = r[g.off + i] - alpha * AMp[g.off + i];
        s[g.off + i] = si;
        acc[0] = si * si;
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

__global__ void __launch_bounds__(BS) k_b_bicg_xr(int op, const BCase *__restrict__ tab, BMapDev m,
                                                  double *__restrict__ x, const double *__restrict__ Mp,
                                                  const double *__restrict__ Ms, double *__restrict__ r,
                                                  const double *__restrict__ s, const double *__restrict__ AMs,
                                                  const double *__restrict__ rs, double *__restrict__ part) {
    const BSeg g = bseg(op, tab[m.cs[blockIdx.x]]);
    const double stopv = g.sc[SC_STOP];
    const bool half = stopv == STOP_HALF_PENDING;
    if (stopv != 0.0 && !half) return;
    const double alpha = g.sc[SC_ALPHA], omega = g.sc[SC_OMEGA];
    const int i = m.loc[blockIdx.x] * BS + threadIdx.x;
    double acc[2] = {0, 0};
    if (i < g.n) {
        const long j = g.off + i;
        const double xh = x[j] + alpha * Mp[j];
        if (half) {
            x[j] = xh;
        } else {
            const double ri = s[j] - omega * AMs[j];
            x[j] = xh + omega * Ms[j];
            r[j] = ri;
            acc[0] = rs[j] * ri;
            acc[1] = ri * ri;
        }
    }
    if (half) return;  // uniform per block
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = acc[0];
        part[m.nblk + blockIdx.x] = acc[1];
    }
}

__global__ void __launch_bounds__(BS) k_b_bicg_p(int op, const BCase *__restrict__ tab, BMapDev m,
                                                 double *__restrict__ p, const double *__restrict__ r,
                                                 const double *__restrict__ AMp) {
    const BSeg g = bseg(op, tab[m.cs[blockIdx.x]]);
    if (g.sc[SC_STOP] != 0.0) return;
    const double beta = g.sc[SC_BETA], omega = g.sc[SC_OMEGA];
    const int i = m.loc[blockIdx.x] * BS + threadIdx.x;
    if (i < g.n) {
        const long j = g.off + i;
        p[j] = r[j] + beta * p[j] - beta * omega * AMp[j];
    }
}

// one block per case: the case's block partials in a fixed order, then finish_update
__global__ void __launch_bounds__(BS) k_b_finish(int op, const BCase *__restrict__ tab, const int *__restrict__ first,
                                                 int nblk, const double *__restrict__ part, int fin, double rtol,
                                                 int maxit, const int *__restrict__ mask) {
    const int c = blockIdx.x;
    if (mask && !mask[c]) return;  // uniform per block
    double *sc = bseg(op, tab[c]).sc;
    if (fin == FIN_BETA && sc[SC_STOP] == STOP_HALF_PENDING) {  // the half-step update is done
        if (threadIdx.x == 0) sc[SC_STOP] = STOP_HALF;
        return;
    }
    if (fin != FIN_INIT && sc[SC_STOP] != 0.0) return;  // uniform per block
    const int ng = (fin == FIN_ALPHA || fin == FIN_SS) ? 1 : 2;  // FIN_RESTART: (r.r, b.b)
    const int b0 = first[c], b1 = first[c + 1];
    double v[2] = {0, 0};
    for (int g = 0; g < ng; g++)
        for (int k = b0 + (int)threadIdx.x; k < b1; k += BS) v[g] += part[g * nblk + k];
    block_sum<2>(v);
    if (threadIdx.x != 0) return;
    const double tot[2] = {v[0], v[1]};
    if (fin == FIN_INIT) {
        sc[SC_ITS] = 0.0;
        sc[SC_HALFK] = -1.0;
        finish_update(fin, tot, sc, 0.0, maxit);
        sc[SC_TOL] = rtol * sqrt(sc[SC_BB]);
        if (sqrt(sc[SC_RR]) <= sc[SC_TOL]) sc[SC_STOP] = STOP_CONVERGED;  // converged at the start
    } else if (fin == FIN_RESTART) {  // new recurrences from the true residual
        sc[SC_RR] = tot[0];
        sc[SC_RHO] = tot[0];
        if (sqrt(tot[0]) <= sc[SC_TOL]) sc[SC_STOP] = STOP_CONVERGED;
    } else {
        finish_update(fin, tot, sc, sc[SC_TOL], maxit);
    }
}

void launch_b_bicg_s(int op, const BCase *tab, const BMapDev &m, const double *r, const double *AMp, double *s_,
                     double *part, hipStream_t s) {
    hipLaunchKernelGGL(k_b_bicg_s, dim3(m.nblk), dim3(BS), 0, s, op, tab, m, r, AMp, s_, part);
    CUIBM_LAUNCH_CHECK();
}

void launch_b_bicg_xr(int op, const BCase *tab, const BMapDev &m, double *x, const double *Mp, const double *Ms,
                      double *r, const double *s_, const double *AMs, const double *rs, double *part,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_b_bicg_xr, dim3(m.nblk), dim3(BS), 0, s, op, tab, m, x, Mp, Ms, r, s_, AMs, rs, part);
    CUIBM_LAUNCH_CHECK();
}

void launch_b_bicg_p(int op, const BCase *tab, const BMapDev &m, double *p, const double *r, const double *AMp,
                     hipStream_t s) {
    hipLaunchKernelGGL(k_b_bicg_p, dim3(m.nblk), dim3(BS), 0, s, op, tab, m, p, r, AMp);
    CUIBM_LAUNCH_CHECK();
}

void launch_b_finish(int op, const BCase *tab, int ncases, const int *first, int nblk, const double *part, int fin,
                     double rtol, int maxit, hipStream_t s, const int *mask) {
    hipLaunchKernelGGL(k_b_finish, dim3(ncases), dim3(BS), 0, s, op, tab, first, nblk, part, fin, rtol, maxit, mask);
    CUIBM_LAUNCH_CHECK();
}

void launch_b_jacobi(long n, const double *dinv, const double *r, double *z, hipStream_t s) {
    hipLaunchKernelGGL(k_jacobi_apply, dim3(grid_for(n)), dim3(BS), 0, s, (int)n, dinv, r, z, nullptr);
    CUIBM_LAUNCH_CHECK();
}

}  // namespace cuibm
