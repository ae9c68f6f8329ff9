// Stencil form of the pressure-Poisson operator and the fused finest-level V-cycle
// kernels (see stencil.hpp for the layout and the bit-identity argument).
#include <algorithm>

#include "stencil.hpp"

namespace cuibm {

namespace {

constexpr int TX = 64, TY = 16;  // cells of a tile (one wavefront wide)
constexpr int W1 = TX + 2, H1 = TY + 2, W2 = TX + 4, H2 = TY + 4;

// blocks arrive at the 8 XCDs round-robin; tile t = xcd * per + k gives each XCD a
// contiguous band of tiles (grid = 8 * per blocks, the tail blocks idle)
__device__ __forceinline__ int xcd_band(int count) {
    const int per = (count + 7) / 8;
    return (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
}
inline int xcd_grid(int count) { return 8 * ((count + 7) / 8); }

// diagonal of row ip = (I, J): LHS2.cu:150-156 accumulation order
__device__ __forceinline__ double p5_diag(const P5Dev &A, int I, int J, int ip) {
    double t = 0;
    if (I != A.nx - 1) t += A.cE[I];
    if (I != 0) t += A.cW[I];
    if (J != A.ny - 1) t += A.cN[J];
    if (J != 0) t += A.cS[J];
    if (ip == 0) t += t;
    if (ip == A.centre) t += t;
    return t;
}

// (A x)_ip in CSR column order S, W, C, E, N; x(o) returns x at offset o from ip in
// the caller's storage, with o in {-sy, -1, 0, 1, sy}
template <class F>
__device__ __forceinline__ double p5_row(const P5Dev &A, int I, int J, double d, F x, int sy) {
    double s = 0;
    if (J != 0) s += (-A.cS[J]) * x(-sy);
    if (I != 0) s += (-A.cW[I]) * x(-1);
    s += d * x(0);
    if (I != A.nx - 1) s += (-A.cE[I]) * x(1);
    if (J != A.ny - 1) s += (-A.cN[J]) * x(sy);
    return s;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

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

// coupling magnitudes, the expressions of LHS2.cu:150-153
__global__ void k_p5_coef(int nx, int ny, const double *__restrict__ dx, const double *__restrict__ dy, double dt,
                          double *cE, double *cW, double *cN, double *cS) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nx) {
        cE[i] = (i != nx - 1) ? dt / (dx[i] * (dx[i] + dx[i + 1]) * 0.5) : 0.0;
        cW[i] = (i != 0) ? dt / (dx[i] * (dx[i] + dx[i - 1]) * 0.5) : 0.0;
    }
    if (i < ny) {
        cN[i] = (i != ny - 1) ? dt / (dy[i] * (dy[i] + dy[i + 1]) * 0.5) : 0.0;
        cS[i] = (i != 0) ? dt / (dy[i] * (dy[i] + dy[i - 1]) * 0.5) : 0.0;
    }
}

// y = A x with partials of (y . a) and (y . y) over a fixed grid of RED_GRID blocks
__global__ void __launch_bounds__(BS) k_p5_spmv_dot(P5Dev A, const double *__restrict__ x, double *__restrict__ y,
                                                    const double *__restrict__ a, int want_yy,
                                                    double *__restrict__ part) {
    const int n = A.nx * A.ny, blk = xcd_band(RED_GRID);
    double acc[2] = {0, 0};
    for (int ip = blk * BS + threadIdx.x; ip < n; ip += RED_GRID * BS) {
        const int I = ip % A.nx, J = ip / A.nx;
        const double yi = p5_row(A, I, J, p5_diag(A, I, J, ip), [&](int o) { return x[ip + o]; }, A.nx);
        if (y) y[ip] = yi;
        acc[0] += yi * a[ip];
        acc[1] += yi * yi;
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        part[blk] = acc[0];
        if (want_yy) part[RED_GRID + blk] = acc[1];
    }
}

// r = b - A x, p = rs = r, partials of (r . r) and (b . b)
__global__ void __launch_bounds__(BS) k_p5_init(P5Dev A, const double *__restrict__ x, const double *__restrict__ b,
                                                double *__restrict__ r, double *__restrict__ p,
                                                double *__restrict__ rs, double *__restrict__ part) {
    const int n = A.nx * A.ny, blk = xcd_band(RED_GRID);
    double acc[2] = {0, 0};
    for (int ip = blk * BS + threadIdx.x; ip < n; ip += RED_GRID * BS) {
        const int I = ip % A.nx, J = ip / A.nx;
        const double bi = b[ip];
        const double ri = bi - p5_row(A, I, J, p5_diag(A, I, J, ip), [&](int o) { return x[ip + o]; }, A.nx);
        r[ip] = ri;
        p[ip] = ri;
        rs[ip] = ri;
        acc[0] += ri * ri;
        acc[1] += bi * bi;
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        part[blk] = acc[0];
        part[RED_GRID + blk] = acc[1];
    }
}

__device__ __forceinline__ bool inside(const P5Dev &A, int I, int J) { return I >= 0 && I < A.nx && J >= 0 && J < A.ny; }

// pre-smoothing (two sweeps from x = 0) and residual on one tile
__global__ void __launch_bounds__(256) k_p5_pre(P5Dev A, double w, const double *__restrict__ b,
                                                double *__restrict__ x2g, double *__restrict__ rg, int tiles_x,
                                                int ntiles) {
    __shared__ double sb[H2 * W2], sw[H2 * W2], sx1[H2 * W2], sx2[H1 * W1];
    const int t = xcd_band(ntiles);
    if (t >= ntiles) return;  // uniform per block
    const int I0 = (t % tiles_x) * TX, J0 = (t / tiles_x) * TY;
    for (int k = threadIdx.x; k < H2 * W2; k += blockDim.x) {
        const int I = I0 - 2 + k % W2, J = J0 - 2 + k / W2;
        double bb = 0, wd = 0, x1 = 0;
        if (inside(A, I, J)) {
            const int ip = J * A.nx + I;
            bb = b[ip];
            wd = w * (1.0 / p5_diag(A, I, J, ip));
            x1 = wd * bb;
        }
        sb[k] = bb;
        sw[k] = wd;
        sx1[k] = x1;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < H1 * W1; k += blockDim.x) {
        const int li = k % W1 + 1, lj = k / W1 + 1;
        const int I = I0 - 2 + li, J = J0 - 2 + lj;
        double v = 0;
        if (inside(A, I, J)) {
            const int c = lj * W2 + li;
            const double s = p5_row(A, I, J, p5_diag(A, I, J, J * A.nx + I), [&](int o) { return sx1[c + o]; }, W2);
            v = sx1[c] + sw[c] * (sb[c] - s);
        }
        sx2[k] = v;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < TX * TY; k += blockDim.x) {
        const int li = k % TX, lj = k / TX;
        const int I = I0 + li, J = J0 + lj;
        if (!inside(A, I, J)) continue;
        const int ip = J * A.nx + I;
        const int c1 = (lj + 1) * W1 + li + 1, c2 = (lj + 2) * W2 + li + 2;
        const double s = p5_row(A, I, J, p5_diag(A, I, J, ip), [&](int o) { return sx2[c1 + o]; }, W1);
        x2g[ip] = sx2[c1];
        rg[ip] = sb[c2] - s;
    }
}

// x = x2 + P xc, then two weighted-Jacobi sweeps, on one tile
template <class V>
__global__ void __launch_bounds__(256) k_p5_post(P5Dev A, double w, const int *__restrict__ prp,
                                                 const int *__restrict__ pci, const V *__restrict__ pv,
                                                 const double *__restrict__ xc, const double *__restrict__ x2,
                                                 const double *__restrict__ b, double *__restrict__ xg, int tiles_x,
                                                 int ntiles) {
    __shared__ double sx[H2 * W2], sb[H1 * W1], sw[H1 * W1], sy[H1 * W1];
    const int t = xcd_band(ntiles);
    if (t >= ntiles) return;
    const int I0 = (t % tiles_x) * TX, J0 = (t / tiles_x) * TY;
    for (int k = threadIdx.x; k < H2 * W2; k += blockDim.x) {
        const int I = I0 - 2 + k % W2, J = J0 - 2 + k / W2;
        double v = 0;
        if (inside(A, I, J)) {
            const int ip = J * A.nx + I;
            double s = 0;
            for (int q = prp[ip]; q < prp[ip + 1]; q++) s += (double)pv[q] * xc[pci[q]];
            v = x2[ip] + s;
        }
        sx[k] = v;
    }
    for (int k = threadIdx.x; k < H1 * W1; k += blockDim.x) {
        const int I = I0 - 1 + k % W1, J = J0 - 1 + k / W1;
        double bb = 0, wd = 0;
        if (inside(A, I, J)) {
            const int ip = J * A.nx + I;
            bb = b[ip];
            wd = w * (1.0 / p5_diag(A, I, J, ip));
        }
        sb[k] = bb;
        sw[k] = wd;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < H1 * W1; k += blockDim.x) {
        const int li = k % W1 + 1, lj = k / W1 + 1;
        const int I = I0 - 2 + li, J = J0 - 2 + lj;
        double v = 0;
        if (inside(A, I, J)) {
            const int c = lj * W2 + li;
            const double s = p5_row(A, I, J, p5_diag(A, I, J, J * A.nx + I), [&](int o) { return sx[c + o]; }, W2);
            v = sx[c] + sw[k] * (sb[k] - s);
        }
        sy[k] = v;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < TX * TY; k += blockDim.x) {
        const int li = k % TX, lj = k / TX;
        const int I = I0 + li, J = J0 + lj;
        if (!inside(A, I, J)) continue;
        const int ip = J * A.nx + I;
        const int c1 = (lj + 1) * W1 + li + 1;
        const double s = p5_row(A, I, J, p5_diag(A, I, J, ip), [&](int o) { return sy[c1 + o]; }, W1);
        xg[ip] = sy[c1] + sw[c1] * (sb[c1] - s);
    }
}

// one sweep xn = x + omega D^-1 (b - A x), a row per thread
__global__ void __launch_bounds__(BS) k_p5_sweep(P5Dev A, double w, const double *__restrict__ x,
                                                 const double *__restrict__ b, double *__restrict__ xn, int nblk) {
    const int blk = xcd_band(nblk);
    const int ip = blk * BS + threadIdx.x, n = A.nx * A.ny;
    if (blk >= nblk || ip >= n) return;
    const int I = ip % A.nx, J = ip / A.nx;
    const double d = p5_diag(A, I, J, ip);
    const double s = p5_row(A, I, J, d, [&](int o) { return x[ip + o]; }, A.nx);
    xn[ip] = x[ip] + w * (1.0 / d) * (b[ip] - s);
}

__global__ void __launch_bounds__(BS) k_p5_apply(P5Dev A, const double *__restrict__ x, double *__restrict__ y,
                                                 int nblk) {
    const int blk = xcd_band(nblk);
    const int ip = blk * BS + threadIdx.x, n = A.nx * A.ny;
    if (blk >= nblk || ip >= n) return;
    const int I = ip % A.nx, J = ip / A.nx;
    y[ip] = p5_row(A, I, J, p5_diag(A, I, J, ip), [&](int o) { return x[ip + o]; }, A.nx);
}

}  // namespace

PoissonStencil::PoissonStencil(const GridDev &g, double dt, bool centre_pin, hipStream_t s) : nx_(g.nx), ny_(g.ny) {
    // LHS2.cu:169-172: the centre pin applies to an interior cell only
    const int c = (ny_ / 2) * nx_ + nx_ / 2, cI = c % nx_, cJ = c / nx_;
    const bool interior = cI > 0 && cI < nx_ - 1 && cJ > 0 && cJ < ny_ - 1;
    centre_ = (centre_pin && interior && c != 0) ? c : -1;
    for (auto *a : {&cE_, &cW_}) a->alloc(nx_);
    for (auto *a : {&cN_, &cS_}) a->alloc(ny_);
    const int m = std::max(nx_, ny_);
    hipLaunchKernelGGL(k_p5_coef, dim3((m + BS - 1) / BS), dim3(BS), 0, s, nx_, ny_, g.dx, g.dy, dt, cE_.p(), cW_.p(),
                       cN_.p(), cS_.p());
    CUIBM_LAUNCH_CHECK();
}

void PoissonStencil::spmv_dot(const double *x, double *y, const double *a, bool want_yy, double *part,
                              hipStream_t s) const {
    hipLaunchKernelGGL(k_p5_spmv_dot, dim3(RED_GRID), dim3(BS), 0, s, dev(), x, y, a, (int)want_yy, part);
    CUIBM_LAUNCH_CHECK();
}

void PoissonStencil::init(const double *x, const double *b, double *r, double *p, double *rs, double *part,
                          hipStream_t s) const {
    hipLaunchKernelGGL(k_p5_init, dim3(RED_GRID), dim3(BS), 0, s, dev(), x, b, r, p, rs, part);
    CUIBM_LAUNCH_CHECK();
}

void PoissonStencil::apply(const double *x, double *y, hipStream_t s) const {
    const int nblk = (n() + BS - 1) / BS;
    hipLaunchKernelGGL(k_p5_apply, dim3(xcd_grid(nblk)), dim3(BS), 0, s, dev(), x, y, nblk);
    CUIBM_LAUNCH_CHECK();
}

static void tiles(const P5Dev &A, int &tx, int &nt) {
    tx = (A.nx + TX - 1) / TX;
    nt = tx * ((A.ny + TY - 1) / TY);
}

void launch_p5_pre(const P5Dev &A, double omega, const double *b, double *x2, double *r, hipStream_t s) {
    int tx, nt;
    tiles(A, tx, nt);
    hipLaunchKernelGGL(k_p5_pre, dim3(xcd_grid(nt)), dim3(256), 0, s, A, omega, b, x2, r, tx, nt);
    CUIBM_LAUNCH_CHECK();
}

template <class V>
void launch_p5_post(const P5Dev &A, double omega, const DCsr<V> &P, const double *xc, const double *x2,
                    const double *b, double *x, hipStream_t s) {
    int tx, nt;
    tiles(A, tx, nt);
    hipLaunchKernelGGL(k_p5_post<V>, dim3(xcd_grid(nt)), dim3(256), 0, s, A, omega, P.rp.p(), P.ci.p(), P.v.p(), xc,
                       x2, b, x, tx, nt);
    CUIBM_LAUNCH_CHECK();
}
template void launch_p5_post<double>(const P5Dev &, double, const DCsr<double> &, const double *, const double *,
                                     const double *, double *, hipStream_t);
template void launch_p5_post<float>(const P5Dev &, double, const DCsr<float> &, const double *, const double *,
                                    const double *, double *, hipStream_t);

void launch_p5_sweep(const P5Dev &A, double omega, const double *x, const double *b, double *xn, hipStream_t s) {
    const int nblk = (A.nx * A.ny + BS - 1) / BS;
    hipLaunchKernelGGL(k_p5_sweep, dim3(xcd_grid(nblk)), dim3(BS), 0, s, A, omega, x, b, xn, nblk);
    CUIBM_LAUNCH_CHECK();
}

}  // namespace cuibm
