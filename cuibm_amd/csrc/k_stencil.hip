// Stencil form of the pressure-Poisson operator and the fused finest-level V-cycle
// kernels (see stencil.hpp for the layout and the bit-identity argument).
#include <hip/hip_ext.h>

#include <algorithm>

#include "batch.hpp"
#include "stencil.hpp"

namespace cuibm {

namespace {

constexpr int TX = 64, TY = 8;  // cells of a tile (one wavefront wide; 64x8 measured best of 32-128 x 4-32)
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
    const double stopv = A.stop ? *A.stop : 0.0;  // the Krylov solve's stop code, tested after the loads
    const int n = A.nx * A.ny, blk = xcd_band(RED_GRID);
    double acc[2] = {0, 0};
    bool tested = false;
    for (int ip = blk * BS + threadIdx.x; ip < n; ip += RED_GRID * BS) {
        const int I = ip % A.nx, J = ip / A.nx;
        const double yi = p5_row(A, I, J, p5_diag(A, I, J, ip), [&](int o) { return x[ip + o]; }, A.nx);
        if (!tested) {
            if (stopv != 0.0) return;  // every thread returns here or below: no barrier is reached
            tested = true;
        }
        if (y) y[ip] = yi;
        acc[0] += yi * a[ip];
        acc[1] += yi * yi;
    }
    if (stopv != 0.0) return;
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

// D^-1 of the operator (the same 1.0 / diagonal a CSR sweep would use)
__global__ void k_p5_dinv(P5Dev A, double *__restrict__ dinv) {
    const int n = A.nx * A.ny;
    for (int ip = blockIdx.x * blockDim.x + threadIdx.x; ip < n; ip += gridDim.x * blockDim.x)
        dinv[ip] = 1.0 / p5_diag(A, ip % A.nx, ip / A.nx, ip);
}

// the coupling magnitudes of the columns and rows one tile (with a halo of 2) touches,
// staged in LDS in the compute type T; li, lj below are halo-2 tile coordinates
// (I = I0 - 2 + li)
template <class T>
struct TileCoef {
    T e[W2], w[W2], n[H2], s[H2];
};

template <class T>
__device__ __forceinline__ void load_coef(const P5Dev &A, int I0, int J0, TileCoef<T> &c) {
    for (int k = threadIdx.x; k < W2; k += blockDim.x) {
        const int I = I0 - 2 + k;
        const bool in = I >= 0 && I < A.nx;
        c.e[k] = in ? (T)A.cE[I] : T(0);
        c.w[k] = in ? (T)A.cW[I] : T(0);
    }
    for (int k = threadIdx.x; k < H2; k += blockDim.x) {
        const int J = J0 - 2 + k;
        const bool in = J >= 0 && J < A.ny;
        c.n[k] = in ? (T)A.cN[J] : T(0);
        c.s[k] = in ? (T)A.cS[J] : T(0);
    }
}

template <class T>
__device__ __forceinline__ T t_diag(const P5Dev &A, const TileCoef<T> &c, int li, int lj, int I, int J, int ip) {
    T t = 0;
    if (I != A.nx - 1) t += c.e[li];
    if (I != 0) t += c.w[li];
    if (J != A.ny - 1) t += c.n[lj];
    if (J != 0) t += c.s[lj];
    if (ip == 0) t += t;
    if (ip == A.centre) t += t;
    return t;
}

template <class T, class F>
__device__ __forceinline__ T t_row(const P5Dev &A, const TileCoef<T> &c, int li, int lj, int I, int J, T d, F x,
                                   int sy) {
    T s = 0;
    if (J != 0) s += (-c.s[lj]) * x(-sy);
    if (I != 0) s += (-c.w[li]) * x(-1);
    s += d * x(0);
    if (I != A.nx - 1) s += (-c.e[li]) * x(1);
    if (J != A.ny - 1) s += (-c.n[lj]) * x(sy);
    return s;
}

// D^-1 in the compute type: the FP64 array, or its FP32 copy for the FP32 V-cycle
template <class T>
__device__ __forceinline__ const T *dinv_of(const P5Dev &A);
template <>
__device__ __forceinline__ const double *dinv_of<double>(const P5Dev &A) { return A.dinv; }
template <>
__device__ __forceinline__ const float *dinv_of<float>(const P5Dev &A) { return A.dinv32; }

constexpr int N2 = H2 * W2, N1 = H1 * W1, R2 = (N2 + 255) / 256, R1 = (N1 + 255) / 256;

// pre-smoothing (two sweeps from x = 0) and residual on one tile, in the compute type T
// (b is the FP64 Krylov vector; x2 and r are level vectors of type T); all global loads
// of a thread are issued before the first is used
template <class T>
__device__ __forceinline__ void p5_pre_tile(const P5Dev &A, T w, const double *__restrict__ b, T *__restrict__ x2g,
                                            T *__restrict__ rg, int tiles_x, int t) {
    const double stopv = A.stop ? *A.stop : 0.0;  // the Krylov solve's stop code, tested after the loads
    __shared__ TileCoef<T> cf;
    __shared__ T sb[N2], swd[N2], sx2[N1];
    const int I0 = (t % tiles_x) * TX, J0 = (t / tiles_x) * TY;
    load_coef(A, I0, J0, cf);
    const T *dinv = dinv_of<T>(A);
    // the deflation constant c, wanted only where a pin lies in the tile or its halo
    double cdefl = 0.0;
    if (A.defl) {
        auto near = [&](int ip) {
            const int I = ip % A.nx, J = ip / A.nx;
            return I >= I0 - 2 && I < I0 + TX + 2 && J >= J0 - 2 && J < J0 + TY + 2;
        };
        if (near(0) || (A.centre > 0 && near(A.centre))) {
            if (A.defl_part) {  // uniform per block: the partials in a fixed order, as one finish block would
                __shared__ double cs;
                double acc[1] = {0};
                for (int k = threadIdx.x; k < RED_GRID; k += 256) acc[0] += A.defl_part[k];
                block_sum<1>(acc);
                if (threadIdx.x == 0) {
                    cs = acc[0] * A.defl_inv_wa;
                    if (t == 0) A.defl[0] = cs;  // for the post-smoothing kernel
                }
                __syncthreads();
                cdefl = cs;
            } else {
                cdefl = A.defl[0];
            }
        }
    }
    T bv[R2], dv[R2];
#pragma unroll
    for (int m = 0; m < R2; m++) {
        const int k = threadIdx.x + 256 * m;
        const int I = I0 - 2 + k % W2, J = J0 - 2 + k / W2;
        const bool in = k < N2 && inside(A, I, J);
        const int ip = in ? J * A.nx + I : 0;
        double bb = in ? b[ip] : 0.0;
        if (A.defl && in && (ip == 0 || ip == A.centre)) bb -= cdefl * (ip == 0 ? A.defl_a0 : A.defl_ac);
        bv[m] = (T)bb;
        dv[m] = in ? dinv[ip] : T(0);
    }
    if (stopv != 0.0) return;  // uniform across the block, before the first barrier
#pragma unroll
    for (int m = 0; m < R2; m++) {
        const int k = threadIdx.x + 256 * m;
        if (k < N2) {
            sb[k] = bv[m];
            swd[k] = w * dv[m];
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < N1; k += blockDim.x) {
        const int li = k % W1 + 1, lj = k / W1 + 1;
        const int I = I0 - 2 + li, J = J0 - 2 + lj;
        T v = 0;
        if (inside(A, I, J)) {
            const int c = lj * W2 + li;
            const T d = t_diag(A, cf, li, lj, I, J, J * A.nx + I);
            const T s = t_row(A, cf, li, lj, I, J, d, [&](int o) { return swd[c + o] * sb[c + o]; }, W2);
            v = swd[c] * sb[c] + swd[c] * (sb[c] - s);
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
        const T d = t_diag(A, cf, li + 2, lj + 2, I, J, ip);
        const T s = t_row(A, cf, li + 2, lj + 2, I, J, d, [&](int o) { return sx2[c1 + o]; }, W1);
        x2g[ip] = sx2[c1];
        rg[ip] = sb[c2] - s;
    }
}

template <class T>
__global__ void __launch_bounds__(256) k_p5_pre(P5Dev A, T w, const double *__restrict__ b, T *__restrict__ x2g,
                                                T *__restrict__ rg, int tiles_x, int ntiles) {
    const int t = xcd_band(ntiles);
    if (t >= ntiles) return;  // uniform per block
    p5_pre_tile<T>(A, w, b, x2g, rg, tiles_x, t);
}

// two weighted-Jacobi sweeps on one tile in the compute type T: x (halo 2, a level
// vector) -> y (halo 1) -> xo (tile, the FP64 Krylov vector); b is the FP64 Krylov vector
template <class T>
__device__ __forceinline__ void p5_sweep2_tile(const P5Dev &A, T w, const T *__restrict__ x,
                                               const double *__restrict__ b, double *__restrict__ xo, int tiles_x,
                                               int t) {
    const double stopv = A.stop ? *A.stop : 0.0;  // the Krylov solve's stop code, tested after the loads
    __shared__ TileCoef<T> cf;
    __shared__ T sx[N2], sb[N1], swd[N1];
    const int I0 = (t % tiles_x) * TX, J0 = (t / tiles_x) * TY;
    load_coef(A, I0, J0, cf);
    const T *dinv = dinv_of<T>(A);
    T xv[R2], bv[R1], dv[R1];
#pragma unroll
    for (int m = 0; m < R2; m++) {
        const int k = threadIdx.x + 256 * m;
        const int I = I0 - 2 + k % W2, J = J0 - 2 + k / W2;
        const bool in = k < N2 && inside(A, I, J);
        xv[m] = in ? x[J * A.nx + I] : T(0);
    }
#pragma unroll
    for (int m = 0; m < R1; m++) {
        const int k = threadIdx.x + 256 * m;
        const int I = I0 - 1 + k % W1, J = J0 - 1 + k / W1;
        const bool in = k < N1 && inside(A, I, J);
        const int ip = in ? J * A.nx + I : 0;
        double bb = in ? b[ip] : 0.0;
        if (A.defl && in && (ip == 0 || ip == A.centre)) bb -= A.defl[0] * (ip == 0 ? A.defl_a0 : A.defl_ac);
        bv[m] = (T)bb;
        dv[m] = in ? dinv[ip] : T(0);
    }
    const double cdefl = A.defl ? A.defl[0] : 0.0;
    if (stopv != 0.0) return;  // uniform across the block, before the first barrier
#pragma unroll
    for (int m = 0; m < R2; m++) {
        const int k = threadIdx.x + 256 * m;
        if (k < N2) sx[k] = xv[m];
    }
#pragma unroll
    for (int m = 0; m < R1; m++) {
        const int k = threadIdx.x + 256 * m;
        if (k < N1) {
            sb[k] = bv[m];
            swd[k] = w * dv[m];
        }
    }
    __syncthreads();
    // first sweep on the halo-1 region, kept in registers until every read of sx is done
    T yv[R1];
#pragma unroll
    for (int m = 0; m < R1; m++) {
        const int k = threadIdx.x + 256 * m;
        T v = 0;
        if (k < N1) {
            const int li = k % W1 + 1, lj = k / W1 + 1;
            const int I = I0 - 2 + li, J = J0 - 2 + lj;
            if (inside(A, I, J)) {
                const int c = lj * W2 + li;
                const T d = t_diag(A, cf, li, lj, I, J, J * A.nx + I);
                const T s = t_row(A, cf, li, lj, I, J, d, [&](int o) { return sx[c + o]; }, W2);
                v = sx[c] + swd[k] * (sb[k] - s);
            }
        }
        yv[m] = v;
    }
    __syncthreads();
    T *sy = sx;  // reuse: the halo-1 iterate in halo-1 layout
#pragma unroll
    for (int m = 0; m < R1; m++) {
        const int k = threadIdx.x + 256 * m;
        if (k < N1) sy[k] = yv[m];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < TX * TY; k += blockDim.x) {
        const int li = k % TX, lj = k / TX;
        const int I = I0 + li, J = J0 + lj;
        if (!inside(A, I, J)) continue;
        const int ip = J * A.nx + I;
        const int c1 = (lj + 1) * W1 + li + 1;
        const T d = t_diag(A, cf, li + 2, lj + 2, I, J, ip);
        const T s = t_row(A, cf, li + 2, lj + 2, I, J, d, [&](int o) { return sy[c1 + o]; }, W1);
        xo[ip] = (double)(sy[c1] + swd[c1] * (sb[c1] - s)) + cdefl;
    }
}

template <class T>
__global__ void __launch_bounds__(256) k_p5_sweep2(P5Dev A, T w, const T *__restrict__ x,
                                                   const double *__restrict__ b, double *__restrict__ xo, int tiles_x,
                                                   int ntiles) {
    const int t = xcd_band(ntiles);
    if (t >= ntiles) return;
    p5_sweep2_tile<T>(A, w, x, b, xo, tiles_x, t);
}

// partial sums of w_i r_i, w_i = dx[I] dy[J], over a fixed grid of RED_GRID blocks
__global__ void __launch_bounds__(BS) k_p5_wdot(P5Dev A, const double *__restrict__ r, double *__restrict__ part) {
    const double stopv = A.stop ? *A.stop : 0.0;
    if (stopv != 0.0) return;
    const int n = A.nx * A.ny, blk = xcd_band(RED_GRID);
    double acc[1] = {0};
    for (int ip = blk * BS + threadIdx.x; ip < n; ip += RED_GRID * BS)
        acc[0] += (A.dx[ip % A.nx] * A.dy[ip / A.nx]) * r[ip];
    block_sum<1>(acc);
    if (threadIdx.x == 0) part[blk] = acc[0];
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
    xn[ip] = x[ip] + w * A.dinv[ip] * (b[ip] - s);
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

PoissonStencil::PoissonStencil(const GridDev &g, double dt, bool centre_pin, hipStream_t s)
    : nx_(g.nx), ny_(g.ny), dx_(g.dx), dy_(g.dy) {
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
    dinv_.alloc((size_t)nx_ * ny_);
    P5Dev d = dev();
    d.dinv = nullptr;
    hipLaunchKernelGGL(k_p5_dinv, dim3(std::min((nx_ * ny_ + BS - 1) / BS, 65536)), dim3(BS), 0, s, d, dinv_.p());
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

// ---- velocity operator (LHS1) in stencil form ---------------------------------------

// the coupling magnitudes and diagonal terms of LHS1.cu:24-351, one thread per I or J:
// u block: E/W from dx[I], dx[I+1] for I < nx-1, N/S from dy[J], dy[J+-1] with the
// doubled-width ghost term at J = ny-1 (N) and J = 0 (S); v block: E/W from dx[I],
// dx[I+-1] with the ghost term at I = nx-1 (E) and I = 0 (W), N/S from dy[J], dy[J+1]
__global__ void k_v5_coef(int nx, int ny, const double *__restrict__ dx, const double *__restrict__ dy, double nu,
                          double dt, double *__restrict__ c, int m) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double *uoE = c, *uoW = c + m, *udE = c + 2 * m, *udW = c + 3 * m;
    double *uoN = c + 4 * m, *uoS = c + 5 * m, *udN = c + 6 * m, *udS = c + 7 * m;
    double *voE = c + 8 * m, *voW = c + 9 * m, *vdE = c + 10 * m, *vdW = c + 11 * m;
    double *voN = c + 12 * m, *voS = c + 13 * m, *vdN = c + 14 * m, *vdS = c + 15 * m;
    if (k < nx - 1) {  // u, per I
        const int I = k;
        const double ce = 0.5 * dt * nu * (1 / (dx[I + 1] * (dx[I + 1] + dx[I]) * 0.5));
        const double cw = 0.5 * dt * nu * (1 / (dx[I] * (dx[I + 1] + dx[I]) * 0.5));
        uoE[I] = ce; udE[I] = ce;
        uoW[I] = cw; udW[I] = cw;
    }
    if (k < ny) {  // u, per J
        const int J = k;
        const double nN = J != ny - 1 ? 0.5 * dt * nu * (1 / (dy[J] * (dy[J + 1] + dy[J]) * 0.5)) : 0.0;
        const double nS = J != 0 ? 0.5 * dt * nu * (1 / (dy[J] * (dy[J - 1] + dy[J]) * 0.5)) : 0.0;
        const double gh = 0.5 * dt * nu * (1 / (dy[J] * (dy[J] + dy[J]) * 0.5));
        uoN[J] = nN; udN[J] = J != ny - 1 ? nN : gh;
        uoS[J] = nS; udS[J] = J != 0 ? nS : gh;
    }
    if (k < nx) {  // v, per I
        const int I = k;
        const double nE = I != nx - 1 ? 0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I + 1]) * 0.5)) : 0.0;
        const double nW = I != 0 ? 0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I - 1]) * 0.5)) : 0.0;
        const double gh = 0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I]) * 0.5));
        voE[I] = nE; vdE[I] = I != nx - 1 ? nE : gh;
        voW[I] = nW; vdW[I] = I != 0 ? nW : gh;
    }
    if (k < ny - 1) {  // v, per J
        const int J = k;
        const double cn = 0.5 * dt * nu * (1 / (dy[J + 1] * (dy[J] + dy[J + 1]) * 0.5));
        const double cs = 0.5 * dt * nu * (1 / (dy[J] * (dy[J] + dy[J + 1]) * 0.5));
        voN[J] = cn; vdN[J] = cn;
        voS[J] = cs; vdS[J] = cs;
    }
}

// (LHS1 x)_i for the packed index i, in the CSR column order S, W, C, E, N
__device__ __forceinline__ double v5_row(const V5Dev &A, int i, const double *__restrict__ x) {
    const V5Block &B = i < A.v.off0 ? A.u : A.v;
    const int k = i - B.off0, I = k % B.ni, J = k / B.ni;
    const double d = (((1.0 + B.dE[I]) + B.dW[I]) + B.dN[J]) + B.dS[J];
    double s = 0;
    if (J != 0) s += (-B.oS[J]) * x[i - B.ni];
    if (I != 0) s += (-B.oW[I]) * x[i - 1];
    s += d * x[i];
    if (I != B.ni - 1) s += (-B.oE[I]) * x[i + 1];
    if (J != B.nj - 1) s += (-B.oN[J]) * x[i + B.ni];
    return s;
}

// y = A x with block partials of (y . a) and (y . y): the loop, grid and reduction of
// the CSR k_spmv_dot, so the partials are the CSR path's
__global__ void __launch_bounds__(BS) k_v5_spmv_dot(V5Dev A, int n, const double *__restrict__ x,
                                                    double *__restrict__ y, const double *__restrict__ a, int want_yy,
                                                    double *__restrict__ part) {
    const double stopv = A.stop ? *A.stop : 0.0;  // the Krylov solve's stop code, tested after the loads
    double acc[2] = {0, 0};
    bool tested = false;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double yi = v5_row(A, i, x);
        if (!tested) {
            if (stopv != 0.0) return;  // every thread returns here or below: no barrier is reached
            tested = true;
        }
        if (y) y[i] = yi;
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

// r = b - A x, p = rs = r, partials of (r . r) and (b . b): as the CSR k_bicg_init
__global__ void __launch_bounds__(BS) k_v5_init(V5Dev A, int n, const double *__restrict__ x,
                                                const double *__restrict__ b, double *__restrict__ r,
                                                double *__restrict__ p, double *__restrict__ rs,
                                                double *__restrict__ part) {
    double acc[2] = {0, 0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double bi = b[i];
        const double ri = bi - v5_row(A, i, x);
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

VelocityStencil::VelocityStencil(const GridDev &g, double nu, double dt, hipStream_t s)
    : nx_(g.nx), ny_(g.ny), nu_((g.nx - 1) * g.ny), nv_(g.nx * (g.ny - 1)) {
    const int m = std::max(nx_, ny_);
    c_.alloc((size_t)16 * m);
    hipLaunchKernelGGL(k_v5_coef, dim3((m + BS - 1) / BS), dim3(BS), 0, s, nx_, ny_, g.dx, g.dy, nu, dt, c_.p(), m);
    CUIBM_LAUNCH_CHECK();
}

V5Dev VelocityStencil::dev() const {
    const int m = std::max(nx_, ny_);
    const double *c = c_.p();
    V5Dev d;
    d.u = V5Block{nx_ - 1, ny_, 0, c, c + m, c + 4 * m, c + 5 * m, c + 2 * m, c + 3 * m, c + 6 * m, c + 7 * m};
    d.v = V5Block{nx_, ny_ - 1, nu_, c + 8 * m, c + 9 * m, c + 12 * m, c + 13 * m, c + 10 * m, c + 11 * m, c + 14 * m,
                  c + 15 * m};
    d.stop = stop_;
    return d;
}

void VelocityStencil::spmv_dot(const double *x, double *y, const double *a, bool want_yy, double *part,
                               hipStream_t s) const {
    hipLaunchKernelGGL(k_v5_spmv_dot, dim3(RED_GRID), dim3(BS), 0, s, dev(), n(), x, y, a, (int)want_yy, part);
    CUIBM_LAUNCH_CHECK();
}

void VelocityStencil::init(const double *x, const double *b, double *r, double *p, double *rs, double *part,
                           hipStream_t s) const {
    hipLaunchKernelGGL(k_v5_init, dim3(RED_GRID), dim3(BS), 0, s, dev(), n(), x, b, r, p, rs, part);
    CUIBM_LAUNCH_CHECK();
}

static void tiles(const P5Dev &A, int &tx, int &nt) {
    tx = (A.nx + TX - 1) / TX;
    nt = tx * ((A.ny + TY - 1) / TY);
}

template <class T>
static void p5_pre(const P5Dev &A, T omega, const double *b, T *x2, T *r, hipStream_t s, KernelProbe *probe) {
    int tx, nt;
    tiles(A, tx, nt);
    if (probe && probe->enabled()) {
        hipEvent_t e0, e1;
        probe->pair(e0, e1);
        hipExtLaunchKernelGGL(k_p5_pre<T>, dim3(xcd_grid(nt)), dim3(256), 0, s, e0, e1, 0, A, omega, b, x2, r, tx,
                              nt);
    } else {
        hipLaunchKernelGGL(k_p5_pre<T>, dim3(xcd_grid(nt)), dim3(256), 0, s, A, omega, b, x2, r, tx, nt);
    }
    CUIBM_LAUNCH_CHECK();
}

template <class T>
static void p5_sweep2(const P5Dev &A, T omega, const T *x, const double *b, double *xo, hipStream_t s) {
    int tx, nt;
    tiles(A, tx, nt);
    hipLaunchKernelGGL(k_p5_sweep2<T>, dim3(xcd_grid(nt)), dim3(256), 0, s, A, omega, x, b, xo, tx, nt);
    CUIBM_LAUNCH_CHECK();
}

void launch_p5_pre(const P5Dev &A, double omega, const double *b, double *x2, double *r, hipStream_t s,
                   KernelProbe *probe) {
    p5_pre<double>(A, omega, b, x2, r, s, probe);
}

void launch_p5_pre(const P5Dev &A, float omega, const double *b, float *x2, float *r, hipStream_t s,
                   KernelProbe *probe) {
    if (!A.dinv32) throw Error("FP32 finest-level smoother without its FP32 D^-1", ERR_ARGUMENT);
    p5_pre<float>(A, omega, b, x2, r, s, probe);
}

void launch_p5_sweep2(const P5Dev &A, double omega, const double *x, const double *b, double *xo, hipStream_t s) {
    p5_sweep2<double>(A, omega, x, b, xo, s);
}

void launch_p5_sweep2(const P5Dev &A, float omega, const float *x, const double *b, double *xo, hipStream_t s) {
    if (!A.dinv32) throw Error("FP32 finest-level smoother without its FP32 D^-1", ERR_ARGUMENT);
    p5_sweep2<float>(A, omega, x, b, xo, s);
}

void launch_p5_wdot(const P5Dev &A, const double *r, double *part, hipStream_t s) {
    if (!A.dx || !A.dy) throw Error("deflation needs the cell widths of the stencil operator", ERR_ARGUMENT);
    hipLaunchKernelGGL(k_p5_wdot, dim3(RED_GRID), dim3(BS), 0, s, A, r, part);
    CUIBM_LAUNCH_CHECK();
}

void launch_p5_sweep(const P5Dev &A, double omega, const double *x, const double *b, double *xn, hipStream_t s) {
    const int nblk = (A.nx * A.ny + BS - 1) / BS;
    hipLaunchKernelGGL(k_p5_sweep, dim3(xcd_grid(nblk)), dim3(BS), 0, s, A, omega, x, b, xn, nblk);
    CUIBM_LAUNCH_CHECK();
}

// ---- lockstep batches: the stencil operators and the finest AMG level of every case ----

int p5_tiles(int nx, int ny) { return ((nx + TX - 1) / TX) * ((ny + TY - 1) / TY); }

template <int OP>
__device__ __forceinline__ double b_row(const BCase &C, int i, const double *__restrict__ xc) {
    if (OP == BOP_VEL) return v5_row(C.v5, i, xc);
    const P5Dev &A = C.p5;
    const int I = i % A.nx, J = i / A.nx;
    return p5_row(A, I, J, p5_diag(A, I, J, i), [&](int o) { return xc[i + o]; }, A.nx);
}

template <int OP>
__global__ void __launch_bounds__(BS) k_b_init(const BCase *__restrict__ tab, BMapDev m, const double *__restrict__ x,
                                               const double *__restrict__ b, double *__restrict__ r,
                                               double *__restrict__ p, double *__restrict__ rs,
                                               double *__restrict__ part, const int *__restrict__ mask) {
    if (mask && !mask[m.cs[blockIdx.x]]) return;  // uniform per block
    const BCase &C = tab[m.cs[blockIdx.x]];
    const long off = OP == BOP_VEL ? C.offUV : C.offP;
    const int n = OP == BOP_VEL ? C.numUV : C.numP;
    const int i = m.loc[blockIdx.x] * BS + threadIdx.x;
    double acc[2] = {0, 0};
    if (i < n) {
        const double bi = b[off + i];
        const double ri = bi - b_row<OP>(C, i, x + off);
        r[off + i] = ri;
        p[off + i] = ri;
        rs[off + i] = ri;
        acc[0] = ri * ri;
        acc[1] = bi * bi;
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = acc[0];
        part[m.nblk + blockIdx.x] = acc[1];
    }
}

template <int OP>
__global__ void __launch_bounds__(BS) k_b_spmv_dot(const BCase *__restrict__ tab, BMapDev m,
                                                   const double *__restrict__ x, double *__restrict__ y,
                                                   const double *__restrict__ a, int want_yy,
                                                   double *__restrict__ part) {
    const BCase &C = tab[m.cs[blockIdx.x]];
    const double *sc = OP == BOP_VEL ? C.scV : C.scP;
    if (sc[SC_STOP] != 0.0) return;  // uniform per block
    const long off = OP == BOP_VEL ? C.offUV : C.offP;
    const int n = OP == BOP_VEL ? C.numUV : C.numP;
    const int i = m.loc[blockIdx.x] * BS + threadIdx.x;
    double acc[2] = {0, 0};
    if (i < n) {
        const double yi = b_row<OP>(C, i, x + off);
        y[off + i] = yi;
        acc[0] = yi * a[off + i];
        if (want_yy) acc[1] = yi * yi;
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = acc[0];
        part[m.nblk + blockIdx.x] = acc[1];
    }
}

void launch_b_init(int op, const BCase *tab, const BMapDev &m, const double *x, const double *b, double *r, double *p,
                   double *rs, double *part, hipStream_t s, const int *mask) {
    if (op == BOP_VEL)
        hipLaunchKernelGGL(k_b_init<BOP_VEL>, dim3(m.nblk), dim3(BS), 0, s, tab, m, x, b, r, p, rs, part, mask);
    else
        hipLaunchKernelGGL(k_b_init<BOP_POIS>, dim3(m.nblk), dim3(BS), 0, s, tab, m, x, b, r, p, rs, part, mask);
    CUIBM_LAUNCH_CHECK();
}

void launch_b_spmv_dot(int op, const BCase *tab, const BMapDev &m, const double *x, double *y, const double *a,
                       bool want_yy, double *part, hipStream_t s) {
    if (op == BOP_VEL)
        hipLaunchKernelGGL(k_b_spmv_dot<BOP_VEL>, dim3(m.nblk), dim3(BS), 0, s, tab, m, x, y, a, (int)want_yy, part);
    else
        hipLaunchKernelGGL(k_b_spmv_dot<BOP_POIS>, dim3(m.nblk), dim3(BS), 0, s, tab, m, x, y, a, (int)want_yy, part);
    CUIBM_LAUNCH_CHECK();
}

// deflation of the pinned constant mode per case (the FP32 V-cycle)
__global__ void __launch_bounds__(BS) k_b_wdot(const BCase *__restrict__ tab, BMapDev m, const double *__restrict__ r,
                                               double *__restrict__ part) {
    const BCase &C = tab[m.cs[blockIdx.x]];
    if (C.p5.stop && *C.p5.stop != 0.0) return;  // uniform per block
    const int ip = m.loc[blockIdx.x] * BS + threadIdx.x;
    double acc[1] = {0};
    if (ip < C.numP) acc[0] = (C.g.dx[ip % C.g.nx] * C.g.dy[ip / C.g.nx]) * r[C.offP + ip];
    block_sum<1>(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

__global__ void __launch_bounds__(BS) k_b_wdot_finish(const BCase *__restrict__ tab, const int *__restrict__ first,
                                                      const double *__restrict__ part,
                                                      const double *__restrict__ inv_wa, double *__restrict__ defl) {
    const int c = blockIdx.x;
    const BCase &C = tab[c];
    if (C.p5.stop && *C.p5.stop != 0.0) return;  // uniform per block
    double acc[1] = {0};
    for (int k = first[c] + (int)threadIdx.x; k < first[c + 1]; k += BS) acc[0] += part[k];
    block_sum<1>(acc);
    if (threadIdx.x == 0) defl[c] = acc[0] * inv_wa[c];
}

void launch_b_wdot(const BCase *tab, const BMapDev &m, const double *r, double *part, hipStream_t s) {
    hipLaunchKernelGGL(k_b_wdot, dim3(m.nblk), dim3(BS), 0, s, tab, m, r, part);
    CUIBM_LAUNCH_CHECK();
}

void launch_b_wdot_finish(const BCase *tab, int ncases, const int *first, const double *part, const double *inv_wa,
                          double *defl, hipStream_t s) {
    hipLaunchKernelGGL(k_b_wdot_finish, dim3(ncases), dim3(BS), 0, s, tab, first, part, inv_wa, defl);
    CUIBM_LAUNCH_CHECK();
}

// the finest-level tile kernels, one block per tile of one case
template <class T>
__global__ void __launch_bounds__(256) k_b_p5_pre(const BCase *__restrict__ tab, BMapDev m, T w,
                                                  const double *__restrict__ b, T *__restrict__ x2g,
                                                  T *__restrict__ rg) {
    const BCase &C = tab[m.cs[blockIdx.x]];
    const P5Dev A = C.p5;
    p5_pre_tile<T>(A, w, b + C.offP, x2g + C.offP, rg + C.offP, (A.nx + TX - 1) / TX, m.loc[blockIdx.x]);
}

template <class T>
__global__ void __launch_bounds__(256) k_b_p5_sweep2(const BCase *__restrict__ tab, BMapDev m, T w,
                                                     const T *__restrict__ x, const double *__restrict__ b,
                                                     double *__restrict__ xo) {
    const BCase &C = tab[m.cs[blockIdx.x]];
    const P5Dev A = C.p5;
    p5_sweep2_tile<T>(A, w, x + C.offP, b + C.offP, xo + C.offP, (A.nx + TX - 1) / TX, m.loc[blockIdx.x]);
}

void launch_b_p5_pre(const BCase *tab, const BMapDev &tiles, float omega, const double *b, float *x2, float *r,
                     hipStream_t s) {
    hipLaunchKernelGGL(k_b_p5_pre<float>, dim3(tiles.nblk), dim3(256), 0, s, tab, tiles, omega, b, x2, r);
    CUIBM_LAUNCH_CHECK();
}
void launch_b_p5_pre(const BCase *tab, const BMapDev &tiles, double omega, const double *b, double *x2, double *r,
                     hipStream_t s) {
    hipLaunchKernelGGL(k_b_p5_pre<double>, dim3(tiles.nblk), dim3(256), 0, s, tab, tiles, omega, b, x2, r);
    CUIBM_LAUNCH_CHECK();
}
void launch_b_p5_sweep2(const BCase *tab, const BMapDev &tiles, float omega, const float *x, const double *b,
                        double *xo, hipStream_t s) {
    hipLaunchKernelGGL(k_b_p5_sweep2<float>, dim3(tiles.nblk), dim3(256), 0, s, tab, tiles, omega, x, b, xo);
    CUIBM_LAUNCH_CHECK();
}
void launch_b_p5_sweep2(const BCase *tab, const BMapDev &tiles, double omega, const double *x, const double *b,
                        double *xo, hipStream_t s) {
    hipLaunchKernelGGL(k_b_p5_sweep2<double>, dim3(tiles.nblk), dim3(256), 0, s, tab, tiles, omega, x, b, xo);
    CUIBM_LAUNCH_CHECK();
}

}  // namespace cuibm
