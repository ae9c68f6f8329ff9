// Smoothed-aggregation AMG V-cycle on the device, with the hierarchy's matrix values
// in FP64 or FP32.
//
// One application is one V-cycle from a zero initial guess: on every level but the
// coarsest, npre weighted-Jacobi sweeps (omega = 0.8; the first is x = omega D^-1 b),
// residual, restriction with R = P^T (P^T D^-1 on a row-scaled finest level),
// recursion, prolongation-correction, npost sweeps; the coarsest level (<= 2500 unknowns) applies a dense inverse, one
// wavefront per row.
//
// Mixed precision stores only the matrix values (A_l, P_l, R_l, D_l^-1 and the
// coarse inverse) as float; the level vectors and every product, sum and update stay
// in FP64. Rounding the matrices perturbs a linear operator, so the preconditioner
// stays linear and the outer BiCGSTAB converges as with the FP64 hierarchy. Rounding
// the level vectors instead makes the V-cycle slightly non-linear; on the pinned
// Neumann Poisson operator this lets BiCGSTAB iteration counts grow by orders of
// magnitude over a run and end in breakdown.
#include <algorithm>
#include <cstdint>

#include "krylov.hpp"
#include "stencil.hpp"

namespace cuibm {

namespace {

int grid_for(long n) {
    const long b = (n + BS - 1) / BS;
    return (int)(b < 65536 ? b : 65536);
}

template <class V>
__global__ void k_jacobi_first(int n, const V *__restrict__ dinv, const double *__restrict__ b,
                               double *__restrict__ x, double omega, const double *__restrict__ stop) {
    const double stopv = stop ? *stop : 0.0;  // read at once, tested after the first loads
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double v = omega * (double)dinv[i] * b[i];
        if (stopv != 0.0) return;
        x[i] = v;
    }
}

// One CSR row product per group of LPR lanes (LPR = 1: a thread per row), with the
// per-operator epilogue fused:
//   OP_SPMV    y = A x
//   OP_RESID   y = b - A x
//   OP_JACOBI  y = x + omega D^-1 (b - A x)        (x is the operand of A)
//   OP_PROLONG y += A x                            (A = P, x the coarse correction)
// The lanes of a group add in a fixed shuffle tree, so results depend only on the
// matrix and LPR, never on scheduling.
enum { OP_SPMV = 0, OP_RESID, OP_JACOBI, OP_PROLONG };

template <class V, int LPR, int OP>
__global__ void __launch_bounds__(BS) k_csr(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                            const V *__restrict__ v, const double *__restrict__ x,
                                            const double *__restrict__ b, const V *__restrict__ dinv,
                                            double *__restrict__ y, double omega, const double *__restrict__ stop) {
    const double stopv = stop ? *stop : 0.0;  // read at once, tested after the first row's loads
    const int lane = threadIdx.x & (LPR - 1);
    const int per_block = BS / LPR;
    for (int i = blockIdx.x * per_block + (int)threadIdx.x / LPR; i < n; i += gridDim.x * per_block) {
        double s = 0;
        const int e = rp[i + 1];
        for (int k = rp[i] + lane; k < e; k += LPR) s += (double)v[k] * x[ci[k]];
        if (LPR > 1)
#pragma unroll
            for (int o = LPR / 2; o > 0; o >>= 1) s += __shfl_down(s, o, LPR);
        if (stopv != 0.0) return;
        if (lane == 0) {
            if (OP == OP_SPMV) y[i] = s;
            else if (OP == OP_RESID) y[i] = b[i] - s;
            else if (OP == OP_JACOBI) y[i] = x[i] + omega * (double)dinv[i] * (b[i] - s);
            else y[i] += s;
        }
    }
}

// Restricted residual y = R (b - A x) as one product with G = [R | -R A]: column c < nf
// of G multiplies b[c], column nf + j multiplies x[j]. Replaces the residual and the
// restriction launches; the residual vector itself is never formed.
template <class V, int LPR>
__global__ void __launch_bounds__(BS) k_rres(int n, int nf, const int *__restrict__ rp, const int *__restrict__ ci,
                                             const V *__restrict__ v, const double *__restrict__ x,
                                             const double *__restrict__ b, double *__restrict__ y,
                                             const double *__restrict__ stop) {
    const double stopv = stop ? *stop : 0.0;  // read at once, tested after the first row's loads
    const int lane = threadIdx.x & (LPR - 1);
    const int per_block = BS / LPR;
    for (int i = blockIdx.x * per_block + (int)threadIdx.x / LPR; i < n; i += gridDim.x * per_block) {
        double s = 0;
        const int e = rp[i + 1];
        for (int k = rp[i] + lane; k < e; k += LPR) {
            const int c = ci[k];
            s += (double)v[k] * (c < nf ? b[c] : x[c - nf]);
        }
        if (LPR > 1)
#pragma unroll
            for (int o = LPR / 2; o > 0; o >>= 1) s += __shfl_down(s, o, LPR);
        if (stopv != 0.0) return;
        if (lane == 0) y[i] = s;
    }
}

template <class V>
void launch_rres(int lpr, const DCsr<V> &G, int nf, const double *x, const double *b, double *y, hipStream_t s,
                 const double *stop) {
    const int n = G.n;
    if (n <= 0) return;
    const long blocks = ((long)n * lpr + BS - 1) / BS;
    const dim3 g((unsigned)std::min<long>(blocks, 65536)), t(BS);
    switch (lpr) {
        case 1: hipLaunchKernelGGL((k_rres<V, 1>), g, t, 0, s, n, nf, G.rp.p(), G.ci.p(), G.v.p(), x, b, y, stop); break;
        case 4: hipLaunchKernelGGL((k_rres<V, 4>), g, t, 0, s, n, nf, G.rp.p(), G.ci.p(), G.v.p(), x, b, y, stop); break;
        case 8: hipLaunchKernelGGL((k_rres<V, 8>), g, t, 0, s, n, nf, G.rp.p(), G.ci.p(), G.v.p(), x, b, y, stop); break;
        case 16: hipLaunchKernelGGL((k_rres<V, 16>), g, t, 0, s, n, nf, G.rp.p(), G.ci.p(), G.v.p(), x, b, y, stop); break;
        default: hipLaunchKernelGGL((k_rres<V, 32>), g, t, 0, s, n, nf, G.rp.p(), G.ci.p(), G.v.p(), x, b, y, stop); break;
    }
    CUIBM_LAUNCH_CHECK();
}

// Both pre-smoothing sweeps of a CSR level in one pass: x2 = x1 + omega D^-1 (b - A x1)
// with x1 = omega D^-1 b evaluated at the row's neighbours on the fly. (Fusing the
// prolongation into the first post-sweep the same way was measured slower: every row
// then gathers its neighbours' whole P rows.)
template <class V, int LPR>
__global__ void __launch_bounds__(BS) k_pre2(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                             const V *__restrict__ v, const V *__restrict__ dinv,
                                             const double *__restrict__ b, double *__restrict__ x2, double omega,
                                             const double *__restrict__ stop) {
    const double stopv = stop ? *stop : 0.0;  // read at once, tested after the first row's loads
    const int lane = threadIdx.x & (LPR - 1);
    const int per_block = BS / LPR;
    for (int i = blockIdx.x * per_block + (int)threadIdx.x / LPR; i < n; i += gridDim.x * per_block) {
        double s = 0;
        const int e = rp[i + 1];
        for (int k = rp[i] + lane; k < e; k += LPR) {
            const int j = ci[k];
            s += (double)v[k] * (omega * (double)dinv[j] * b[j]);
        }
        if (LPR > 1)
#pragma unroll
            for (int o = LPR / 2; o > 0; o >>= 1) s += __shfl_down(s, o, LPR);
        if (stopv != 0.0) return;
        if (lane == 0) {
            const double wd = omega * (double)dinv[i];
            x2[i] = wd * b[i] + wd * (b[i] - s);
        }
    }
}

template <class V>
void launch_pre2(int lpr, const DCsr<V> &A, const V *dinv, const double *b, double *x2, double omega, hipStream_t s,
                 const double *stop) {
    const int n = A.n;
    const long blocks = ((long)n * lpr + BS - 1) / BS;
    const dim3 g((unsigned)std::min<long>(blocks, 65536)), t(BS);
    switch (lpr) {
        case 1: hipLaunchKernelGGL((k_pre2<V, 1>), g, t, 0, s, n, A.rp.p(), A.ci.p(), A.v.p(), dinv, b, x2, omega, stop); break;
        case 4: hipLaunchKernelGGL((k_pre2<V, 4>), g, t, 0, s, n, A.rp.p(), A.ci.p(), A.v.p(), dinv, b, x2, omega, stop); break;
        case 8: hipLaunchKernelGGL((k_pre2<V, 8>), g, t, 0, s, n, A.rp.p(), A.ci.p(), A.v.p(), dinv, b, x2, omega, stop); break;
        case 16: hipLaunchKernelGGL((k_pre2<V, 16>), g, t, 0, s, n, A.rp.p(), A.ci.p(), A.v.p(), dinv, b, x2, omega, stop); break;
        default: hipLaunchKernelGGL((k_pre2<V, 32>), g, t, 0, s, n, A.rp.p(), A.ci.p(), A.v.p(), dinv, b, x2, omega, stop); break;
    }
    CUIBM_LAUNCH_CHECK();
}

// lanes per row for a CSR operator, from its mean row length
int lanes_for(const HCsr &A) {
    const double avg = A.n > 0 ? (double)A.ci.size() / A.n : 0.0;
    if (avg < 6) return 1;
    if (avg < 12) return 4;
    if (avg < 24) return 8;
    if (avg < 48) return 16;
    return 32;
}

template <class V, int OP>
void launch_csr_op(int lpr, const DCsr<V> &A, const double *x, const double *b, const V *dinv, double *y, double omega,
                   hipStream_t s, const double *stop = nullptr) {
    const int n = A.n;
    if (n <= 0) return;
    const long blocks = ((long)n * lpr + BS - 1) / BS;
    const dim3 g((unsigned)std::min<long>(blocks, 65536)), t(BS);
    switch (lpr) {
        case 1: hipLaunchKernelGGL((k_csr<V, 1, OP>), g, t, 0, s, n, A.rp.p(), A.ci.p(), A.v.p(), x, b, dinv, y, omega, stop); break;
        case 4: hipLaunchKernelGGL((k_csr<V, 4, OP>), g, t, 0, s, n, A.rp.p(), A.ci.p(), A.v.p(), x, b, dinv, y, omega, stop); break;
        case 8: hipLaunchKernelGGL((k_csr<V, 8, OP>), g, t, 0, s, n, A.rp.p(), A.ci.p(), A.v.p(), x, b, dinv, y, omega, stop); break;
        case 16: hipLaunchKernelGGL((k_csr<V, 16, OP>), g, t, 0, s, n, A.rp.p(), A.ci.p(), A.v.p(), x, b, dinv, y, omega, stop); break;
        default: hipLaunchKernelGGL((k_csr<V, 32, OP>), g, t, 0, s, n, A.rp.p(), A.ci.p(), A.v.p(), x, b, dinv, y, omega, stop); break;
    }
    CUIBM_LAUNCH_CHECK();
}

// x = Ainv b on the coarsest level. Ainv is row-major with a row stride ld that is a
// multiple of 4 (zero padding), so each lane loads four consecutive matrix values at a
// time (one 16-byte load for FP32, two for FP64) and the matching four values of b, which
// every wavefront reads from cache; one wavefront per row, four partial sums per lane
// combined in a fixed order, then a fixed shuffle tree across the wavefront
template <class V>
__device__ __forceinline__ void load4(const V *p, double (&v)[4]);
template <>
__device__ __forceinline__ void load4<float>(const float *p, double (&v)[4]) {
    const float4 q = *reinterpret_cast<const float4 *>(p);
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
}
template <>
__device__ __forceinline__ void load4<double>(const double *p, double (&v)[4]) {
    const double2 q0 = reinterpret_cast<const double2 *>(p)[0], q1 = reinterpret_cast<const double2 *>(p)[1];
    v[0] = q0.x; v[1] = q0.y; v[2] = q1.x; v[3] = q1.y;
}

template <class V>
__global__ void __launch_bounds__(256) k_dense(int n, int ld, const V *__restrict__ Ainv, const double *__restrict__ b,
                                               double *__restrict__ x, const double *__restrict__ stop) {
    const double stopv = stop ? *stop : 0.0;
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;  // no barrier in this kernel
    const V *a = Ainv + (size_t)row * ld;
    const int n4 = n & ~3;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    bool tested = false;
    for (int j = 4 * lane; j < n4; j += 256) {
        double av[4], bv[4];
        load4<V>(a + j, av);
        load4<double>(b + j, bv);
        if (!tested) {
            if (stopv != 0.0) return;
            tested = true;
        }
        s0 += av[0] * bv[0];
        s1 += av[1] * bv[1];
        s2 += av[2] * bv[2];
        s3 += av[3] * bv[3];
    }
    if (stopv != 0.0) return;
    if (lane < n - n4) s0 += (double)a[n4 + lane] * b[n4 + lane];
    double s = (s0 + s1) + (s2 + s3);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) x[row] = s;
}

}  // namespace

template <class V>
struct AmgLevelDev {
    int n = 0;
    int lanesA = 1, lanesP = 1, lanesR = 1, lanesG = 1;  // lanes per row of each operator
    DCsr<V> A, P, R, G;
    DArray<V> dinv;
    DArray<double> b, x, xt, r;
};

template <class V>
struct AmgDev {
    std::vector<AmgLevelDev<V>> lv;
    DArray<V> coarse_inv;
    int nc = 0, ld = 0;
    AmgParams prm;
    bool use_fine = false;  // finest level in stencil form (fused kernels)
    P5Dev fine{};
    KernelProbe *probe = nullptr;  // events around the finest-level pre-smoothing launches
    const double *stop = nullptr;  // the Krylov solver's stop flag (every kernel checks it)

    void build(const std::vector<HostAmgLevel> &h, hipStream_t s) {
        lv.resize(h.size());
        for (size_t l = 0; l < h.size(); l++) {
            AmgLevelDev<V> &L = lv[l];
            L.n = h[l].A.n;
            const bool stencil_level = (l == 0 && use_fine);
            if (!stencil_level) upload_csr(h[l].A, L.A, s);
            L.lanesA = lanes_for(h[l].A);
            if (l + 1 < h.size()) {
                upload_csr(h[l].P, L.P, s);
                upload_csr(h[l].R, L.R, s);
                L.lanesP = lanes_for(h[l].P);
                L.lanesR = lanes_for(h[l].R);
                if (!stencil_level && h[l].G.n > 0) {
                    upload_csr(h[l].G, L.G, s);
                    L.lanesG = lanes_for(h[l].G);
                }
            }
            if (!stencil_level) {
                std::vector<V> d(L.n, V(0));
                for (int i = 0; i < L.n; i++) {
                    double a = 0;
                    for (int k = h[l].A.rp[i]; k < h[l].A.rp[i + 1]; k++)
                        if (h[l].A.ci[k] == i) a += h[l].A.v[k];
                    d[i] = (V)(1.0 / a);
                }
                L.dinv.upload(d, s);
            }
            if (l > 0) {  // level 0 works on the caller's vectors
                L.b.alloc(L.n);
                L.x.alloc(L.n);
            }
            L.xt.alloc(L.n);
            L.r.alloc(L.n);
        }
        nc = h.back().A.n;
        if (nc > 6144) throw Error("AMG coarsening stalled at " + std::to_string(nc) + " unknowns", ERR_GENERIC);
        const std::vector<double> inv = dense_inverse(h.back().A);
        ld = (nc + 3) & ~3;  // row stride, zero-padded to a multiple of 4 values
        std::vector<V> invt((size_t)nc * ld, V(0));
        for (int i = 0; i < nc; i++)
            for (int j = 0; j < nc; j++) invt[(size_t)i * ld + j] = (V)inv[(size_t)i * nc + j];
        coarse_inv.upload(invt, s);
        CUIBM_HIP(hipStreamSynchronize(s));
    }

    // V-cycle on level l with right-hand side b, result into x (both level-sized)
    void vcycle(size_t l, const double *b, double *x, hipStream_t s) {
        AmgLevelDev<V> &L = lv[l];
        const int n = L.n;
        const double w = prm.omega;
        if (l + 1 == lv.size()) {
            if ((reinterpret_cast<uintptr_t>(b) & 15) != 0)  // the kernel reads b four values at a time
                throw Error("coarsest AMG right-hand side is not 16-byte aligned", ERR_ARGUMENT);
            hipLaunchKernelGGL(k_dense<V>, dim3((n + 3) / 4), dim3(256), 0, s, n, ld, coarse_inv.p(), b, x,
                               stop);
            CUIBM_LAUNCH_CHECK();
            return;
        }
        if (l == 0 && use_fine && prm.npre == 2 && prm.npost == 2) {
            // stencil finest level: x2 -> xt and r -> r, restriction, coarse cycle,
            // prolongation into xt, both post-sweeps straight into x
            launch_p5_pre(fine, w, b, L.xt.p(), L.r.p(), s, probe);
            AmgLevelDev<V> &C = lv[1];
            launch_csr_op<V, OP_SPMV>(L.lanesR, L.R, L.r.p(), nullptr, nullptr, C.b.p(), 0.0, s, stop);
            vcycle(1, C.b.p(), C.x.p(), s);
            launch_csr_op<V, OP_PROLONG>(L.lanesP, L.P, C.x.p(), nullptr, nullptr, L.xt.p(), 0.0, s, stop);
            launch_p5_sweep2(fine, w, L.xt.p(), b, x, s);
            return;
        }
        if (prm.npre == 2 && prm.npost == 2) {
            // x2 -> xt, r -> r, restriction, coarse cycle, prolongation into xt,
            // post-sweeps xt -> r -> x
            AmgLevelDev<V> &C = lv[l + 1];
            launch_pre2<V>(L.lanesA, L.A, L.dinv.p(), b, L.xt.p(), w, s, stop);
            if (L.G.n > 0) {
                launch_rres<V>(L.lanesG, L.G, n, L.xt.p(), b, C.b.p(), s, stop);  // R (b - A x2)
            } else {
                launch_csr_op<V, OP_RESID>(L.lanesA, L.A, L.xt.p(), b, nullptr, L.r.p(), 0.0, s, stop);
                launch_csr_op<V, OP_SPMV>(L.lanesR, L.R, L.r.p(), nullptr, nullptr, C.b.p(), 0.0, s, stop);
            }
            vcycle(l + 1, C.b.p(), C.x.p(), s);
            launch_csr_op<V, OP_PROLONG>(L.lanesP, L.P, C.x.p(), nullptr, nullptr, L.xt.p(), 0.0, s, stop);
            launch_csr_op<V, OP_JACOBI>(L.lanesA, L.A, L.xt.p(), b, L.dinv.p(), L.r.p(), w, s, stop);
            launch_csr_op<V, OP_JACOBI>(L.lanesA, L.A, L.r.p(), b, L.dinv.p(), x, w, s, stop);
            return;
        }
        const int g = grid_for(n);
        // ping-pong so that the last sweep lands in x
        const int swaps = (prm.npre - 1) + prm.npost;
        double *cur = (swaps % 2 == 0) ? x : L.xt.p();
        double *oth = (cur == x) ? L.xt.p() : x;
        hipLaunchKernelGGL(k_jacobi_first<V>, dim3(g), dim3(BS), 0, s, n, L.dinv.p(), b, cur, w, stop);
        CUIBM_LAUNCH_CHECK();
        for (int k = 1; k < prm.npre; k++) {
            launch_csr_op<V, OP_JACOBI>(L.lanesA, L.A, cur, b, L.dinv.p(), oth, w, s, stop);
            std::swap(cur, oth);
        }
        launch_csr_op<V, OP_RESID>(L.lanesA, L.A, cur, b, nullptr, L.r.p(), 0.0, s, stop);
        AmgLevelDev<V> &C = lv[l + 1];
        launch_csr_op<V, OP_SPMV>(L.lanesR, L.R, L.r.p(), nullptr, nullptr, C.b.p(), 0.0, s, stop);
        vcycle(l + 1, C.b.p(), C.x.p(), s);
        launch_csr_op<V, OP_PROLONG>(L.lanesP, L.P, C.x.p(), nullptr, nullptr, cur, 0.0, s, stop);
        for (int k = 0; k < prm.npost; k++) {
            launch_csr_op<V, OP_JACOBI>(L.lanesA, L.A, cur, b, L.dinv.p(), oth, w, s, stop);
            std::swap(cur, oth);
        }
    }
};

struct AmgPrecond::Impl {
    AmgDev<double> d64;
    AmgDev<float> d32;
    std::vector<HostAmgLevel> host;
    KernelProbe probe;
};

KernelProbe &AmgPrecond::pre_probe() { return im_->probe; }

AmgPrecond::AmgPrecond(const HCsr &A, const AmgParams &p, hipStream_t s, const P5Dev *fine)
    : im_(new Impl), prm_(p) {
    im_->host = amg_build(A, p.coarse_max, p.max_levels, p.scale0, p.fused_restriction);
    fine_ = fine != nullptr && im_->host.size() > 1 && p.npre == 2 && p.npost == 2;
    if (fine_) {
        if (fine->nx * fine->ny != A.n) throw Error("stencil and matrix sizes differ", ERR_ARGUMENT);
        fine_dev_ = *fine;
    }
    im_->d32.probe = &im_->probe;
    im_->d64.probe = &im_->probe;
    if (p.precision == 32) {
        im_->d32.prm = p;
        im_->d32.use_fine = fine_;
        if (fine_) im_->d32.fine = *fine;
        im_->d32.build(im_->host, s);
    } else {
        im_->d64.prm = p;
        im_->d64.use_fine = fine_;
        if (fine_) im_->d64.fine = *fine;
        im_->d64.build(im_->host, s);
    }
    // keep only the sizes of the host hierarchy
    for (auto &L : im_->host) {
        L.A.v.clear(); L.A.v.shrink_to_fit();
        L.P.v.clear(); L.P.v.shrink_to_fit();
        L.R.v.clear(); L.R.v.shrink_to_fit();
        L.G.v.clear(); L.G.v.shrink_to_fit();
    }
}

AmgPrecond::~AmgPrecond() = default;

void AmgPrecond::set_stop(const double *stop) {
    stop_ = stop;
    im_->d64.stop = stop;
    im_->d32.stop = stop;
    im_->d64.fine.stop = stop;
    im_->d32.fine.stop = stop;
}

void AmgPrecond::apply(const double *r, double *z, hipStream_t s) {
    if (prm_.precision == 32) im_->d32.vcycle(0, r, z, s);
    else im_->d64.vcycle(0, r, z, s);
}

void AmgPrecond::level0_sweep(const double *b, const double *x, double *xn, hipStream_t s) {
    if (fine_) {
        launch_p5_sweep(fine_dev_, prm_.omega, x, b, xn, s);
    } else if (prm_.precision == 32) {
        auto &L = im_->d32.lv[0];
        launch_csr_op<float, OP_JACOBI>(L.lanesA, L.A, x, b, L.dinv.p(), xn, prm_.omega, s);
    } else {
        auto &L = im_->d64.lv[0];
        launch_csr_op<double, OP_JACOBI>(L.lanesA, L.A, x, b, L.dinv.p(), xn, prm_.omega, s);
    }
}

double AmgPrecond::level0_sweep_bytes() const {
    const double sv = prm_.precision == 32 ? 4.0 : 8.0;
    const auto &A = im_->host[0].A;
    const double n = A.n, nnz = (double)A.ci.size();
    if (fine_) return n * (8 + 8 + 8) + 32.0 * (fine_dev_.nx + fine_dev_.ny);  // x, b, xn, coefficient arrays
    return nnz * (sv + 4) + 4 * (n + 1) + n * (8 + 8 + 8 + sv);  // A, rowptr, x, b, xn, dinv
}

void AmgPrecond::level0_pre(const double *b, double *x2, double *r, hipStream_t s) {
    if (!fine_) throw Error("the finest AMG level is not in stencil form", ERR_ARGUMENT);
    launch_p5_pre(fine_dev_, prm_.omega, b, x2, r, s);
}

void AmgPrecond::level0_post(double *x2, const double *b, double *x, hipStream_t s) {
    if (!fine_) throw Error("the finest AMG level is not in stencil form", ERR_ARGUMENT);
    if (prm_.precision == 32) {
        auto &L = im_->d32.lv[0];
        launch_csr_op<float, OP_PROLONG>(L.lanesP, L.P, im_->d32.lv[1].x.p(), nullptr, nullptr, x2, 0.0, s);
    } else {
        auto &L = im_->d64.lv[0];
        launch_csr_op<double, OP_PROLONG>(L.lanesP, L.P, im_->d64.lv[1].x.p(), nullptr, nullptr, x2, 0.0, s);
    }
    launch_p5_sweep2(fine_dev_, prm_.omega, x2, b, x, s);
}

double AmgPrecond::level0_pre_bytes() const {
    return 4.0 * 8 * im_->host[0].A.n + 32.0 * (fine_dev_.nx + fine_dev_.ny);  // b, D^-1 in; x2, r out
}

double AmgPrecond::level0_post_bytes() const {
    const double sv = prm_.precision == 32 ? 4.0 : 8.0;
    const double n = im_->host[0].A.n, nc = im_->host[1].A.n, nnzP = (double)im_->host[0].P.ci.size();
    const double prolong = nnzP * (sv + 4) + 4 * (n + 1) + nc * 8 + 2 * n * 8;            // P, xc; x2 read-modify-write
    const double sweeps = 4.0 * 8 * n + 32.0 * (fine_dev_.nx + fine_dev_.ny);            // x, b, D^-1 in; x out
    return prolong + sweeps;
}

AmgHierarchyInfo AmgPrecond::info() const {
    AmgHierarchyInfo I;
    I.levels = (int)im_->host.size();
    for (auto &L : im_->host) {
        I.n.push_back(L.A.n);
        I.nnz.push_back((int)L.A.ci.size());
        I.nnzP.push_back((int)L.P.ci.size());
    }
    return I;
}

double AmgPrecond::bytes_per_apply() const {
    // distinct arrays per kernel; matrix values and D^-1 take sv bytes, vectors 8:
    // CSR product = nnz*(sv+4) + 4(n+1) + 8 ncol (x) + 8 n (y)
    const double sv = prm_.precision == 32 ? 4.0 : 8.0;
    double bytes = 0;
    const auto &H = im_->host;
    auto spmv = [&](double n, double ncol, double nnz) { return nnz * (sv + 4) + 4 * (n + 1) + ncol * 8 + n * 8; };
    for (size_t l = 0; l + 1 < H.size(); l++) {
        const double n = H[l].A.n, nnz = (double)H[l].A.ci.size();
        const double nc = H[l + 1].A.n, nnzP = (double)H[l].P.ci.size();
        if (l == 0 && fine_) {
            bytes += level0_pre_bytes() + spmv(nc, n, nnzP) + level0_post_bytes();  // pre, restriction, post
            continue;
        }
        if (prm_.npre == 2 && prm_.npost == 2) {  // the fused (2,2) sequence of vcycle()
            const double nnzG = (double)H[l].G.ci.size();
            bytes += spmv(n, n, nnz) + n * sv + n * 8;                    // both pre-sweeps (+ dinv, b)
            if (nnzG > 0) bytes += nnzG * (sv + 4) + 4 * (nc + 1) + 2 * n * 8 + nc * 8;  // R (b - A x2) via G
            else bytes += spmv(n, n, nnz) + n * 8 + spmv(nc, n, nnzP);    // residual (+ b), restriction
            bytes += spmv(n, nc, nnzP) + n * 8;                           // prolong-add
            bytes += 2 * (spmv(n, n, nnz) + n * sv + n * 8);              // post-sweeps (+ dinv, b)
            continue;
        }
        bytes += n * sv + 2 * n * 8;                                   // first sweep: dinv, b, x
        const int sweeps = prm_.npre - 1 + prm_.npost;
        bytes += sweeps * (spmv(n, n, nnz) + n * sv + n * 8);          // + dinv, b
        bytes += spmv(n, n, nnz) + n * 8;                              // residual (+ b)
        bytes += spmv(nc, n, nnzP);                                    // restriction
        bytes += spmv(n, nc, nnzP) + n * 8;                            // prolong-add (read-modify-write x)
    }
    const double nc = H.back().A.n;
    bytes += nc * nc * sv + 2 * nc * 8;
    return bytes;
}

}  // namespace cuibm
