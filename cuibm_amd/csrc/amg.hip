// Smoothed-aggregation AMG V-cycle on the device, with the hierarchy's matrix values
// in FP64 or FP32.
//
// One application is one V-cycle from a zero initial guess: on every level but the
// coarsest, npre weighted-Jacobi sweeps (omega = 0.8; the first is x = omega D^-1 b),
// residual, restriction with R = P^T (P^T D^-1 on a row-scaled finest level),
// recursion, prolongation-correction, npost sweeps; the coarsest level (<= 2500 unknowns) applies a dense inverse, one
// wavefront per row.
//
// Three precisions, as template parameters V (matrix values: A_l, P_l, R_l, D_l^-1 and
// the coarse inverse) and X (level vectors and the arithmetic of products, sums and
// updates): FP64 (double, double); FP32 values with FP64 vectors (float, double), where
// rounding only perturbs a linear operator; and the FP32 V-cycle (float, float), which
// reads the FP64 Krylov vector and writes its result in FP64 but keeps every level
// vector, smoother sweep and product in FP32 (the coarsest solve accumulates in FP64).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <type_traits>

#include "batch.hpp"
#include "krylov.hpp"
#include "stencil.hpp"

namespace cuibm {

namespace {

int grid_for(long n) {
    const long b = (n + BS - 1) / BS;
    return (int)(b < 65536 ? b : 65536);
}

template <class V, class X>
__global__ void k_jacobi_first(int n, const V *__restrict__ dinv, const X *__restrict__ b, X *__restrict__ x, X omega,
                               const double *__restrict__ stop) {
    const double stopv = stop ? *stop : 0.0;  // read at once, tested after the first loads
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const X v = omega * (X)dinv[i] * b[i];
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

template <class V, class X, int LPR, int OP>
__global__ void __launch_bounds__(BS) k_csr(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                            const V *__restrict__ v, const X *__restrict__ x, const X *__restrict__ b,
                                            const V *__restrict__ dinv, X *__restrict__ y, X omega,
                                            const double *__restrict__ stop, const int *__restrict__ seg,
                                            const double *const *__restrict__ cstop) {
    const double stopv = stop ? *stop : 0.0;  // read at once, tested after the first row's loads
    const int lane = threadIdx.x & (LPR - 1);
    const int per_block = BS / LPR;
    // rows banded by XCD, except in a batch, where the cases that stop early would leave
    // whole XCDs idle (blocks in dispatch order spread each case over all of them)
    const int blk = seg ? (int)blockIdx.x : xcd_block(gridDim.x);
    for (int i = blk * per_block + (int)threadIdx.x / LPR; i < n; i += gridDim.x * per_block) {
        if (seg && *cstop[seg[i / BATCH_PAD]] != 0.0) continue;  // a stopped case of a batch (whole row group)
        // grouped loads up to 4 lanes per row; with 8 or more lanes (the finest-level
        // restriction) a lane has two or three entries and one at a time measured faster
        // (10.5 against 12.5 us per launch at 1024x1024); the sum is the same either way
        X s = 0;
        if (LPR <= 4) s = csr_lane_sum<X, LPR>(ci, v, rp[i] + lane, rp[i + 1], [&](int c) { return x[c]; });
        else
            for (int k = rp[i] + lane, e = rp[i + 1]; k < e; k += LPR) s += (X)v[k] * x[ci[k]];
        if (LPR > 1)
#pragma unroll
            for (int o = LPR / 2; o > 0; o >>= 1) s += __shfl_down(s, o, LPR);
        if (stopv != 0.0) return;
        if (lane == 0) {
            if (OP == OP_SPMV) y[i] = s;
            else if (OP == OP_RESID) y[i] = b[i] - s;
            else if (OP == OP_JACOBI) y[i] = x[i] + omega * (X)dinv[i] * (b[i] - s);
            else y[i] += s;
        }
    }
}

// Restricted residual y = R (b - A x) as one product with G = [R | -R A]: column c < nf
// of G multiplies b[c], column nf + j multiplies x[j]. Replaces the residual and the
// restriction launches; the residual vector itself is never formed.
template <class V, class X, int LPR>
__global__ void __launch_bounds__(BS) k_rres(int n, int nf, const int *__restrict__ rp, const int *__restrict__ ci,
                                             const V *__restrict__ v, const X *__restrict__ x, const X *__restrict__ b,
                                             X *__restrict__ y, const double *__restrict__ stop,
                                             const int *__restrict__ seg, const double *const *__restrict__ cstop) {
    const double stopv = stop ? *stop : 0.0;  // read at once, tested after the first row's loads
    const int lane = threadIdx.x & (LPR - 1);
    const int per_block = BS / LPR;
    // rows banded by XCD, except in a batch, where the cases that stop early would leave
    // whole XCDs idle (blocks in dispatch order spread each case over all of them)
    const int blk = seg ? (int)blockIdx.x : xcd_block(gridDim.x);
    for (int i = blk * per_block + (int)threadIdx.x / LPR; i < n; i += gridDim.x * per_block) {
        if (seg && *cstop[seg[i / BATCH_PAD]] != 0.0) continue;  // a stopped case of a batch (whole row group)
        // column c < nf reads b[c], the others x[c - nf]: the address is chosen, not the load
        X s = csr_lane_sum<X, LPR>(ci, v, rp[i] + lane, rp[i + 1], [&](int c) { return *(c < nf ? b + c : x + (c - nf)); });
        if (LPR > 1)
#pragma unroll
            for (int o = LPR / 2; o > 0; o >>= 1) s += __shfl_down(s, o, LPR);
        if (stopv != 0.0) return;
        if (lane == 0) y[i] = s;
    }
}

template <class V, class X>
void launch_rres(int lpr, const DCsr<V> &G, int nf, const X *x, const X *b, X *y, hipStream_t s, const double *stop,
                 const int *seg = nullptr, const double *const *cstop = nullptr) {
    const int n = G.n;
    if (n <= 0) return;
    const long blocks = ((long)n * lpr + BS - 1) / BS;
    const dim3 g(xcd_blocks(blocks)), t(BS);
    switch (lpr) {
        case 1: hipLaunchKernelGGL((k_rres<V, X, 1>), g, t, 0, s, n, nf, G.rp.p(), G.ci.p(), G.v.p(), x, b, y, stop, seg, cstop); break;
        case 4: hipLaunchKernelGGL((k_rres<V, X, 4>), g, t, 0, s, n, nf, G.rp.p(), G.ci.p(), G.v.p(), x, b, y, stop, seg, cstop); break;
        case 8: hipLaunchKernelGGL((k_rres<V, X, 8>), g, t, 0, s, n, nf, G.rp.p(), G.ci.p(), G.v.p(), x, b, y, stop, seg, cstop); break;
        case 16: hipLaunchKernelGGL((k_rres<V, X, 16>), g, t, 0, s, n, nf, G.rp.p(), G.ci.p(), G.v.p(), x, b, y, stop, seg, cstop); break;
        default: hipLaunchKernelGGL((k_rres<V, X, 32>), g, t, 0, s, n, nf, G.rp.p(), G.ci.p(), G.v.p(), x, b, y, stop, seg, cstop); break;
    }
    CUIBM_LAUNCH_CHECK();
}

// Both pre-smoothing sweeps of a CSR level in one pass: x2 = x1 + omega D^-1 (b - A x1)
// with x1 = omega D^-1 b evaluated at the row's neighbours on the fly. (Fusing the
// prolongation into the first post-sweep the same way was measured slower: every row
// then gathers its neighbours' whole P rows.)
template <class V, class X, int LPR>
__global__ void __launch_bounds__(BS) k_pre2(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                             const V *__restrict__ v, const V *__restrict__ dinv,
                                             const X *__restrict__ b, X *__restrict__ x2, X omega,
                                             const double *__restrict__ stop, const int *__restrict__ seg,
                                             const double *const *__restrict__ cstop) {
    const double stopv = stop ? *stop : 0.0;  // read at once, tested after the first row's loads
    const int lane = threadIdx.x & (LPR - 1);
    const int per_block = BS / LPR;
    // rows banded by XCD, except in a batch, where the cases that stop early would leave
    // whole XCDs idle (blocks in dispatch order spread each case over all of them)
    const int blk = seg ? (int)blockIdx.x : xcd_block(gridDim.x);
    for (int i = blk * per_block + (int)threadIdx.x / LPR; i < n; i += gridDim.x * per_block) {
        if (seg && *cstop[seg[i / BATCH_PAD]] != 0.0) continue;  // a stopped case of a batch (whole row group)
        X s = csr_lane_sum<X, LPR>(ci, v, rp[i] + lane, rp[i + 1], [&](int j) { return omega * (X)dinv[j] * b[j]; });
        if (LPR > 1)
#pragma unroll
            for (int o = LPR / 2; o > 0; o >>= 1) s += __shfl_down(s, o, LPR);
        if (stopv != 0.0) return;
        if (lane == 0) {
            const X wd = omega * (X)dinv[i];
            x2[i] = wd * b[i] + wd * (b[i] - s);
        }
    }
}

template <class V, class X>
void launch_pre2(int lpr, const DCsr<V> &A, const V *dinv, const X *b, X *x2, X omega, hipStream_t s,
                 const double *stop, const int *seg = nullptr, const double *const *cstop = nullptr) {
    const int n = A.n;
    const long blocks = ((long)n * lpr + BS - 1) / BS;
    const dim3 g(xcd_blocks(blocks)), t(BS);
    switch (lpr) {
        case 1: hipLaunchKernelGGL((k_pre2<V, X, 1>), g, t, 0, s, n, A.rp.p(), A.ci.p(), A.v.p(), dinv, b, x2, omega, stop, seg, cstop); break;
        case 4: hipLaunchKernelGGL((k_pre2<V, X, 4>), g, t, 0, s, n, A.rp.p(), A.ci.p(), A.v.p(), dinv, b, x2, omega, stop, seg, cstop); break;
        case 8: hipLaunchKernelGGL((k_pre2<V, X, 8>), g, t, 0, s, n, A.rp.p(), A.ci.p(), A.v.p(), dinv, b, x2, omega, stop, seg, cstop); break;
        case 16: hipLaunchKernelGGL((k_pre2<V, X, 16>), g, t, 0, s, n, A.rp.p(), A.ci.p(), A.v.p(), dinv, b, x2, omega, stop, seg, cstop); break;
        default: hipLaunchKernelGGL((k_pre2<V, X, 32>), g, t, 0, s, n, A.rp.p(), A.ci.p(), A.v.p(), dinv, b, x2, omega, stop, seg, cstop); break;
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

template <class V, class X, int OP>
void launch_csr_op(int lpr, const DCsr<V> &A, const X *x, const X *b, const V *dinv, X *y, X omega, hipStream_t s,
                   const double *stop = nullptr, const int *seg = nullptr, const double *const *cstop = nullptr) {
    const int n = A.n;
    if (n <= 0) return;
    const long blocks = ((long)n * lpr + BS - 1) / BS;
    const dim3 g(xcd_blocks(blocks)), t(BS);
    switch (lpr) {
        case 1: hipLaunchKernelGGL((k_csr<V, X, 1, OP>), g, t, 0, s, n, A.rp.p(), A.ci.p(), A.v.p(), x, b, dinv, y, omega, stop, seg, cstop); break;
        case 4: hipLaunchKernelGGL((k_csr<V, X, 4, OP>), g, t, 0, s, n, A.rp.p(), A.ci.p(), A.v.p(), x, b, dinv, y, omega, stop, seg, cstop); break;
        case 8: hipLaunchKernelGGL((k_csr<V, X, 8, OP>), g, t, 0, s, n, A.rp.p(), A.ci.p(), A.v.p(), x, b, dinv, y, omega, stop, seg, cstop); break;
        case 16: hipLaunchKernelGGL((k_csr<V, X, 16, OP>), g, t, 0, s, n, A.rp.p(), A.ci.p(), A.v.p(), x, b, dinv, y, omega, stop, seg, cstop); break;
        default: hipLaunchKernelGGL((k_csr<V, X, 32, OP>), g, t, 0, s, n, A.rp.p(), A.ci.p(), A.v.p(), x, b, dinv, y, omega, stop, seg, cstop); break;
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

template <class V, class X>
__device__ __forceinline__ void dense_row(int n, int ld, const V *__restrict__ Ainv, const X *__restrict__ b,
                                          X *__restrict__ x, int row, double stopv) {
    const int lane = threadIdx.x & 63;
    if (row >= n) return;  // no barrier in this function
    const V *a = Ainv + (size_t)row * ld;
    const int n4 = n & ~3;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    bool tested = false;
    for (int j = 4 * lane; j < n4; j += 256) {
        double av[4], bv[4];
        load4<V>(a + j, av);
        load4<X>(b + j, bv);
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
    if (lane == 0) x[row] = (X)s;
}

template <class V, class X>
__global__ void __launch_bounds__(256) k_dense(int n, int ld, const V *__restrict__ Ainv, const X *__restrict__ b,
                                               X *__restrict__ x, const double *__restrict__ stop) {
    const double stopv = stop ? *stop : 0.0;
    dense_row<V, X>(n, ld, Ainv, b, x, blockIdx.x * 4 + (threadIdx.x >> 6), stopv);
}

// the dense coarsest solves of several cases in one launch (lockstep batches): case c's
// inverse, its size and row stride, and its offset in the level vectors
template <class V>
struct DenseCase {
    const V *inv;
    int n, ld;
    long off;
    const double *stop;  // the case's Krylov stop code: the solve is skipped once it is set
};

template <class V, class X>
__global__ void __launch_bounds__(256) k_b_dense(const DenseCase<V> *__restrict__ tab, BMapDev m,
                                                 const X *__restrict__ b, X *__restrict__ x) {
    const DenseCase<V> &D = tab[m.cs[blockIdx.x]];
    dense_row<V, X>(D.n, D.ld, D.inv, b + D.off, x + D.off, m.loc[blockIdx.x] * 4 + (threadIdx.x >> 6),
                    D.stop ? *D.stop : 0.0);
}

}  // namespace

template <class V, class X>
struct AmgLevelDev {
    int n = 0;
    int lanesA = 1, lanesP = 1, lanesR = 1, lanesG = 1;  // lanes per row of each operator
    DCsr<V> A, P, R, G;
    DArray<V> dinv;
    DArray<X> b, x, xt, r;
};

template <class V, class X>
struct AmgDev {
    std::vector<AmgLevelDev<V, X>> lv;
    DArray<V> coarse_inv;
    DArray<float> fine_dinv32;  // FP32 D^-1 of the stencil finest level (FP32 V-cycle)
    // FP64 deflation of the pinned constant mode (FP32 V-cycle on a stencil finest level):
    // partial sums, c, and 1 / sum over the pins of area * (A 1)
    DArray<double> defl_part, defl_c;
    double inv_wa = 0;
    int nc = 0, ld = 0;
    AmgParams prm;
    bool use_fine = false;  // finest level in stencil form (fused kernels)
    P5Dev fine{};
    KernelProbe *probe = nullptr;  // events around the finest-level pre-smoothing launches
    const double *stop = nullptr;  // the Krylov solver's stop flag (every kernel checks it)
    static constexpr bool kVec64 = std::is_same<X, double>::value;

    void build(const std::vector<HostAmgLevel> &h, hipStream_t s) {
        lv.resize(h.size());
        for (size_t l = 0; l < h.size(); l++) {
            AmgLevelDev<V, X> &L = lv[l];
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
            // level 0 works on the caller's FP64 vectors, except a CSR finest level of
            // the FP32 V-cycle, which converts them into and out of its own
            if (l > 0 || (!kVec64 && !stencil_level)) {
                L.b.alloc(L.n);
                L.x.alloc(L.n);
            }
            L.xt.alloc(L.n);
            L.r.alloc(L.n);
        }
        if (use_fine && !kVec64) {
            fine_dinv32.alloc(lv[0].n);
            launch_to_float(fine_dinv32.p(), fine.dinv, lv[0].n, s);
            fine.dinv32 = fine_dinv32.p();
        }
        nc = h.back().A.n;
        if (nc > 6144) throw Error("AMG coarsening stalled at " + std::to_string(nc) + " unknowns", ERR_GENERIC);
        std::vector<double> inv = dense_inverse(h.back().A);
        if (fine.defl) deflate_coarse(inv, nc);
        ld = (nc + 3) & ~3;  // row stride, zero-padded to a multiple of 4 values
        std::vector<V> invt((size_t)nc * ld, V(0));
        for (int i = 0; i < nc; i++)
            for (int j = 0; j < nc; j++) invt[(size_t)i * ld + j] = (V)inv[(size_t)i * nc + j];
        coarse_inv.upload(invt, s);
        CUIBM_HIP(hipStreamSynchronize(s));
    }

    // With the finest level deflated (the FP32 V-cycle), the coarsest operator is still
    // weakly pinned: its inverse has one eigenvalue 1/mu far above the others, whose
    // eigenvector is close to the constant, and an FP32 copy of it (and FP32 level vectors
    // holding its huge component) would drown everything else in rounding. The constant
    // mode is already carried in FP64 by the finest-level deflation, so that component is
    // removed here (Wielandt deflation, in FP64, before the FP32 copy):
    //   inv <- inv - sigma v u^T / (u^T v),
    // with sigma, v, u the dominant eigenvalue and right / left eigenvectors of inv by
    // power iteration (the gap to the next eigenvalue makes a few iterations exact).
    static void deflate_coarse(std::vector<double> &inv, int n) {
        auto power = [&](bool left) {
            std::vector<double> x(n, 1.0), y(n);
            for (int it = 0; it < 60; it++) {
                for (int i = 0; i < n; i++) {
                    double a = 0;
                    for (int j = 0; j < n; j++) a += (left ? inv[(size_t)j * n + i] : inv[(size_t)i * n + j]) * x[j];
                    y[i] = a;
                }
                double nrm = 0;
                for (double v : y) nrm += v * v;
                nrm = std::sqrt(nrm);
                if (!(nrm > 0)) break;
                double diff = 0;
                for (int i = 0; i < n; i++) {
                    const double v = y[i] / nrm;
                    diff = std::max(diff, std::fabs(v - x[i]));
                    x[i] = v;
                }
                if (diff < 1e-15) break;
            }
            return x;
        };
        const std::vector<double> v = power(false), u = power(true);
        double uv = 0, uiv = 0;
        for (int i = 0; i < n; i++) {
            uv += u[i] * v[i];
            double a = 0;
            for (int j = 0; j < n; j++) a += inv[(size_t)i * n + j] * v[j];
            uiv += u[i] * a;
        }
        if (!(std::fabs(uv) > 1e-12)) return;  // no separated mode to remove
        const double sigma = uiv / uv;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) inv[(size_t)i * n + j] -= sigma * v[i] * u[j] / uv;
    }

    // the preconditioner: one V-cycle from x = 0 on the FP64 Krylov vectors b and x;
    // weighted: the deflation's partial sums of b are already in defl_part
    void apply(const double *b, double *x, hipStream_t s, bool weighted = false) {
        const X w = (X)prm.omega;
        if (use_fine && prm.npre == 2 && prm.npost == 2 && lv.size() > 1) {
            // stencil finest level: x2 -> xt and r -> r, restriction, coarse cycle,
            // prolongation into xt, both post-sweeps straight into x
            AmgLevelDev<V, X> &L = lv[0], &C = lv[1];
            if (fine.defl && !weighted) launch_p5_wdot(fine, b, defl_part.p(), s);
            launch_p5_pre(fine, w, b, L.xt.p(), L.r.p(), s, probe);
            launch_csr_op<V, X, OP_SPMV>(L.lanesR, L.R, L.r.p(), nullptr, nullptr, C.b.p(), X(0), s, stop);
            vcycle(1, C.b.p(), C.x.p(), s);
            launch_csr_op<V, X, OP_PROLONG>(L.lanesP, L.P, C.x.p(), nullptr, nullptr, L.xt.p(), X(0), s, stop);
            launch_p5_sweep2(fine, w, L.xt.p(), b, x, s);
            return;
        }
        if constexpr (kVec64) {
            vcycle(0, b, x, s);
        } else {
            AmgLevelDev<V, X> &L = lv[0];
            launch_to_float(L.b.p(), b, L.n, s);
            vcycle(0, L.b.p(), L.x.p(), s);
            launch_to_double(x, L.x.p(), L.n, s);
        }
    }

    // V-cycle on level l with right-hand side b, result into x (both level vectors)
    void vcycle(size_t l, const X *b, X *x, hipStream_t s) {
        AmgLevelDev<V, X> &L = lv[l];
        const int n = L.n;
        const X w = (X)prm.omega;
        if (l + 1 == lv.size()) {
            if ((reinterpret_cast<uintptr_t>(b) & 15) != 0)  // the kernel reads b four values at a time
                throw Error("coarsest AMG right-hand side is not 16-byte aligned", ERR_ARGUMENT);
            hipLaunchKernelGGL((k_dense<V, X>), dim3((n + 3) / 4), dim3(256), 0, s, n, ld, coarse_inv.p(), b, x,
                               stop);
            CUIBM_LAUNCH_CHECK();
            return;
        }
        if (prm.npre == 2 && prm.npost == 2) {
            // x2 -> xt, r -> r, restriction, coarse cycle, prolongation into xt,
            // post-sweeps xt -> r -> x
            AmgLevelDev<V, X> &C = lv[l + 1];
            launch_pre2<V, X>(L.lanesA, L.A, L.dinv.p(), b, L.xt.p(), w, s, stop);
            if (L.G.n > 0) {
                launch_rres<V, X>(L.lanesG, L.G, n, L.xt.p(), b, C.b.p(), s, stop);  // R (b - A x2)
            } else {
                launch_csr_op<V, X, OP_RESID>(L.lanesA, L.A, L.xt.p(), b, nullptr, L.r.p(), X(0), s, stop);
                launch_csr_op<V, X, OP_SPMV>(L.lanesR, L.R, L.r.p(), nullptr, nullptr, C.b.p(), X(0), s, stop);
            }
            vcycle(l + 1, C.b.p(), C.x.p(), s);
            launch_csr_op<V, X, OP_PROLONG>(L.lanesP, L.P, C.x.p(), nullptr, nullptr, L.xt.p(), X(0), s, stop);
            launch_csr_op<V, X, OP_JACOBI>(L.lanesA, L.A, L.xt.p(), b, L.dinv.p(), L.r.p(), w, s, stop);
            launch_csr_op<V, X, OP_JACOBI>(L.lanesA, L.A, L.r.p(), b, L.dinv.p(), x, w, s, stop);
            return;
        }
        const int g = grid_for(n);
        // ping-pong so that the last sweep lands in x
        const int swaps = (prm.npre - 1) + prm.npost;
        X *cur = (swaps % 2 == 0) ? x : L.xt.p();
        X *oth = (cur == x) ? L.xt.p() : x;
        hipLaunchKernelGGL((k_jacobi_first<V, X>), dim3(g), dim3(BS), 0, s, n, L.dinv.p(), b, cur, w, stop);
        CUIBM_LAUNCH_CHECK();
        for (int k = 1; k < prm.npre; k++) {
            launch_csr_op<V, X, OP_JACOBI>(L.lanesA, L.A, cur, b, L.dinv.p(), oth, w, s, stop);
            std::swap(cur, oth);
        }
        launch_csr_op<V, X, OP_RESID>(L.lanesA, L.A, cur, b, nullptr, L.r.p(), X(0), s, stop);
        AmgLevelDev<V, X> &C = lv[l + 1];
        launch_csr_op<V, X, OP_SPMV>(L.lanesR, L.R, L.r.p(), nullptr, nullptr, C.b.p(), X(0), s, stop);
        vcycle(l + 1, C.b.p(), C.x.p(), s);
        launch_csr_op<V, X, OP_PROLONG>(L.lanesP, L.P, C.x.p(), nullptr, nullptr, cur, X(0), s, stop);
        for (int k = 0; k < prm.npost; k++) {
            launch_csr_op<V, X, OP_JACOBI>(L.lanesA, L.A, cur, b, L.dinv.p(), oth, w, s, stop);
            std::swap(cur, oth);
        }
    }
};

// the device copy of the hierarchy's symbolic structure and FP64 values, for numeric
// re-set-up (AmgPrecond::refresh)
struct RefreshLevel {
    int n = 0;
    DArray<int> Arp, Aci, Prp, Pci, Rrp, Rci, agg, perm;
    DArray<int> Prow;  // the row of each slot of P
    // the products A P and R (A P) (the next level's A) as plans: slot q of the product
    // sums X[pl[t].x] * Y[pl[t].y] over t = off[q] .. off[q + 1] - 1, in the host order
    DArray<int> APoff, Coff;
    DArray<int2> APpl, Cpl;
    int nnzAP = 0, nnzC = 0;
    DArray<double> Av, Pv, Rv, APv, tval, d;
};
struct RefreshData {
    bool ready = false;
    std::vector<RefreshLevel> lv;  // every level but the coarsest, then the coarsest (A only)
    DArray<double> rowsum, part, w;  // prolongator smoothing: row sums, partial maxima, omega_p
    DArray<int> flag;
};

struct AmgPrecond::Impl {
    AmgDev<double, double> d64;
    AmgDev<float, double> d3264;
    AmgDev<float, float> d32;
    std::vector<HostAmgLevel> host;
    KernelProbe probe;
    RefreshData rf;
    template <class F>
    void each(F f) {
        f(d64);
        f(d3264);
        f(d32);
    }
};

KernelProbe &AmgPrecond::pre_probe() { return im_->probe; }

static void refresh_setup(RefreshData &rf, const std::vector<HostAmgLevel> &host, hipStream_t s);

AmgPrecond::AmgPrecond(const HCsr &A, const AmgParams &p, hipStream_t s, const P5Dev *fine)
    : im_(new Impl), prm_(p) {
    if (p.precision != 64 && p.precision != 32 && p.precision != 3264)
        throw Error("AMG precision must be 64, 32 or 3264", ERR_ARGUMENT);
    // a hierarchy set up again on the device for moving bodies is built on a padded
    // pattern that the moved body's rows keep to (the padding entries are zeros, so the
    // values, aggregates and levels are those of A's own hierarchy)
    const bool pad = p.refreshable && fine == nullptr && !p.fused_restriction && p.pad_nx > 0;
    HCsr padded;
    if (pad) padded = pad_pattern(A, p.pad_nx, p.pad_margin);
    im_->host = amg_build(pad ? padded : A, p.coarse_max, p.max_levels, p.scale0, p.fused_restriction);
    fine_ = fine != nullptr && im_->host.size() > 1 && p.npre == 2 && p.npost == 2;
    if (fine_) {
        if (fine->nx * fine->ny != A.n) throw Error("stencil and matrix sizes differ", ERR_ARGUMENT);
        fine_dev_ = *fine;
    }
    auto setup = [&](auto &d) {
        d.probe = &im_->probe;
        d.prm = p;
        d.use_fine = fine_;
        if (fine_) d.fine = *fine;
        d.build(im_->host, s);
        if (fine_) fine_dev_.dinv32 = d.fine.dinv32;
    };
    if (p.refreshable && !fine_ && !p.fused_restriction && im_->host.size() >= 2) refresh_setup(im_->rf, im_->host, s);
    if (p.precision == 32 && fine_) {
        // The FP32 V-cycle loses the weakly pinned constant mode of the Poisson operator
        // (A = A0 + E with A0 1 = 0 and E nonzero only at the pinned cells), and BiCGSTAB
        // then needs several times the iterations. The cell areas w make W A0 symmetric, so
        // w^T A0 = 0 and w^T A = (w o a)^T with a = A 1 nonzero only at the pins. Writing
        // r = r' + c a with c = w^T r / w^T a gives A^-1 r = A^-1 r' + c 1, where r' has no
        // component along the constant mode; the V-cycle sees r' and c 1 is added in FP64.
        auto &d = im_->d32;
        const HCsr &H = A;
        auto rowsum = [&](int i) {
            double acc = 0;
            for (int k = H.rp[i]; k < H.rp[i + 1]; k++) acc += H.v[k];
            return acc;
        };
        const int nx = fine->nx, cc = fine->centre;
        auto area = [&](int ip) {
            double dx = 0, dy = 0;
            CUIBM_HIP(hipMemcpy(&dx, fine->dx + ip % nx, sizeof(double), hipMemcpyDeviceToHost));
            CUIBM_HIP(hipMemcpy(&dy, fine->dy + ip / nx, sizeof(double), hipMemcpyDeviceToHost));
            return dx * dy;
        };
        if (!fine->dx || !fine->dy) throw Error("deflation needs the cell widths of the stencil operator", ERR_ARGUMENT);
        CUIBM_HIP(hipStreamSynchronize(s));
        const double a0 = rowsum(0), ac = cc > 0 ? rowsum(cc) : 0.0;
        const double wa = area(0) * a0 + (cc > 0 ? area(cc) * ac : 0.0);
        if (!(wa > 0)) throw Error("the Poisson operator has no pinned cell to deflate", ERR_ARGUMENT);
        d.defl_part.alloc(RED_GRID);
        d.defl_c.alloc(1);
        launch_fill(d.defl_c.p(), 0.0, 1, s);
        d.inv_wa = 1.0 / wa;
        P5Dev f = *fine;
        f.defl = d.defl_c.p();
        f.defl_a0 = a0;
        f.defl_ac = ac;
        f.defl_part = d.defl_part.p();
        f.defl_inv_wa = d.inv_wa;
        d.probe = &im_->probe;
        d.prm = p;
        d.use_fine = true;
        d.fine = f;
        d.build(im_->host, s);
        fine_dev_ = d.fine;
    } else if (p.precision == 32) setup(im_->d32);
    else if (p.precision == 3264) setup(im_->d3264);
    else setup(im_->d64);
    // keep only the sizes of the host hierarchy
    for (auto &L : im_->host) {
        L.A.v.clear(); L.A.v.shrink_to_fit();
        L.P.v.clear(); L.P.v.shrink_to_fit();
        L.R.v.clear(); L.R.v.shrink_to_fit();
        L.G.v.clear(); L.G.v.shrink_to_fit();
    }
}

AmgPrecond::~AmgPrecond() = default;

double AmgPrecond::deflation_inv_wa() const { return fine_dev_.defl ? im_->d32.inv_wa : 0.0; }

void AmgPrecond::set_stop(const double *stop) {
    stop_ = stop;
    im_->each([&](auto &d) {
        d.stop = stop;
        d.fine.stop = stop;
    });
}

void AmgPrecond::apply(const double *r, double *z, hipStream_t s) {
    if (prm_.precision == 32) im_->d32.apply(r, z, s);
    else if (prm_.precision == 3264) im_->d3264.apply(r, z, s);
    else im_->d64.apply(r, z, s);
}

InputWeights AmgPrecond::input_weights() const {
    InputWeights w;
    if (!fine_dev_.defl) return w;
    w.wx = fine_dev_.dx;
    w.wy = fine_dev_.dy;
    w.nx = fine_dev_.nx;
    w.part = im_->d32.defl_part.p();
    return w;
}

void AmgPrecond::apply_weighted(const double *r, double *z, hipStream_t s) {
    if (fine_dev_.defl) im_->d32.apply(r, z, s, true);
    else apply(r, z, s);
}

void AmgPrecond::level0_sweep(const double *b, const double *x, double *xn, hipStream_t s) {
    if (fine_) {
        launch_p5_sweep(fine_dev_, prm_.omega, x, b, xn, s);
    } else if (prm_.precision == 64) {
        auto &L = im_->d64.lv[0];
        launch_csr_op<double, double, OP_JACOBI>(L.lanesA, L.A, x, b, L.dinv.p(), xn, prm_.omega, s);
    } else if (prm_.precision == 3264) {  // the FP32-valued sweep on FP64 vectors
        auto &L = im_->d3264.lv[0];
        launch_csr_op<float, double, OP_JACOBI>(L.lanesA, L.A, x, b, L.dinv.p(), xn, prm_.omega, s);
    } else {
        throw Error("level-0 sweep timing of the FP32 V-cycle needs its stencil finest level", ERR_UNSUPPORTED);
    }
}

// bytes of a value of the matrices (sv) and of a level vector (sx)
static double value_bytes(int precision) { return precision == 64 ? 8.0 : 4.0; }
static double vector_bytes(int precision) { return precision == 32 ? 4.0 : 8.0; }

double AmgPrecond::level0_sweep_bytes() const {
    const double sv = value_bytes(prm_.precision);
    const auto &A = im_->host[0].A;
    const double n = A.n, nnz = (double)A.ci.size();
    if (fine_) return n * (8 + 8 + 8) + 32.0 * (fine_dev_.nx + fine_dev_.ny);  // x, b, xn, coefficient arrays
    return nnz * (sv + 4) + 4 * (n + 1) + n * (8 + 8 + 8 + sv);  // A, rowptr, x, b, xn, dinv
}

void AmgPrecond::level0_pre(const double *b, double *x2, double *r, hipStream_t s) {
    if (!fine_) throw Error("the finest AMG level is not in stencil form", ERR_ARGUMENT);
    if (prm_.precision == 32)  // FP32 level vectors in the caller's (larger) buffers
        launch_p5_pre(fine_dev_, (float)prm_.omega, b, reinterpret_cast<float *>(x2), reinterpret_cast<float *>(r),
                      s);
    else
        launch_p5_pre(fine_dev_, prm_.omega, b, x2, r, s);
}

void AmgPrecond::level0_post(double *x2, const double *b, double *x, hipStream_t s) {
    if (!fine_) throw Error("the finest AMG level is not in stencil form", ERR_ARGUMENT);
    if (prm_.precision == 32) {
        auto &L = im_->d32.lv[0];
        float *x2f = reinterpret_cast<float *>(x2);
        launch_csr_op<float, float, OP_PROLONG>(L.lanesP, L.P, im_->d32.lv[1].x.p(), nullptr, nullptr, x2f, 0.0f, s);
        launch_p5_sweep2(fine_dev_, (float)prm_.omega, x2f, b, x, s);
    } else if (prm_.precision == 3264) {
        auto &L = im_->d3264.lv[0];
        launch_csr_op<float, double, OP_PROLONG>(L.lanesP, L.P, im_->d3264.lv[1].x.p(), nullptr, nullptr, x2, 0.0, s);
        launch_p5_sweep2(fine_dev_, prm_.omega, x2, b, x, s);
    } else {
        auto &L = im_->d64.lv[0];
        launch_csr_op<double, double, OP_PROLONG>(L.lanesP, L.P, im_->d64.lv[1].x.p(), nullptr, nullptr, x2, 0.0, s);
        launch_p5_sweep2(fine_dev_, prm_.omega, x2, b, x, s);
    }
}

double AmgPrecond::level0_pre_bytes() const {
    const double sx = vector_bytes(prm_.precision);
    // b (the FP64 Krylov vector) and D^-1 in; x2, r out
    return (8.0 + 3 * sx) * im_->host[0].A.n + 32.0 * (fine_dev_.nx + fine_dev_.ny);
}

double AmgPrecond::level0_post_bytes() const {
    const double sv = value_bytes(prm_.precision), sx = vector_bytes(prm_.precision);
    const double n = im_->host[0].A.n, nc = im_->host[1].A.n, nnzP = (double)im_->host[0].P.ci.size();
    const double prolong = nnzP * (sv + 4) + 4 * (n + 1) + nc * sx + 2 * n * sx;  // P, xc; x2 read-modify-write
    // x2 and D^-1 in as level vectors, b in and the result out as FP64 Krylov vectors
    const double sweeps = (2 * sx + 16.0) * n + 32.0 * (fine_dev_.nx + fine_dev_.ny);
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

std::vector<double> AmgPrecond::launch_bytes() const {
    // distinct arrays per kernel; matrix values and D^-1 take sv bytes, level vectors sx:
    // CSR product = nnz*(sv+4) + 4(n+1) + sx ncol (x) + sx n (y)
    const double sv = value_bytes(prm_.precision), sx = vector_bytes(prm_.precision);
    std::vector<double> L;
    const auto &H = im_->host;
    auto spmv = [&](double n, double ncol, double nnz) { return nnz * (sv + 4) + 4 * (n + 1) + ncol * sx + n * sx; };
    for (size_t l = 0; l + 1 < H.size(); l++) {
        const double n = H[l].A.n, nnz = (double)H[l].A.ci.size();
        const double nc = H[l + 1].A.n, nnzP = (double)H[l].P.ci.size();
        if (l == 0 && fine_) {
            // (with deflation, the Krylov kernels that produce b also form its weighted partial
            // sums, and the pre-smoothing kernel finishes them: no launch of its own)
            L.push_back(level0_pre_bytes());
            L.push_back(spmv(nc, n, nnzP));                                          // restriction
            L.push_back(nnzP * (sv + 4) + 4 * (n + 1) + nc * sx + 2 * n * sx);        // prolong-add
            L.push_back((2 * sx + 16.0) * n + 32.0 * (fine_dev_.nx + fine_dev_.ny));  // both post-sweeps
            continue;
        }
        if (l == 0 && prm_.precision == 32) {  // FP64 <-> FP32 conversions of b and x
            L.push_back(n * (8 + sx));
            L.push_back(n * (8 + sx));
        }
        if (prm_.npre == 2 && prm_.npost == 2) {  // the fused (2,2) sequence of vcycle()
            const double nnzG = (double)H[l].G.ci.size();
            L.push_back(spmv(n, n, nnz) + n * sv + n * sx);  // both pre-sweeps (+ dinv, b)
            if (nnzG > 0) {
                L.push_back(nnzG * (sv + 4) + 4 * (nc + 1) + 2 * n * sx + nc * sx);  // R (b - A x2) via G
            } else {
                L.push_back(spmv(n, n, nnz) + n * sx);  // residual (+ b)
                L.push_back(spmv(nc, n, nnzP));         // restriction
            }
            L.push_back(spmv(n, nc, nnzP) + n * sx);  // prolong-add
            for (int k = 0; k < 2; k++) L.push_back(spmv(n, n, nnz) + n * sv + n * sx);  // post-sweeps (+ dinv, b)
            continue;
        }
        L.push_back(n * sv + 2 * n * sx);  // first sweep: dinv, b, x
        const int sweeps = prm_.npre - 1 + prm_.npost;
        for (int k = 0; k < sweeps; k++) L.push_back(spmv(n, n, nnz) + n * sv + n * sx);  // + dinv, b
        L.push_back(spmv(n, n, nnz) + n * sx);  // residual (+ b)
        L.push_back(spmv(nc, n, nnzP));         // restriction
        L.push_back(spmv(n, nc, nnzP) + n * sx);  // prolong-add (read-modify-write x)
    }
    const double nc = H.back().A.n;
    L.push_back(nc * nc * sv + 2 * nc * sx);
    return L;
}

double AmgPrecond::bytes_per_apply() const {
    double bytes = 0;
    for (double b : launch_bytes()) bytes += b;
    return bytes;
}

// ---- numeric re-set-up of the hierarchy on the device (same pattern, new values) ----

namespace {

// A's entries into the kept finest pattern S (zeros elsewhere), one row per thread;
// flag = 1 where an entry of A has no slot in S (columns of S rows are sorted)
__global__ void k_rf_scatter(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                             const double *__restrict__ v, const int *__restrict__ Srp, const int *__restrict__ Sci,
                             double *__restrict__ Sv, int *__restrict__ flag) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int s0 = Srp[i], s1 = Srp[i + 1];
        for (int q = s0; q < s1; q++) Sv[q] = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; k++) {
            const int c = ci[k];
            int lo = s0, hi = s1 - 1, at = -1;
            while (lo <= hi) {
                const int mid = (lo + hi) >> 1;
                const int m = Sci[mid];
                if (m == c) {
                    at = mid;
                    break;
                }
                if (m < c) lo = mid + 1;
                else hi = mid - 1;
            }
            if (at < 0) flag[0] = 1;
            else Sv[at] += v[k];
        }
    }
}

// Inverse of the coarsest operator (n <= RF_GJ_MAX) by Gauss-Jordan elimination with
// partial pivoting in one workgroup, M and X = I in LDS: every value goes through
// dense_inverse's operations in its order (the pivot is the first row of largest
// magnitude; entries left of and at the pivot column are not read again and are not
// updated). flag = 1 on a zero pivot.
constexpr int RF_GJ_MAX = 60;
template <class V>
__global__ void __launch_bounds__(1024) k_rf_gj(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                                const double *__restrict__ v, V *__restrict__ inv, int ld,
                                                int *__restrict__ flag) {
    __shared__ double M[RF_GJ_MAX * RF_GJ_MAX], X[RF_GJ_MAX * RF_GJ_MAX];
    __shared__ int piv_s;
    const int t = threadIdx.x, nt = blockDim.x, nn = n * n;
    for (int e = t; e < nn; e += nt) {
        M[e] = 0.0;
        X[e] = (e / n == e % n) ? 1.0 : 0.0;
    }
    __syncthreads();
    if (t < n)
        for (int k = rp[t]; k < rp[t + 1]; k++) M[t * n + ci[k]] += v[k];
    __syncthreads();
    for (int c = 0; c < n; c++) {
        if (t < 64) {  // wave 0: lanes hold rows c .. c + 63 (n - c <= 60)
            const int r = c + t;
            double a = r < n ? fabs(M[r * n + c]) : -1.0;
            int idx = r < n ? r : n;
            for (int o = 32; o > 0; o >>= 1) {
                const double a2 = __shfl_xor(a, o);
                const int i2 = __shfl_xor(idx, o);
                if (a2 > a || (a2 == a && i2 < idx)) {
                    a = a2;
                    idx = i2;
                }
            }
            if (t == 0) {
                piv_s = idx;
                if (M[idx * n + c] == 0.0) flag[0] = 1;
            }
        }
        __syncthreads();
        const int piv = piv_s;
        if (piv != c && t < 2 * n) {
            double *A = t < n ? M : X;
            const int k = t < n ? t : t - n;
            const double x = A[c * n + k];
            A[c * n + k] = A[piv * n + k];
            A[piv * n + k] = x;
        }
        __syncthreads();
        const double iv = 1.0 / M[c * n + c];
        if (t < 2 * n) {
            const int k = t < n ? t : t - n;
            if (t >= n) X[c * n + k] = X[c * n + k] * iv;
            else if (k > c) M[c * n + k] = M[c * n + k] * iv;
        }
        __syncthreads();
        for (int e = t; e < 2 * nn; e += nt) {
            const bool isx = e >= nn;
            const int ee = isx ? e - nn : e;
            const int r = ee / n, k = ee % n;
            if (r == c || (!isx && k <= c)) continue;
            const double f = M[r * n + c];
            if (f == 0.0) continue;
            if (isx) X[r * n + k] = X[r * n + k] - f * X[c * n + k];
            else M[r * n + k] = M[r * n + k] - f * M[c * n + k];
        }
        __syncthreads();
    }
    for (int e = t; e < nn; e += nt) inv[(size_t)(e / n) * ld + e % n] = (V)X[e];
}

// d_i = sum of the diagonal entries of row i; s_i = sum_k |a_ik (1 / d_i)| (amg_build's order)
__global__ void k_rf_diag(int n, const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ v,
                          double *__restrict__ d, double *__restrict__ rs) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double di = 0;
        for (int k = rp[i]; k < rp[i + 1]; k++)
            if (ci[k] == i) di += v[k];
        double t = 0;
        for (int k = rp[i]; k < rp[i + 1]; k++) t += fabs(v[k] * (1.0 / di));
        d[i] = di;
        rs[i] = t;
    }
}

// w = 4 / (3 max_i s_i): partial maxima over a fixed grid, then one block
__global__ void __launch_bounds__(BS) k_rf_max(int n, const double *__restrict__ rs, double *__restrict__ part) {
    __shared__ double sh[BS];
    double m = 0;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) m = fmax(m, rs[i]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int o = BS / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ void __launch_bounds__(BS) k_rf_omega(int nb, const double *__restrict__ part, double *__restrict__ w) {
    __shared__ double sh[BS];
    double m = 0;
    for (int b = threadIdx.x; b < nb; b += BS) m = fmax(m, part[b]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int o = BS / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) w[0] = 4.0 / (3.0 * sh[0]);
}

// P = (I - w D^-1 A) T on P's cached pattern, one thread per slot: for slot (i, c),
// (A T)_ic summed in the order of row i of A, then -w (1 / d_i) (A T)_ic, plus t_i where c
// is row i's own aggregate
__global__ void k_rf_p(int nnz, const int *__restrict__ Prow, const int *__restrict__ Pci,
                       const int *__restrict__ Arp, const int *__restrict__ Aci, const double *__restrict__ Av,
                       const int *__restrict__ agg, const double *__restrict__ tval, const double *__restrict__ d,
                       const double *__restrict__ w, double *__restrict__ Pv) {
    const double wp = w[0];
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nnz; q += gridDim.x * blockDim.x) {
        const int i = Prow[q], c = Pci[q];
        double at = 0;
        bool any = false;
        for (int k = Arp[i]; k < Arp[i + 1]; k++) {
            const int j = Aci[k];
            if (agg[j] == c) {
                at += Av[k] * tval[j];
                any = true;
            }
        }
        if (!any) {  // only the tentative entry
            Pv[q] = tval[i];
            continue;
        }
        const double val = -wp * (1.0 / d[i]) * at;
        Pv[q] = c == agg[i] ? tval[i] + val : val;
    }
}

// R = P^T through the cached permutation, scaled by 1 / d_col on a row-scaled finest level
__global__ void k_rf_r(int nnz, const int *__restrict__ perm, const double *__restrict__ Pv,
                       const int *__restrict__ Rci, const double *__restrict__ d, double *__restrict__ Rv) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nnz; q += gridDim.x * blockDim.x) {
        const double v = Pv[perm[q]];
        Rv[q] = d ? v * (1.0 / d[Rci[q]]) : v;
    }
}

// Z = X Y on Z's cached pattern through its plan, one thread per slot: the products of
// the slot's (X, Y) entry pairs are summed in the order the host product accumulates them
__global__ void k_rf_plan(int nnz, const int *__restrict__ off, const int2 *__restrict__ pl,
                          const double *__restrict__ Xv, const double *__restrict__ Yv, double *__restrict__ Zv) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nnz; q += gridDim.x * blockDim.x) {
        double acc = 0;
        for (int t = off[q]; t < off[q + 1]; t++) {
            const int2 e = pl[t];
            acc += Xv[e.x] * Yv[e.y];
        }
        Zv[q] = acc;
    }
}

template <class V>
__global__ void k_rf_cast(long n, const double *__restrict__ x, V *__restrict__ y) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = (V)x[i];
}
template <class V>
__global__ void k_rf_dinv(int n, const double *__restrict__ d, V *__restrict__ dinv) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dinv[i] = (V)(1.0 / d[i]);
}

int rf_blocks(long n) { return (int)std::max(1L, std::min((n + BS - 1) / BS, 4096L)); }

template <class T>
void upload_v(DArray<T> &d, const std::vector<T> &h, hipStream_t s) {
    d.upload(h.empty() ? std::vector<T>{T(0)} : h, s);
}

std::vector<int> slot_rows(const std::vector<int> &rp) {
    std::vector<int> r(rp.empty() ? 0 : rp.back());
    for (size_t i = 0; i + 1 < rp.size(); i++)
        for (int q = rp[i]; q < rp[i + 1]; q++) r[q] = (int)i;
    return r;
}

// the plan of Z = X Y on Z's pattern (multiply()'s): for each slot of Z, the pairs (entry
// of X row i, entry of Y row k) contributing to it, in X row order as multiply() sums them
void product_plan(const HCsr &X, const HCsr &Y, const std::vector<int> &Zrp, const std::vector<int> &Zci,
                  std::vector<int> &off, std::vector<int2> &pl) {
    const size_t nnz = Zci.size();
    std::vector<int> pos(std::max(Y.ncols, 1), -1), cnt(nnz + 1, 0);
    auto each = [&](auto f) {
        for (int i = 0; i < X.n; i++) {
            for (int q = Zrp[i]; q < Zrp[i + 1]; q++) pos[Zci[q]] = q;
            for (int ka = X.rp[i]; ka < X.rp[i + 1]; ka++) {
                const int k = X.ci[ka];
                for (int kb = Y.rp[k]; kb < Y.rp[k + 1]; kb++) f(pos[Y.ci[kb]], ka, kb);
            }
            for (int q = Zrp[i]; q < Zrp[i + 1]; q++) pos[Zci[q]] = -1;
        }
    };
    each([&](int q, int, int) {
        if (q < 0) throw Error("AMG refresh: product entry outside the kept pattern", ERR_GENERIC);
        cnt[q + 1]++;
    });
    for (size_t q = 0; q < nnz; q++) cnt[q + 1] += cnt[q];
    off = cnt;
    pl.resize(std::max<size_t>(off.back(), 1));
    std::vector<int> next(off.begin(), off.end() - 1);
    each([&](int q, int ka, int kb) { pl[next[q]++] = make_int2(ka, kb); });
}

}  // namespace

// the symbolic structure on the device, from the host hierarchy's patterns
static void refresh_setup(RefreshData &rf, const std::vector<HostAmgLevel> &host, hipStream_t s) {
    const int L = (int)host.size();
    rf.lv.clear();
    rf.lv.resize(L);
    for (int l = 0; l < L; l++) {
        const HostAmgLevel &H = host[l];
        RefreshLevel &R = rf.lv[l];
        R.n = H.A.n;
        upload_v(R.Arp, H.A.rp, s);
        upload_v(R.Aci, H.A.ci, s);
        R.Av.alloc(std::max<size_t>(H.A.ci.size(), 1));
        R.d.alloc(std::max(H.A.n, 1));
        if (l + 1 == L) continue;
        upload_v(R.Prp, H.P.rp, s);
        upload_v(R.Pci, H.P.ci, s);
        upload_v(R.Prow, slot_rows(H.P.rp), s);
        upload_v(R.Rrp, H.R.rp, s);
        upload_v(R.Rci, H.R.ci, s);
        upload_v(R.agg, H.agg, s);
        std::vector<int> size(H.P.ncols, 0);
        for (int a : H.agg) size[a]++;
        std::vector<double> t(H.A.n);
        for (int i = 0; i < H.A.n; i++) t[i] = 1.0 / std::sqrt((double)size[H.agg[i]]);
        upload_v(R.tval, t, s);
        // R = P^T: the P entry behind each R entry, in the host transpose's order
        std::vector<int> next(H.R.rp.begin(), H.R.rp.end() - 1), perm(H.P.ci.size());
        for (int i = 0; i < H.P.n; i++)
            for (int k = H.P.rp[i]; k < H.P.rp[i + 1]; k++) perm[next[H.P.ci[k]]++] = k;
        upload_v(R.perm, perm, s);
        std::vector<int> aprp, apci, off;
        std::vector<int2> pl;
        multiply_pattern(H.A, H.P, aprp, apci);
        R.nnzAP = (int)apci.size();
        product_plan(H.A, H.P, aprp, apci, off, pl);
        upload_v(R.APoff, off, s);
        upload_v(R.APpl, pl, s);
        HCsr AP;  // the pattern of A P, as the right operand of R (A P)
        AP.n = H.A.n;
        AP.ncols = H.P.ncols;
        AP.rp = aprp;
        AP.ci = apci;
        const HCsr &Ac = host[l + 1].A;
        R.nnzC = (int)Ac.ci.size();
        product_plan(H.R, AP, Ac.rp, Ac.ci, off, pl);
        upload_v(R.Coff, off, s);
        upload_v(R.Cpl, pl, s);
        R.Pv.alloc(std::max<size_t>(H.P.ci.size(), 1));
        R.Rv.alloc(std::max<size_t>(H.R.ci.size(), 1));
        R.APv.alloc(std::max<size_t>(apci.size(), 1));
    }
    const int n0 = host[0].A.n;
    rf.rowsum.alloc(std::max(n0, 1));
    rf.part.alloc(4096);
    rf.w.alloc(1);
    rf.flag.alloc(2);
    rf.ready = true;
}

template <class V, class X>
static void refresh_levels(AmgDev<V, X> &dv, RefreshData &rf, const std::vector<HostAmgLevel> &host, bool scale0,
                           hipStream_t s) {
    const int L = (int)host.size();
    for (int l = 0; l + 1 < L; l++) {
        RefreshLevel &R = rf.lv[l], &C = rf.lv[l + 1];
        const HostAmgLevel &H = host[l];
        const int n = R.n;
        hipLaunchKernelGGL(k_rf_diag, dim3(rf_blocks(n)), dim3(BS), 0, s, n, R.Arp.p(), R.Aci.p(), R.Av.p(), R.d.p(),
                           rf.rowsum.p());
        const int nb = std::min(rf_blocks(n), 4096);
        hipLaunchKernelGGL(k_rf_max, dim3(nb), dim3(BS), 0, s, n, rf.rowsum.p(), rf.part.p());
        hipLaunchKernelGGL(k_rf_omega, dim3(1), dim3(BS), 0, s, nb, rf.part.p(), rf.w.p());
        const int nnzP0 = (int)H.P.ci.size();
        hipLaunchKernelGGL(k_rf_p, dim3(rf_blocks(nnzP0)), dim3(BS), 0, s, nnzP0, R.Prow.p(), R.Pci.p(), R.Arp.p(),
                           R.Aci.p(), R.Av.p(), R.agg.p(), R.tval.p(), R.d.p(), rf.w.p(), R.Pv.p());
        const int nnzR = (int)H.R.ci.size();
        hipLaunchKernelGGL(k_rf_r, dim3(rf_blocks(nnzR)), dim3(BS), 0, s, nnzR, R.perm.p(), R.Pv.p(), R.Rci.p(),
                           (scale0 && l == 0) ? R.d.p() : nullptr, R.Rv.p());
        hipLaunchKernelGGL(k_rf_plan, dim3(rf_blocks(R.nnzAP)), dim3(BS), 0, s, R.nnzAP, R.APoff.p(), R.APpl.p(),
                           R.Av.p(), R.Pv.p(), R.APv.p());
        hipLaunchKernelGGL(k_rf_plan, dim3(rf_blocks(R.nnzC)), dim3(BS), 0, s, R.nnzC, R.Coff.p(), R.Cpl.p(),
                           R.Rv.p(), R.APv.p(), C.Av.p());
        // the device hierarchy in its value type
        AmgLevelDev<V, X> &D = dv.lv[l];
        const long nnzA = (long)H.A.ci.size(), nnzP = (long)H.P.ci.size();
        hipLaunchKernelGGL((k_rf_cast<V>), dim3(rf_blocks(nnzA)), dim3(BS), 0, s, nnzA, R.Av.p(), D.A.v.p());
        hipLaunchKernelGGL((k_rf_dinv<V>), dim3(rf_blocks(n)), dim3(BS), 0, s, n, R.d.p(), D.dinv.p());
        hipLaunchKernelGGL((k_rf_cast<V>), dim3(rf_blocks(nnzP)), dim3(BS), 0, s, nnzP, R.Pv.p(), D.P.v.p());
        hipLaunchKernelGGL((k_rf_cast<V>), dim3(rf_blocks(nnzR)), dim3(BS), 0, s, (long)nnzR, R.Rv.p(), D.R.v.p());
        CUIBM_LAUNCH_CHECK();
    }
    RefreshLevel &Cl = rf.lv[L - 1];
    if (dv.nc <= RF_GJ_MAX && !dv.fine.defl) {
        hipLaunchKernelGGL((k_rf_gj<V>), dim3(1), dim3(1024), 0, s, Cl.n, Cl.Arp.p(), Cl.Aci.p(), Cl.Av.p(),
                           dv.coarse_inv.p(), dv.ld, rf.flag.p() + 1);
        CUIBM_LAUNCH_CHECK();
        return;
    }
    // a larger coarsest level: its dense inverse on the host, as build() forms it
    HCsr Ac;
    Ac.n = Cl.n;
    Ac.ncols = Cl.n;
    Ac.rp = host[L - 1].A.rp;
    Ac.ci = host[L - 1].A.ci;
    Ac.v.resize(Ac.ci.size());
    if (!Ac.v.empty()) Cl.Av.download(Ac.v.data(), Ac.v.size(), s);
    CUIBM_HIP(hipStreamSynchronize(s));
    std::vector<double> inv = dense_inverse(Ac);
    if (dv.fine.defl) AmgDev<V, X>::deflate_coarse(inv, dv.nc);
    std::vector<V> invt((size_t)dv.nc * dv.ld, V(0));
    for (int i = 0; i < dv.nc; i++)
        for (int j = 0; j < dv.nc; j++) invt[(size_t)i * dv.ld + j] = (V)inv[(size_t)i * dv.nc + j];
    dv.coarse_inv.upload(invt, s);
    CUIBM_HIP(hipStreamSynchronize(s));
}

bool AmgPrecond::refresh(const DCsr<double> &A, hipStream_t s) {
    if (fine_) return false;  // a stencil finest level is rebuilt from its operator instead
    RefreshData &rf = im_->rf;
    if (!rf.ready || A.n != rf.lv[0].n) return false;
    launch_fill_i(rf.flag.p(), 0, 2, s);
    hipLaunchKernelGGL(k_rf_scatter, dim3(rf_blocks(A.n)), dim3(BS), 0, s, A.n, A.rp.p(), A.ci.p(), A.v.p(),
                       rf.lv[0].Arp.p(), rf.lv[0].Aci.p(), rf.lv[0].Av.p(), rf.flag.p());
    CUIBM_LAUNCH_CHECK();
    int flags[2] = {0, 0};
    rf.flag.download(flags, 1, s);
    if (flags[0]) return false;
    if (prm_.precision == 32) refresh_levels(im_->d32, rf, im_->host, prm_.scale0, s);
    else if (prm_.precision == 3264) refresh_levels(im_->d3264, rf, im_->host, prm_.scale0, s);
    else refresh_levels(im_->d64, rf, im_->host, prm_.scale0, s);
    rf.flag.download(flags, 2, s);
    if (flags[1]) throw Error("singular coarsest AMG operator", ERR_GENERIC);
    return true;
}

// ---- lockstep batches: the V-cycles of many cases as one preconditioner ----
//
// Cases are ordered by hierarchy depth (deepest first), so the cases present on level l
// are a prefix of the case list, and within it the cases that still smooth on level l
// (depth > l + 1) come before those whose coarsest (dense) level is l. Every case segment
// is padded to a multiple of BATCH_PAD rows (empty CSR rows, zero D^-1), so the
// concatenated block-diagonal CSR operators run through the single-case kernels and
// each dense solve reads aligned vectors.

namespace {

template <class V>
struct HCat {
    int n = 0;
    std::vector<int> rp{0}, ci;
    std::vector<V> v;
    // append the rows of src (columns mapped by col), then empty rows up to `rows`
    template <class F>
    void append(const std::vector<int> &srp, const std::vector<int> &sci, const std::vector<V> &sv, int rows, F col) {
        const int sn = (int)srp.size() - 1;
        for (int i = 0; i < sn; i++) {
            for (int k = srp[i]; k < srp[i + 1]; k++) {
                ci.push_back(col(sci[k]));
                v.push_back(sv[k]);
            }
            rp.push_back((int)ci.size());
        }
        for (int i = sn; i < rows; i++) rp.push_back((int)ci.size());
        n += rows;
    }
    void upload(DCsr<V> &d, int ncols, hipStream_t s) const {
        // every row pointer and column inside the operator's shape (host check before any launch)
        if ((int)rp.size() != n + 1 || rp.back() != (int)ci.size())
            throw Error("batched V-cycle: inconsistent concatenated row pointers", ERR_GENERIC);
        for (int i = 0; i < n; i++)
            if (rp[i + 1] < rp[i]) throw Error("batched V-cycle: decreasing row pointers", ERR_GENERIC);
        for (int j : ci)
            if (j < 0 || j >= ncols) throw Error("batched V-cycle: a column outside the operator", ERR_GENERIC);
        d.n = n;
        d.ncols = ncols;
        d.nnz = (int)ci.size();
        d.rp.upload(rp, s);
        d.ci.upload(ci.empty() ? std::vector<int>{0} : ci, s);
        d.v.upload(v.empty() ? std::vector<V>{V(0)} : v, s);
    }
    int lanes() const {
        HCsr h;
        h.n = n;
        h.ci.assign(ci.size(), 0);
        return lanes_for(h);
    }
};

}  // namespace

template <class V, class X>
struct BatchDev {
    struct Lvl {
        int n = 0, nS = 0;  // padded rows of the present cases, and of the smoothing prefix
        DCsr<V> A, G, P, R;
        int lanesA = 1, lanesG = 1, lanesP = 1, lanesR = 1;
        DArray<V> dinv;
        DArray<X> b, x, xt, r;
        DArray<DenseCase<V>> dense;
        BMap dmap;
        DArray<int> seg;  // the case of each BATCH_PAD-row segment of the level vectors
    };
    std::vector<Lvl> lv;  // lv[0]: the finest level (P0, R0 and its two vectors)
    int ncases = 0;
    const BCase *tab = nullptr;
    BMapDev pmap{}, tiles{};
    const int *pfirst = nullptr;
    BMap tilemap;
    DArray<double> part, inv_wa;
    double *defl = nullptr;
    bool deflate = false;
    double omega = 0.8;
    DArray<const double *> cstop;  // each case's Poisson stop code: the CSR levels skip stopped cases

    void build(const std::vector<AmgDev<V, X> *> &cd, const std::vector<long> &offP,
               const std::vector<const double *> &stops, hipStream_t s) {
        ncases = (int)cd.size();
        int depth = 0;
        for (auto *d : cd) depth = std::max(depth, (int)d->lv.size());
        for (auto *d : cd)
            if (d->lv.size() < 2 || !d->use_fine) throw Error("batched V-cycle: every case needs a stencil finest level and a coarser one", ERR_UNSUPPORTED);
        for (int c = 1; c < ncases; c++)
            if (cd[c]->lv.size() > cd[c - 1]->lv.size()) throw Error("batched V-cycle: cases must come deepest first", ERR_ARGUMENT);
        lv.resize(depth);
        // segment offsets per level: level 0 uses the pressure offsets
        std::vector<std::vector<long>> off(depth, std::vector<long>(ncases, -1));
        for (int c = 0; c < ncases; c++) off[0][c] = offP[c];
        lv[0].n = (int)(offP.back() + batch_pad(cd.back()->lv[0].n));
        for (int l = 1; l < depth; l++) {
            long o = 0;
            for (int c = 0; c < ncases; c++)
                if ((int)cd[c]->lv.size() > l) {
                    off[l][c] = o;
                    o += batch_pad(cd[c]->lv[l].n);
                }
            lv[l].n = (int)o;
        }
        cstop.upload(std::vector<const double *>(stops.begin(), stops.end()), s);
        for (int l = 0; l < depth; l++) {
            std::vector<int> seg;
            for (int c = 0; c < ncases; c++)
                if ((int)cd[c]->lv.size() > l) seg.insert(seg.end(), batch_pad(cd[c]->lv[l].n) / BATCH_PAD, c);
            if ((long)seg.size() * BATCH_PAD != lv[l].n) throw Error("batched V-cycle: level segments", ERR_GENERIC);
            lv[l].seg.upload(seg.empty() ? std::vector<int>{0} : seg, s);
        }
        for (int l = 0; l < depth; l++) {
            Lvl &L = lv[l];
            HCat<V> A, G, P, R;
            std::vector<V> dinv;
            std::vector<DenseCase<V>> dense;
            std::vector<long> dcount;
            std::vector<int> dcase;
            int nS = 0;
            for (int c = 0; c < ncases; c++) {
                AmgDev<V, X> &d = *cd[c];
                if ((int)d.lv.size() <= l) continue;
                const int rows = (int)batch_pad(d.lv[l].n);
                if ((int)d.lv.size() - 1 == l) {  // this case's coarsest level
                    dense.push_back(DenseCase<V>{d.coarse_inv.p(), d.nc, d.ld, off[l][c], stops[c]});
                    dcount.push_back(d.nc);
                    continue;
                }
                nS = (int)(off[l][c] + rows);
                AmgLevelDev<V, X> &S = d.lv[l];
                const long o = off[l][c], oc = off[l + 1][c];
                if (l > 0) {
                    if (S.G.n == 0) throw Error("batched V-cycle: a CSR level without its fused restriction", ERR_UNSUPPORTED);
                    A.append(S.A.rp.to_host(s), S.A.ci.to_host(s), S.A.v.to_host(s), rows, [&](int j) { return (int)(o + j); });
                    std::vector<V> dv = S.dinv.to_host(s);
                    dv.resize(rows, V(0));
                    dinv.insert(dinv.end(), dv.begin(), dv.end());
                }
                const int crow = (int)batch_pad(d.lv[l + 1].n);
                P.append(S.P.rp.to_host(s), S.P.ci.to_host(s), S.P.v.to_host(s), rows, [&](int j) { return (int)(oc + j); });
                R.append(S.R.rp.to_host(s), S.R.ci.to_host(s), S.R.v.to_host(s), crow, [&](int j) { return (int)(o + j); });
                if (l > 0) {
                    const int nf = S.n;
                    G.append(S.G.rp.to_host(s), S.G.ci.to_host(s), S.G.v.to_host(s), crow, [&, nf](int j) {
                        return j < nf ? (int)(o + j) : (int)(-1 - (o + j - nf));  // x part: marked, shifted below
                    });
                }
            }
            L.nS = nS;
            // the x part of G's columns goes after the nS b columns
            for (int &j : G.ci)
                if (j < 0) j = nS + (-1 - j);
            if (nS > 0) {
                if (l > 0) {
                    A.upload(L.A, L.n, s);
                    L.lanesA = A.lanes();
                    G.upload(L.G, 2 * nS, s);
                    L.lanesG = G.lanes();
                    L.dinv.upload(dinv, s);
                }
                P.upload(L.P, lv[l + 1].n, s);
                L.lanesP = P.lanes();
                R.upload(L.R, nS, s);
                L.lanesR = R.lanes();
            }
            if (!dense.empty()) {
                L.dense.upload(dense, s);
                L.dmap.build(dcount, 4, s);
            }
            if (l > 0) {
                for (auto *a : {&L.b, &L.x, &L.xt, &L.r}) {
                    a->alloc(std::max(L.n, 1));
                    launch_fill_x(a->p(), L.n, s);
                }
            } else {
                for (auto *a : {&L.xt, &L.r}) {
                    a->alloc(std::max(L.n, 1));
                    launch_fill_x(a->p(), L.n, s);
                }
            }
        }
        CUIBM_HIP(hipStreamSynchronize(s));
    }

    static void launch_fill_x(float *p, long n, hipStream_t s) { if (n) launch_fill_f(p, 0.0f, n, s); }
    static void launch_fill_x(double *p, long n, hipStream_t s) { if (n) launch_fill(p, 0.0, n, s); }

    void vcycle(size_t l, hipStream_t s) {
        Lvl &L = lv[l];
        const X w = (X)omega;
        if (L.nS > 0) {
            Lvl &C = lv[l + 1];
            const double *const *cs = cstop.p();
            launch_pre2<V, X>(L.lanesA, L.A, L.dinv.p(), L.b.p(), L.xt.p(), w, s, nullptr, L.seg.p(), cs);
            launch_rres<V, X>(L.lanesG, L.G, L.nS, L.xt.p(), L.b.p(), C.b.p(), s, nullptr, C.seg.p(), cs);
            vcycle(l + 1, s);
            launch_csr_op<V, X, OP_PROLONG>(L.lanesP, L.P, C.x.p(), nullptr, nullptr, L.xt.p(), X(0), s, nullptr,
                                            L.seg.p(), cs);
            launch_csr_op<V, X, OP_JACOBI>(L.lanesA, L.A, L.xt.p(), L.b.p(), L.dinv.p(), L.r.p(), w, s, nullptr,
                                           L.seg.p(), cs);
            launch_csr_op<V, X, OP_JACOBI>(L.lanesA, L.A, L.r.p(), L.b.p(), L.dinv.p(), L.x.p(), w, s, nullptr,
                                           L.seg.p(), cs);
        }
        if (L.dmap.nblk() > 0) {
            hipLaunchKernelGGL((k_b_dense<V, X>), dim3(L.dmap.nblk()), dim3(256), 0, s, L.dense.p(), L.dmap.dev(),
                               L.b.p(), L.x.p());
            CUIBM_LAUNCH_CHECK();
        }
    }

    void apply(const double *b, double *x, hipStream_t s) {
        Lvl &F = lv[0], &C = lv[1];
        const X w = (X)omega;
        if (deflate) {
            launch_b_wdot(tab, pmap, b, part.p(), s);
            launch_b_wdot_finish(tab, ncases, pfirst, part.p(), inv_wa.p(), defl, s);
        }
        launch_b_p5_pre(tab, tiles, w, b, F.xt.p(), F.r.p(), s);
        launch_csr_op<V, X, OP_SPMV>(F.lanesR, F.R, F.r.p(), nullptr, nullptr, C.b.p(), X(0), s, nullptr, C.seg.p(),
                                     cstop.p());
        vcycle(1, s);
        launch_csr_op<V, X, OP_PROLONG>(F.lanesP, F.P, C.x.p(), nullptr, nullptr, F.xt.p(), X(0), s, nullptr,
                                        F.seg.p(), cstop.p());
        launch_b_p5_sweep2(tab, tiles, w, F.xt.p(), b, x, s);
    }
};

struct BatchAmg::Impl {
    int precision = 64;
    BatchDev<double, double> d64;
    BatchDev<float, float> d32;
};

BatchAmg::BatchAmg(const std::vector<AmgPrecond *> &amg, const std::vector<long> &offP, const BCase *tab_dev,
                   const std::vector<BCase> &tab_host, const BMapDev &pmap, const int *pfirst, hipStream_t s)
    : im_(new Impl) {
    if (amg.empty()) throw Error("batched V-cycle without cases", ERR_ARGUMENT);
    const int prec = amg[0]->precision();
    for (auto *a : amg) {
        if (a->precision() != prec) throw Error("batched V-cycle: cases of different AMG precisions", ERR_ARGUMENT);
        if (!a->has_stencil_level()) throw Error("batched V-cycle: a case without its stencil finest level", ERR_UNSUPPORTED);
    }
    if (prec != 64 && prec != 32) throw Error("batched V-cycle: AMG precision 64 or 32", ERR_UNSUPPORTED);
    im_->precision = prec;
    const int n = (int)amg.size();
    std::vector<long> tcount(n);
    for (int c = 0; c < n; c++) tcount[c] = p5_tiles(tab_host[c].p5.nx, tab_host[c].p5.ny);
    const bool deflate = tab_host[0].p5.defl != nullptr;
    std::vector<double> iw(n, 0.0);
    for (int c = 0; c < n; c++) {
        if ((tab_host[c].p5.defl != nullptr) != deflate)
            throw Error("batched V-cycle: deflation on some cases only", ERR_ARGUMENT);
        iw[c] = amg[c]->deflation_inv_wa();
    }
    std::vector<const double *> stops(n);
    for (int c = 0; c < n; c++) stops[c] = tab_host[c].p5.stop;
    auto setup = [&](auto &B, const auto &cd) {
        B.omega = amg[0]->prm_.omega;
        B.tab = tab_dev;
        B.pmap = pmap;
        B.pfirst = pfirst;
        B.build(cd, offP, stops, s);
        B.tilemap.build(tcount, 1, s);
        B.tiles = B.tilemap.dev();
        B.deflate = deflate;
        if (deflate) {
            B.inv_wa.upload(iw, s);
            B.part.alloc(std::max(pmap.nblk, 1));
            B.defl = const_cast<double *>(tab_host[0].p5.defl);  // the batch's array: case c at defl + c
        }
    };
    if (prec == 32) {
        std::vector<AmgDev<float, float> *> cd;
        for (auto *a : amg) cd.push_back(&a->im_->d32);
        setup(im_->d32, cd);
    } else {
        std::vector<AmgDev<double, double> *> cd;
        for (auto *a : amg) cd.push_back(&a->im_->d64);
        setup(im_->d64, cd);
    }
    CUIBM_HIP(hipStreamSynchronize(s));
}

BatchAmg::~BatchAmg() = default;

void BatchAmg::apply(const double *r, double *z, hipStream_t s) {
    if (im_->precision == 32) im_->d32.apply(r, z, s);
    else im_->d64.apply(r, z, s);
}

int BatchAmg::levels() const { return im_->precision == 32 ? (int)im_->d32.lv.size() : (int)im_->d64.lv.size(); }

}  // namespace cuibm
