// Smoothed-aggregation AMG V-cycle on the device, with the hierarchy's matrix values
// in FP64 or FP32.
//
// One application is one V-cycle from a zero initial guess: on every level but the
// coarsest, npre weighted-Jacobi sweeps (omega = 0.8; the first is x = omega D^-1 b),
// residual, restriction with R = P^T, recursion, prolongation-correction, npost
// sweeps; the coarsest level (<= 256 unknowns) applies a dense inverse with one
// block.
//
// Mixed precision stores only the matrix values (A_l, P_l, R_l, D_l^-1 and the
// coarse inverse) as float; the level vectors and every product, sum and update stay
// in FP64. Rounding the matrices perturbs a linear operator, so the preconditioner
// stays linear and the outer BiCGSTAB converges as with the FP64 hierarchy. Rounding
// the level vectors instead makes the V-cycle slightly non-linear; on the pinned
// Neumann Poisson operator this lets BiCGSTAB iteration counts grow by orders of
// magnitude over a run and end in breakdown.
#include <algorithm>

#include "krylov.hpp"

namespace cuibm {

namespace {

int grid_for(long n) {
    const long b = (n + BS - 1) / BS;
    return (int)(b < 65536 ? b : 65536);
}

template <class V>
__global__ void k_jacobi_first(int n, const V *__restrict__ dinv, const double *__restrict__ b,
                               double *__restrict__ x, double omega) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        x[i] = omega * (double)dinv[i] * b[i];
}

// xn = x + omega D^-1 (b - A x)
template <class V>
__global__ void __launch_bounds__(BS) k_jacobi(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                               const V *__restrict__ v, const V *__restrict__ dinv,
                                               const double *__restrict__ b, const double *__restrict__ x,
                                               double *__restrict__ xn, double omega) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double s = 0;
        for (int k = rp[i]; k < rp[i + 1]; k++) s += (double)v[k] * x[ci[k]];
        xn[i] = x[i] + omega * (double)dinv[i] * (b[i] - s);
    }
}

// r = b - A x
template <class V>
__global__ void __launch_bounds__(BS) k_resid(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                              const V *__restrict__ v, const double *__restrict__ b,
                                              const double *__restrict__ x, double *__restrict__ r) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double s = 0;
        for (int k = rp[i]; k < rp[i + 1]; k++) s += (double)v[k] * x[ci[k]];
        r[i] = b[i] - s;
    }
}

// x += P xc
template <class V>
__global__ void __launch_bounds__(BS) k_prolong_add(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                                    const V *__restrict__ v, const double *__restrict__ xc,
                                                    double *__restrict__ x) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double s = 0;
        for (int k = rp[i]; k < rp[i + 1]; k++) s += (double)v[k] * xc[ci[k]];
        x[i] += s;
    }
}

// x = Ainv b on the coarsest level (one block; b staged through LDS in tiles)
template <class V>
__global__ void __launch_bounds__(256) k_dense(int n, const V *__restrict__ Ainv, const double *__restrict__ b,
                                               double *__restrict__ x) {
    __shared__ double sb[256];
    for (int i0 = 0; i0 < n; i0 += blockDim.x) {
        const int i = i0 + threadIdx.x;
        double s = 0;
        for (int j0 = 0; j0 < n; j0 += 256) {
            __syncthreads();
            if (j0 + (int)threadIdx.x < n) sb[threadIdx.x] = b[j0 + threadIdx.x];
            __syncthreads();
            const int je = min(256, n - j0);
            if (i < n)
                for (int j = 0; j < je; j++) s += (double)Ainv[(size_t)i * n + j0 + j] * sb[j];
        }
        if (i < n) x[i] = s;
    }
}

}  // namespace

template <class V>
struct AmgLevelDev {
    int n = 0;
    DCsr<V> A, P, R;
    DArray<V> dinv;
    DArray<double> b, x, xt, r;
};

template <class V>
struct AmgDev {
    std::vector<AmgLevelDev<V>> lv;
    DArray<V> coarse_inv;
    int nc = 0;
    AmgParams prm;

    void build(const std::vector<HostAmgLevel> &h, hipStream_t s) {
        lv.resize(h.size());
        for (size_t l = 0; l < h.size(); l++) {
            AmgLevelDev<V> &L = lv[l];
            L.n = h[l].A.n;
            upload_csr(h[l].A, L.A, s);
            if (l + 1 < h.size()) {
                upload_csr(h[l].P, L.P, s);
                upload_csr(h[l].R, L.R, s);
            }
            std::vector<V> d(L.n, V(0));
            for (int i = 0; i < L.n; i++) {
                double a = 0;
                for (int k = h[l].A.rp[i]; k < h[l].A.rp[i + 1]; k++)
                    if (h[l].A.ci[k] == i) a += h[l].A.v[k];
                d[i] = (V)(1.0 / a);
            }
            L.dinv.upload(d, s);
            if (l > 0) {  // level 0 works on the caller's vectors
                L.b.alloc(L.n);
                L.x.alloc(L.n);
            }
            L.xt.alloc(L.n);
            L.r.alloc(L.n);
        }
        nc = h.back().A.n;
        if (nc > 4096) throw Error("AMG coarsening stalled at " + std::to_string(nc) + " unknowns", ERR_GENERIC);
        const std::vector<double> inv = dense_inverse(h.back().A);
        std::vector<V> invt(inv.begin(), inv.end());
        coarse_inv.upload(invt, s);
        CUIBM_HIP(hipStreamSynchronize(s));
    }

    // V-cycle on level l with right-hand side b, result into x (both level-sized)
    void vcycle(size_t l, const double *b, double *x, hipStream_t s) {
        AmgLevelDev<V> &L = lv[l];
        const int n = L.n;
        const double w = prm.omega;
        if (l + 1 == lv.size()) {
            hipLaunchKernelGGL(k_dense<V>, dim3(1), dim3(256), 0, s, n, coarse_inv.p(), b, x);
            CUIBM_LAUNCH_CHECK();
            return;
        }
        const int g = grid_for(n);
        // ping-pong so that the last sweep lands in x
        const int swaps = (prm.npre - 1) + prm.npost;
        double *cur = (swaps % 2 == 0) ? x : L.xt.p();
        double *oth = (cur == x) ? L.xt.p() : x;
        hipLaunchKernelGGL(k_jacobi_first<V>, dim3(g), dim3(BS), 0, s, n, L.dinv.p(), b, cur, w);
        CUIBM_LAUNCH_CHECK();
        for (int k = 1; k < prm.npre; k++) {
            hipLaunchKernelGGL(k_jacobi<V>, dim3(g), dim3(BS), 0, s, n, L.A.rp.p(), L.A.ci.p(), L.A.v.p(), L.dinv.p(),
                               b, cur, oth, w);
            CUIBM_LAUNCH_CHECK();
            std::swap(cur, oth);
        }
        hipLaunchKernelGGL(k_resid<V>, dim3(g), dim3(BS), 0, s, n, L.A.rp.p(), L.A.ci.p(), L.A.v.p(), b, cur, L.r.p());
        CUIBM_LAUNCH_CHECK();
        AmgLevelDev<V> &C = lv[l + 1];
        launch_spmv<V, double>(L.R, L.r.p(), C.b.p(), s);
        vcycle(l + 1, C.b.p(), C.x.p(), s);
        hipLaunchKernelGGL(k_prolong_add<V>, dim3(g), dim3(BS), 0, s, n, L.P.rp.p(), L.P.ci.p(), L.P.v.p(), C.x.p(),
                           cur);
        CUIBM_LAUNCH_CHECK();
        for (int k = 0; k < prm.npost; k++) {
            hipLaunchKernelGGL(k_jacobi<V>, dim3(g), dim3(BS), 0, s, n, L.A.rp.p(), L.A.ci.p(), L.A.v.p(), L.dinv.p(),
                               b, cur, oth, w);
            CUIBM_LAUNCH_CHECK();
            std::swap(cur, oth);
        }
    }
};

struct AmgPrecond::Impl {
    AmgDev<double> d64;
    AmgDev<float> d32;
    std::vector<HostAmgLevel> host;
};

AmgPrecond::AmgPrecond(const HCsr &A, const AmgParams &p, hipStream_t s) : im_(new Impl), prm_(p) {
    im_->host = amg_build(A, p.coarse_max, p.max_levels);
    if (p.precision == 32) {
        im_->d32.prm = p;
        im_->d32.build(im_->host, s);
    } else {
        im_->d64.prm = p;
        im_->d64.build(im_->host, s);
    }
    // keep only the sizes of the host hierarchy
    for (auto &L : im_->host) {
        L.A.v.clear(); L.A.v.shrink_to_fit();
        L.P.v.clear(); L.P.v.shrink_to_fit();
        L.R.v.clear(); L.R.v.shrink_to_fit();
    }
}

AmgPrecond::~AmgPrecond() = default;

void AmgPrecond::apply(const double *r, double *z, hipStream_t s) {
    if (prm_.precision == 32) im_->d32.vcycle(0, r, z, s);
    else im_->d64.vcycle(0, r, z, s);
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
