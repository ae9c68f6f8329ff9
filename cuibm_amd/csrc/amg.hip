// Smoothed-aggregation AMG V-cycle on the device (FP64 or FP32 hierarchy).
//
// One application is one V-cycle from a zero initial guess: on every level but the
// coarsest, npre weighted-Jacobi sweeps (omega = 0.8; the first is x = omega D^-1 b),
// residual, restriction with R = P^T, recursion, prolongation-correction, npost
// sweeps; the coarsest level (<= 256 unknowns) applies a dense inverse with one
// block. In FP32 mode the hierarchy values and vectors are float and the FP64
// outer Krylov vectors are converted on entry and exit (defect correction).
#include <algorithm>

#include "krylov.hpp"

namespace cuibm {

namespace {

int grid_for(long n) {
    const long b = (n + BS - 1) / BS;
    return (int)(b < 65536 ? b : 65536);
}

template <class T>
__global__ void k_jacobi_first(int n, const T *__restrict__ dinv, const T *__restrict__ b, T *__restrict__ x, T omega) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = omega * dinv[i] * b[i];
}

// xn = x + omega D^-1 (b - A x)
template <class T>
__global__ void __launch_bounds__(BS) k_jacobi(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                               const T *__restrict__ v, const T *__restrict__ dinv,
                                               const T *__restrict__ b, const T *__restrict__ x, T *__restrict__ xn,
                                               T omega) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        T s = 0;
        for (int k = rp[i]; k < rp[i + 1]; k++) s += v[k] * x[ci[k]];
        xn[i] = x[i] + omega * dinv[i] * (b[i] - s);
    }
}

// r = b - A x
template <class T>
__global__ void __launch_bounds__(BS) k_resid(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                              const T *__restrict__ v, const T *__restrict__ b,
                                              const T *__restrict__ x, T *__restrict__ r) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        T s = 0;
        for (int k = rp[i]; k < rp[i + 1]; k++) s += v[k] * x[ci[k]];
        r[i] = b[i] - s;
    }
}

// x += P xc
template <class T>
__global__ void __launch_bounds__(BS) k_prolong_add(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                                    const T *__restrict__ v, const T *__restrict__ xc,
                                                    T *__restrict__ x) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        T s = 0;
        for (int k = rp[i]; k < rp[i + 1]; k++) s += v[k] * xc[ci[k]];
        x[i] += s;
    }
}

// x = Ainv b on the coarsest level (one block; b staged through LDS in tiles)
template <class T>
__global__ void __launch_bounds__(256) k_dense(int n, const T *__restrict__ Ainv, const T *__restrict__ b,
                                               T *__restrict__ x) {
    __shared__ T sb[256];
    for (int i0 = 0; i0 < n; i0 += blockDim.x) {
        const int i = i0 + threadIdx.x;
        T s = 0;
        for (int j0 = 0; j0 < n; j0 += 256) {
            __syncthreads();
            if (j0 + (int)threadIdx.x < n) sb[threadIdx.x] = b[j0 + threadIdx.x];
            __syncthreads();
            const int je = min(256, n - j0);
            if (i < n)
                for (int j = 0; j < je; j++) s += Ainv[(size_t)i * n + j0 + j] * sb[j];
        }
        if (i < n) x[i] = s;
    }
}

template <class T>
__global__ void k_fill_t(int n, T *x, T v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = v;
}

}  // namespace

template <class T>
struct AmgLevelDev {
    int n = 0;
    DCsr<T> A, P, R;
    DArray<T> dinv, b, x, xt, r;
};

template <class T>
struct AmgDev {
    std::vector<AmgLevelDev<T>> lv;
    DArray<T> coarse_inv;
    int nc = 0;
    AmgParams prm;
    DArray<T> b0, x0;  // FP32 entry/exit buffers

    void build(const std::vector<HostAmgLevel> &h, hipStream_t s) {
        lv.resize(h.size());
        for (size_t l = 0; l < h.size(); l++) {
            AmgLevelDev<T> &L = lv[l];
            L.n = h[l].A.n;
            upload_csr(h[l].A, L.A, s);
            if (l + 1 < h.size()) {
                upload_csr(h[l].P, L.P, s);
                upload_csr(h[l].R, L.R, s);
            }
            std::vector<T> d(L.n, T(0));
            for (int i = 0; i < L.n; i++) {
                double a = 0;
                for (int k = h[l].A.rp[i]; k < h[l].A.rp[i + 1]; k++)
                    if (h[l].A.ci[k] == i) a += h[l].A.v[k];
                d[i] = (T)(1.0 / a);
            }
            L.dinv.upload(d, s);
            L.b.alloc(L.n);
            L.x.alloc(L.n);
            L.xt.alloc(L.n);
            L.r.alloc(L.n);
        }
        nc = h.back().A.n;
        if (nc > 4096) throw Error("AMG coarsening stalled at " + std::to_string(nc) + " unknowns", ERR_GENERIC);
        const std::vector<double> inv = dense_inverse(h.back().A);
        std::vector<T> invt(inv.begin(), inv.end());
        coarse_inv.upload(invt, s);
        CUIBM_HIP(hipStreamSynchronize(s));
    }

    // V-cycle on level l with right-hand side b, result into x (both level-sized)
    void vcycle(size_t l, const T *b, T *x, hipStream_t s) {
        AmgLevelDev<T> &L = lv[l];
        const int n = L.n;
        const T w = (T)prm.omega;
        if (l + 1 == lv.size()) {
            hipLaunchKernelGGL(k_dense<T>, dim3(1), dim3(256), 0, s, n, coarse_inv.p(), b, x);
            CUIBM_LAUNCH_CHECK();
            return;
        }
        const int g = grid_for(n);
        // ping-pong so that the last sweep lands in x
        const int swaps = (prm.npre - 1) + prm.npost;
        T *cur = (swaps % 2 == 0) ? x : L.xt.p();
        T *oth = (cur == x) ? L.xt.p() : x;
        hipLaunchKernelGGL(k_jacobi_first<T>, dim3(g), dim3(BS), 0, s, n, L.dinv.p(), b, cur, w);
        CUIBM_LAUNCH_CHECK();
        for (int k = 1; k < prm.npre; k++) {
            hipLaunchKernelGGL(k_jacobi<T>, dim3(g), dim3(BS), 0, s, n, L.A.rp.p(), L.A.ci.p(), L.A.v.p(), L.dinv.p(),
                               b, cur, oth, w);
            CUIBM_LAUNCH_CHECK();
            std::swap(cur, oth);
        }
        hipLaunchKernelGGL(k_resid<T>, dim3(g), dim3(BS), 0, s, n, L.A.rp.p(), L.A.ci.p(), L.A.v.p(), b, cur, L.r.p());
        CUIBM_LAUNCH_CHECK();
        AmgLevelDev<T> &C = lv[l + 1];
        launch_spmv<T, T>(L.R, L.r.p(), C.b.p(), s);
        vcycle(l + 1, C.b.p(), C.x.p(), s);
        hipLaunchKernelGGL(k_prolong_add<T>, dim3(g), dim3(BS), 0, s, n, L.P.rp.p(), L.P.ci.p(), L.P.v.p(), C.x.p(),
                           cur);
        CUIBM_LAUNCH_CHECK();
        for (int k = 0; k < prm.npost; k++) {
            hipLaunchKernelGGL(k_jacobi<T>, dim3(g), dim3(BS), 0, s, n, L.A.rp.p(), L.A.ci.p(), L.A.v.p(), L.dinv.p(),
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
        im_->d32.b0.alloc(A.n);
        im_->d32.x0.alloc(A.n);
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
    if (prm_.precision == 32) {
        AmgDev<float> &d = im_->d32;
        const long n = d.lv[0].n;
        launch_to_float(d.b0.p(), r, n, s);
        d.vcycle(0, d.b0.p(), d.x0.p(), s);
        launch_to_double(z, d.x0.p(), n, s);
    } else {
        im_->d64.vcycle(0, r, z, s);
    }
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
    // distinct arrays per kernel: CSR spmv = nnz*(s+4) + 4(n+1) + n*s (x) + n*s (y)
    const double s = prm_.precision == 32 ? 4.0 : 8.0;
    double bytes = 0;
    const auto &H = im_->host;
    auto spmv = [&](double n, double ncol, double nnz) { return nnz * (s + 4) + 4 * (n + 1) + ncol * s + n * s; };
    for (size_t l = 0; l + 1 < H.size(); l++) {
        const double n = H[l].A.n, nnz = (double)H[l].A.ci.size();
        const double nc = H[l + 1].A.n, nnzP = (double)H[l].P.ci.size();
        bytes += 3 * n * s;                                            // first sweep: dinv, b, x
        const int sweeps = prm_.npre - 1 + prm_.npost;
        bytes += sweeps * (spmv(n, n, nnz) + 2 * n * s);               // + dinv, b
        bytes += spmv(n, n, nnz) + n * s;                              // residual (+ b)
        bytes += spmv(nc, n, nnzP);                                    // restriction
        bytes += spmv(n, nc, nnzP) + n * s;                            // prolong-add (read-modify-write x)
    }
    const double nc = H.back().A.n;
    bytes += nc * nc * s + 2 * nc * s;
    if (prm_.precision == 32) bytes += 2 * H[0].A.n * (8 + 4);
    return bytes;
}

}  // namespace cuibm
