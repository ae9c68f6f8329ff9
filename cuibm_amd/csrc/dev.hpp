// HIP helpers: error checking, owning device arrays, launch geometry.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

namespace cuibm {

// Library error with a C-ABI return code (negative)
struct Error : std::runtime_error {
    int code;
    explicit Error(const std::string &m, int c = -1) : std::runtime_error(m), code(c) {}
};

enum ErrorCode {
    ERR_GENERIC = -1,
    ERR_INPUT = -2,
    ERR_HIP = -3,
    ERR_NOT_CONVERGED = -4,
    ERR_UNSUPPORTED = -5,
    ERR_ARGUMENT = -6,
};

#define CUIBM_HIP(expr)                                                                                  \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            throw ::cuibm::Error(std::string(#expr) + ": " + hipGetErrorString(e_) + " (" __FILE__ ":" + \
                                     std::to_string(__LINE__) + ")",                                     \
                                 ::cuibm::ERR_HIP);                                                      \
    } while (0)

#define CUIBM_LAUNCH_CHECK() CUIBM_HIP(hipGetLastError())

constexpr int BS = 256;
inline int nblk(long n, int bs = BS) { return (int)((n + bs - 1) / bs); }

// The MI355X hands workgroups to its 8 XCDs (each with its own L2) round-robin. With a
// grid that is a multiple of 8, block b -> (b % 8) * (grid / 8) + b / 8 is a permutation
// that gives each XCD a contiguous range of the launch's work, so rows that gather
// from neighbouring indices find them in their own XCD's L2.
__device__ __forceinline__ int xcd_block(int grid) { return (int)(blockIdx.x % 8) * (grid / 8) + (int)(blockIdx.x / 8); }
// grid of at least `blocks` workgroups, a multiple of 8, at most 65536
inline unsigned xcd_blocks(long blocks) {
    const long b = std::min<long>(std::max<long>(blocks, 1), 65536);
    return (unsigned)((b + 7) / 8 * 8);
}

// s = sum over k = b, b + STRIDE, ... < e of v[k] g(ci[k]), added in k order in type S
// (STRIDE > 1: one lane's share of a row that several lanes split). The index and value
// loads of a group of entries (5, a 5-point row, for one lane per row; 4 for a lane of a
// split row) are issued together, then the gathers g(c), before the first add: a loop
// that waits for each entry's loads in turn keeps a row about two memory latencies per
// entry long, and with every block of a launch resident the cache lines of a row's first
// entries can leave L1 and L2 before its last entries read them. The loads are
// unconditional (a short group re-reads its last entry, and g must read only what its
// column allows) so that no branch splits them; only the adds are masked.
template <class S, int STRIDE, class V, class Gather>
__device__ __forceinline__ S csr_lane_sum(const int *__restrict__ ci, const V *__restrict__ v, int b, int e,
                                          Gather g) {
    constexpr int G = STRIDE == 1 ? 5 : 4;
    S s = 0;
    for (int k0 = b; k0 < e; k0 += G * STRIDE) {
        const int last = min(G, (e - k0 + STRIDE - 1) / STRIDE) - 1;
        int c[G];
        S a[G], xv[G];
#pragma unroll
        for (int j = 0; j < G; j++) {
            const int k = k0 + min(j, last) * STRIDE;
            c[j] = ci[k];
            a[j] = (S)v[k];
        }
#pragma unroll
        for (int j = 0; j < G; j++) xv[j] = g(c[j]);
#pragma unroll
        for (int j = 0; j < G; j++)
            if (j <= last) s += a[j] * xv[j];
    }
    return s;
}

// s = sum over k in [b, e) of v[k] x[ci[k]], added in k order in type S
template <class S, class V, class X>
__device__ __forceinline__ S csr_row_sum(const int *__restrict__ ci, const V *__restrict__ v,
                                         const X *__restrict__ x, int b, int e) {
    return csr_lane_sum<S, 1>(ci, v, b, e, [&](int c) { return (S)x[c]; });
}

template <class T>
class DArray {
  public:
    DArray() = default;
    explicit DArray(size_t n) { alloc(n); }
    DArray(const DArray &) = delete;
    DArray &operator=(const DArray &) = delete;
    DArray(DArray &&o) noexcept : p_(o.p_), n_(o.n_), cap_(o.cap_) { o.p_ = nullptr; o.n_ = o.cap_ = 0; }
    DArray &operator=(DArray &&o) noexcept {
        if (this != &o) {
            release();
            p_ = o.p_;
            n_ = o.n_;
            cap_ = o.cap_;
            o.p_ = nullptr;
            o.n_ = o.cap_ = 0;
        }
        return *this;
    }
    ~DArray() { release(); }
    // n elements; an allocation that already holds n is kept (arrays resized every step,
    // such as a moving body's matrices and node lists, allocate only when they grow)
    void alloc(size_t n) {
        if (p_ && n <= cap_) {
            n_ = n;
            return;
        }
        release();
        if (n) CUIBM_HIP(hipMalloc((void **)&p_, n * sizeof(T)));
        n_ = cap_ = n;
    }
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = cap_ = 0;
    }
    void upload(const T *h, size_t n, hipStream_t s) {
        alloc(n);
        if (n) CUIBM_HIP(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T> &h, hipStream_t s) { upload(h.data(), h.size(), s); }
    void download(T *h, size_t n, hipStream_t s) const {
        if (n > n_) throw Error("download of " + std::to_string(n) + " > " + std::to_string(n_), ERR_ARGUMENT);
        if (n) CUIBM_HIP(hipMemcpyAsync(h, p_, n * sizeof(T), hipMemcpyDeviceToHost, s));
        CUIBM_HIP(hipStreamSynchronize(s));
    }
    std::vector<T> to_host(hipStream_t s) const {
        std::vector<T> h(n_);
        download(h.data(), n_, s);
        return h;
    }
    T *p() const { return p_; }
    size_t n() const { return n_; }
    bool empty() const { return n_ == 0; }

  private:
    T *p_ = nullptr;
    size_t n_ = 0, cap_ = 0;
};

// device-side CSR matrix; values in V (double or float)
template <class V>
struct DCsr {
    int n = 0, ncols = 0, nnz = 0;
    DArray<int> rp, ci;
    DArray<V> v;
};

// host-side CSR
struct HCsr {
    int n = 0, ncols = 0;
    std::vector<int> rp, ci;
    std::vector<double> v;
    int nnz() const { return (int)ci.size(); }
};

template <class V>
void upload_csr(const HCsr &h, DCsr<V> &d, hipStream_t s) {
    d.n = h.n;
    d.ncols = h.ncols;
    d.nnz = h.nnz();
    d.rp.upload(h.rp, s);
    d.ci.upload(h.ci, s);
    std::vector<V> vv(h.v.begin(), h.v.end());
    d.v.upload(vv, s);
    CUIBM_HIP(hipStreamSynchronize(s));
}

}  // namespace cuibm
