// BiCGSTAB for callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc): every device buffer of the solve
// lives in one file-scope device array, `scratch`, of fixed size.
//
// Layout of `scratch` for a system of n unknowns (stride = n rounded up to 32 doubles):
//   [0, stride)            x, staged from the caller's host vector
//   [stride, 2 stride)     b, staged likewise
//   [2 stride, 3 stride)   D^-1 of the Jacobi preconditioner
//   [3 stride, ...)        the BiCGSTAB workspace: 8 vectors, 2 x 2048 reduction
//                          partials, the 16-double scalar block
// Each GPU holds its own copy of the array at its own address, looked up once per device
// with hipGetSymbolAddress; host code passes only that address to kernels and copies.
// Calls on one GPU share the array, so calls that overlap on one GPU race.
#include <mutex>

#include "krylov.hpp"
#include "noalloc.hpp"

namespace cuibm {

static_assert(NOALLOC_MAX_N % 32 == 0, "stride of the largest system must equal NOALLOC_MAX_N");
constexpr size_t NOALLOC_DOUBLES = 3 * (size_t)NOALLOC_MAX_N + 8 * (size_t)NOALLOC_MAX_N + 2 * RED_GRID + SC_COUNT;

static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};
}  // namespace

// device address of this GPU's `scratch` (looked up on the first call on each device)
static double *scratch_ptr() {
    int dev = 0;
    CUIBM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error("device ordinal out of range", ERR_ARGUMENT);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        double *p = nullptr;
        CUIBM_HIP(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error("hipGetSymbolAddress returned a null address for the scratch array", ERR_HIP);
        g_ptr[dev] = p;
    }
    return g_ptr[dev];
}

SolveStat bicgstab_noalloc(const DCsr<double> &A, double *x, const double *b, int precond, double rtol, int maxit,
                           hipStream_t s) {
    const int n = A.n;
    if (n <= 0 || n > NOALLOC_MAX_N)
        throw Error("cuibm_bicgstab_noalloc: n = " + std::to_string(n) + " is outside 1.." +
                        std::to_string(NOALLOC_MAX_N) + " (the size of the library's scratch array)",
                    ERR_ARGUMENT);
    if (precond != 0 && precond != 1)
        throw Error("cuibm_bicgstab_noalloc: only NONE (0) and DIAGONAL (1) preconditioners allocate nothing",
                    ERR_UNSUPPORTED);
    double *ws = scratch_ptr();
    const size_t st = ((size_t)n + 31) / 32 * 32;
    double *dx = ws, *db = ws + st, *dinv = ws + 2 * st, *kws = ws + 3 * st;
    // the reduction partials and the scalar block start from zero on every call (the
    // array keeps the previous call's contents); a fill kernel, not a memset node
    launch_fill(kws + 8 * st, 0.0, 2 * RED_GRID + SC_COUNT, s);
    CUIBM_HIP(hipMemcpyAsync(dx, x, n * sizeof(double), hipMemcpyHostToDevice, s));
    CUIBM_HIP(hipMemcpyAsync(db, b, n * sizeof(double), hipMemcpyHostToDevice, s));
    double host_sc[SC_COUNT] = {};
    Bicgstab ks(n, kws, host_sc);
    const CsrOp op(A);
    SolveStat st_out;
    if (precond == 1) {
        JacobiPrecond M(A, dinv, s);
        st_out = ks.solve(op, dx, db, M, rtol, maxit, s);
    } else {
        IdentityPrecond M(n);
        st_out = ks.solve(op, dx, db, M, rtol, maxit, s);
    }
    CUIBM_HIP(hipMemcpyAsync(x, dx, n * sizeof(double), hipMemcpyDeviceToHost, s));
    CUIBM_HIP(hipStreamSynchronize(s));
    return st_out;
}

}  // namespace cuibm
