// Linear solvers: preconditioned BiCGSTAB with device-resident recurrence scalars,
// Jacobi, and the smoothed-aggregation AMG V-cycle (FP64 or FP32 hierarchy).
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "dev.hpp"

namespace cuibm {

constexpr int RED_GRID = 2048;  // blocks of every reducing kernel

// slots of the device scalar block. SC_STOP is the device-side end of the iteration:
// 0 running, STOP_* below otherwise; every kernel of the iteration returns at once when
// it is set, so the host may queue several iterations before reading the block back.
enum { SC_BB = 0, SC_RHO, SC_ALPHA, SC_OMEGA, SC_BETA, SC_SS, SC_RR, SC_STOP, SC_ITS, SC_HALFK, SC_D0, SC_D1,
       SC_TOL, SC_COUNT = 16 };  // SC_TOL: the absolute tolerance of a batched solve
// STOP_HALF_PENDING: the half step converged in this iteration and the x update of its
// second half is still due (the iteration's FIN_BETA finish turns it into STOP_HALF), so
// the launches of an iteration do not depend on its index
enum { STOP_CONVERGED = 1, STOP_HALF = 2, STOP_BREAK_HALF = 3, STOP_BREAK_FULL = 4, STOP_MAXIT = 5,
       STOP_HALF_PENDING = 6 };
// finish-kernel operations
// FIN_RESTART (batched solves): a restart of the cases marked in a mask, from the true
// residual; the iteration count, the tolerance and the half-step record are kept
enum { FIN_INIT = 0, FIN_ALPHA, FIN_SS, FIN_OMEGA, FIN_BETA, FIN_DOT2, FIN_RESTART };

// ---- k_blas.hip
void launch_fill(double *x, double v, long n, hipStream_t s);
void launch_fill_f(float *x, float v, long n, hipStream_t s);
void launch_fill_i(int *x, int v, long n, hipStream_t s);
void launch_copy(double *y, const double *x, long n, hipStream_t s, const double *stop = nullptr);
void launch_to_float(float *y, const double *x, long n, hipStream_t s);
void launch_to_double(double *y, const float *x, long n, hipStream_t s);
template <class V, class X>
void launch_spmv(const DCsr<V> &A, const X *x, X *y, hipStream_t s);
void launch_spmv_dot(const DCsr<double> &A, const double *x, double *y, const double *a, bool want_yy, double *part,
                     hipStream_t s, const double *stop = nullptr);
void launch_bicg_init(const DCsr<double> &A, const double *x, const double *b, double *r, double *p, double *rs,
                      double *part, hipStream_t s);
// weights w_i = wx[i % nx] * wy[i / nx] of a preconditioner's input (cell areas), whose
// RED_GRID partial sums of w_i v_i the kernel producing the input v writes into part
struct InputWeights {
    const double *wx = nullptr, *wy = nullptr;
    int nx = 0;
    double *part = nullptr;
};
// s = r - alpha AMp with partials of (s.s); with w, also the weighted partials of s
void launch_bicg_s(int n, const double *r, const double *AMp, double *s_, const double *sc, double *part, hipStream_t s,
                   const InputWeights *w = nullptr);
// x += alpha Mp + omega Ms, r = s - omega AMs with partials; after a half-step exit of
// the same iteration (STOP_HALF_PENDING), x += alpha Mp only
void launch_bicg_xr(int n, double *x, const double *Mp, const double *Ms, double *r, const double *s_,
                    const double *AMs, const double *rs, const double *sc, double *part, hipStream_t s);
// p = r + beta (p - omega AMp); with w, also the weighted partials of p
void launch_bicg_p(int n, double *p, const double *r, const double *AMp, const double *sc, hipStream_t s,
                   const InputWeights *w = nullptr);
// sums the partials and updates the scalars; FIN_SS and FIN_BETA also apply the stopping
// tests of the monitor (tol) and the iteration limit (maxit) and set SC_STOP
void launch_bicg_finish(const double *part, int op, double *sc, hipStream_t s, double tol = 0, int maxit = 0);
void launch_dot2(long n, const double *x, const double *y, double *part, hipStream_t s);
void launch_jacobi_apply(int n, const double *dinv, const double *r, double *z, hipStream_t s,
                         const double *stop = nullptr);
void launch_diag_inv(const DCsr<double> &A, double *dinv, hipStream_t s);

// ---- linear operators (the A of BiCGSTAB)
class LinOp {
  public:
    virtual ~LinOp() = default;
    virtual int n() const = 0;
    // y = A x (y may be null) with partials of (y . a) and, if want_yy, (y . y)
    virtual void spmv_dot(const double *x, double *y, const double *a, bool want_yy, double *part,
                          hipStream_t s) const = 0;
    // r = b - A x, p = rs = r, partials of (r . r) and (b . b)
    virtual void init(const double *x, const double *b, double *r, double *p, double *rs, double *part,
                      hipStream_t s) const = 0;
    // algorithmic bytes of one product y = A x
    virtual double spmv_bytes() const = 0;
    // device flag (the solver's SC_STOP slot): the products return at once while it is set
    void set_stop(const double *stop) const { stop_ = stop; }

  protected:
    mutable const double *stop_ = nullptr;
};

class CsrOp : public LinOp {
  public:
    explicit CsrOp(const DCsr<double> &A) : A_(A) {}
    int n() const override { return A_.n; }
    void spmv_dot(const double *x, double *y, const double *a, bool want_yy, double *part,
                  hipStream_t s) const override {
        launch_spmv_dot(A_, x, y, a, want_yy, part, s, stop_);
    }
    void init(const double *x, const double *b, double *r, double *p, double *rs, double *part,
              hipStream_t s) const override {
        launch_bicg_init(A_, x, b, r, p, rs, part, s);
    }
    double spmv_bytes() const override { return A_.nnz * 12.0 + 4.0 * (A_.n + 1) + 16.0 * A_.n; }

  private:
    const DCsr<double> &A_;
};

// ---- preconditioners
class Preconditioner {
  public:
    virtual ~Preconditioner() = default;
    virtual void apply(const double *r, double *z, hipStream_t s) = 0;
    // the weights whose partial sums of each input the preconditioner forms (part null:
    // none), and apply() for an input whose producer has already written them into part
    virtual InputWeights input_weights() const { return {}; }
    virtual void apply_weighted(const double *r, double *z, hipStream_t s) { apply(r, z, s); }
    // a preconditioner whose rounding is far coarser than the solve's (the FP32 V-cycle)
    // is not the same linear map from one apply to the next; BiCGSTAB then restarts from
    // the true residual b - A x whenever its residual has fallen by this factor since the
    // last (re)start (reliable updates); 0: never
    virtual double restart_reduction() const { return 0.0; }
    virtual std::string name() const = 0;
    // bytes moved by one apply, counted as distinct arrays touched per kernel
    virtual double bytes_per_apply() const = 0;
    // device flag (the solver's SC_STOP slot): every kernel of an apply returns at once
    // while it is set
    virtual void set_stop(const double *stop) { stop_ = stop; }

  protected:
    const double *stop_ = nullptr;
};

class IdentityPrecond : public Preconditioner {
  public:
    explicit IdentityPrecond(int n) : n_(n) {}
    void apply(const double *r, double *z, hipStream_t s) override { launch_copy(z, r, n_, s, stop_); }
    std::string name() const override { return "NONE"; }
    double bytes_per_apply() const override { return 16.0 * n_; }

  private:
    int n_;
};

class JacobiPrecond : public Preconditioner {
  public:
    JacobiPrecond(const DCsr<double> &A, hipStream_t s);
    // D^-1 kept in caller-owned device memory of A.n doubles (allocates nothing)
    JacobiPrecond(const DCsr<double> &A, double *dinv, hipStream_t s);
    void apply(const double *r, double *z, hipStream_t s) override;
    // D^-1 of a new matrix of the same size, in place
    void refresh(const DCsr<double> &A, hipStream_t s);
    std::string name() const override { return "DIAGONAL"; }
    double bytes_per_apply() const override { return 24.0 * n_; }

  private:
    int n_;
    DArray<double> own_;
    double *dinv_ = nullptr;
};

// smoothed-aggregation AMG (setup on the host, V-cycle on the device)
struct AmgParams {
    // 64: FP64 hierarchy; 32: matrix values, level vectors and smoother arithmetic in FP32
    // (the V-cycle reads the FP64 Krylov vector and writes its result in FP64); 3264:
    // FP32 matrix values with FP64 level vectors and arithmetic
    int precision = 64;
    int coarse_max = 2500;  // dense coarsest solve (one 25 MB FP32 / 50 MB FP64 matrix at most)
    int max_levels = 25;
    double omega = 0.8;
    int npre = 2, npost = 2;
    // restrict the finest level with P^T D^-1 instead of P^T: the hierarchy of the
    // row-scaled D^-1 A, for matrices whose rows mix scales
    bool scale0 = false;
    // build G = [R | -R A] and restrict residuals in one product (an extra SpGEMM in the
    // set-up: left out for hierarchies rebuilt every step)
    bool fused_restriction = true;
    // keep the symbolic structure on the device for refresh() (taken only with a CSR
    // finest level, no fused restriction and at least two levels)
    bool refreshable = false;
    // with refreshable: the unknowns are the cells of a pad_nx-wide grid, and the
    // hierarchy is built on pad_pattern(A, pad_nx, pad_margin), so that matrices whose
    // body rows moved keep to its pattern (0: the pattern of A itself)
    int pad_nx = 0;
    int pad_margin = 8;
    // reliable-update factor of the BiCGSTAB it preconditions (restart_reduction());
    // < 0: 1e-3 for the FP32 V-cycle (the best of 1e-1 .. 1e-6 on the 1024^2 cylinder), 0
    // otherwise
    double restart = -1.0;
};

// HIP events around the launches of one kernel inside the live solve: mark() before
// and after each launch while enabled; total_ms() waits and sums the pairs
class KernelProbe {
  public:
    ~KernelProbe() {
        for (hipEvent_t e : ev_) (void)hipEventDestroy(e);
    }
    void enable(bool on) {
        on_ = on;
        used_ = 0;
    }
    bool enabled() const { return on_; }
    void mark(hipStream_t s) {
        if (!on_) return;
        if (used_ == ev_.size()) {
            hipEvent_t e;  // timestamps only: no system-scope release fence at the record
            CUIBM_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
            ev_.push_back(e);
        }
        CUIBM_HIP(hipEventRecord(ev_[used_++], s));
    }
    // a start/stop pair for one launch made with hipExtLaunchKernel, which stamps them
    // with the kernel's own begin and end (no dispatch latency inside the pair)
    void pair(hipEvent_t &start, hipEvent_t &stop) {
        while (ev_.size() < used_ + 2) {
            hipEvent_t e;
            CUIBM_HIP(hipEventCreate(&e));
            ev_.push_back(e);
        }
        start = ev_[used_++];
        stop = ev_[used_++];
    }
    int launches() const { return (int)(used_ / 2); }
    double total_ms() const {
        double t = 0;
        for (size_t i = 0; i + 1 < used_; i += 2) {
            CUIBM_HIP(hipEventSynchronize(ev_[i + 1]));
            float ms = 0;
            CUIBM_HIP(hipEventElapsedTime(&ms, ev_[i], ev_[i + 1]));
            t += ms;
        }
        return t;
    }

  private:
    std::vector<hipEvent_t> ev_;
    size_t used_ = 0;
    bool on_ = false;
};

struct AmgHierarchyInfo {
    int levels = 0;
    std::vector<int> n, nnz, nnzP;
};

// the pressure-Poisson operator as a 5-point stencil (see stencil.hpp)
struct P5Dev {
    int nx, ny;
    int centre;                        // cell whose diagonal is doubled besides cell 0, or -1
    const double *cE, *cW, *cN, *cS;   // coupling magnitudes; 0 where the neighbour is missing
    const double *dinv;                // 1 / diagonal, per cell
    const double *stop = nullptr;      // the kernels return at once while *stop != 0
    const float *dinv32 = nullptr;     // the same in FP32 (V-cycle with FP32 level vectors)
    const double *dx = nullptr, *dy = nullptr;  // cell widths (the cell areas weight the deflation)
    // FP64 deflation of the pinned constant mode for the FP32 V-cycle: defl[0] = c, the
    // input is r - c A 1 (A 1 is a0 at cell 0 and ac at the centre pin) and c is added to
    // the output. With defl_part, the pre-smoothing kernel forms c itself from those
    // RED_GRID partial sums of w_i r_i (c = their sum times defl_inv_wa) in the tiles whose
    // halo holds a pin, and the tile of cell 0 stores it in defl[0] for the post-smoothing
    // kernel; without, c is read from defl[0]
    double *defl = nullptr;
    double defl_a0 = 0, defl_ac = 0;
    const double *defl_part = nullptr;
    double defl_inv_wa = 0;
};

class AmgPrecond : public Preconditioner {
  public:
    // `fine`, when given, is the stencil form of A: the finest level then runs the
    // fused stencil kernels instead of CSR products with A (the hierarchy below is
    // still built from the CSR A)
    AmgPrecond(const HCsr &A, const AmgParams &p, hipStream_t s, const P5Dev *fine = nullptr);
    ~AmgPrecond() override;
    void apply(const double *r, double *z, hipStream_t s) override;
    // the FP32 V-cycle's deflation weights (cell areas) on a stencil finest level
    InputWeights input_weights() const override;
    void apply_weighted(const double *r, double *z, hipStream_t s) override;
    double restart_reduction() const override {
        return prm_.restart >= 0 ? prm_.restart : prm_.precision == 32 ? 1e-3 : 0.0;
    }
    void set_stop(const double *stop) override;
    std::string name() const override {
        return prm_.precision == 32 ? "SMOOTHED_AGGREGATION(FP32)"
                                    : prm_.precision == 3264 ? "SMOOTHED_AGGREGATION(FP32 values)" : "SMOOTHED_AGGREGATION";
    }
    double bytes_per_apply() const override;
    // algorithmic bytes of each launch of one V-cycle, in launch order (their sum is
    // bytes_per_apply())
    std::vector<double> launch_bytes() const;
    AmgHierarchyInfo info() const;
    // one weighted-Jacobi sweep xn = x + omega D^-1 (b - A x) on the finest level, as
    // run inside the V-cycle (for timing), and its algorithmic bytes
    void level0_sweep(const double *b, const double *x, double *xn, hipStream_t s);
    double level0_sweep_bytes() const;
    // the two fused finest-level kernels of the stencil form (timing): pre-smoothing
    // + residual (b -> x2, r) and prolongation + post-smoothing (x2, b -> x, with the
    // current coarse correction); bytes are their algorithmic bytes
    bool has_stencil_level() const { return fine_; }
    // events around every finest-level pre-smoothing launch of the live V-cycles
    KernelProbe &pre_probe();
    void level0_pre(const double *b, double *x2, double *r, hipStream_t s);
    void level0_post(double *x2, const double *b, double *x, hipStream_t s);
    double level0_pre_bytes() const;
    double level0_post_bytes() const;
    // the stencil finest level as the V-cycle sees it (FP32 D^-1, deflation constants),
    // and 1 / w^T a of the deflation (0 without deflation)
    const P5Dev &fine_dev() const { return fine_dev_; }
    // Numeric set-up on the device for a new matrix whose entries all lie in the pattern
    // the hierarchy was built on (A's own, or its padded superset): A is scattered into
    // that pattern (zeros elsewhere), the aggregates and every level's pattern are kept,
    // and D, the prolongator smoothing, P, R (P^T, or P^T D^-1 on a row-scaled finest
    // level), A P, R A P and the coarsest inverse (Gauss-Jordan with partial pivoting in
    // one workgroup, up to 64 unknowns; on the host above that) are recomputed on the
    // device, each value by the host set-up's sequence of operations. With unchanged
    // aggregates the hierarchy is the one the host would build from the same aggregates;
    // a fresh host set-up may aggregate differently where entries became zero or moved.
    // Returns false when an entry of A lies outside the pattern (the caller then builds
    // a new hierarchy). CSR finest level only.
    bool refresh(const DCsr<double> &A, hipStream_t s);
    double deflation_inv_wa() const;
    int precision() const { return prm_.precision; }

  private:
    friend class BatchAmg;
    struct Impl;
    std::unique_ptr<Impl> im_;
    AmgParams prm_;
    bool fine_ = false;
    P5Dev fine_dev_{};
};

// host smoothed-aggregation set-up (amg_setup.cpp)
struct HostAmgLevel {
    HCsr A, P, R;
    std::vector<int> agg;  // the aggregate of each row (every level but the coarsest)
    // G = [R | -R A] (A.n + A.n columns): the restricted residual R (b - A x) as one
    // product with the pair (b, x), on every level but the coarsest
    HCsr G;
};
std::vector<HostAmgLevel> amg_build(const HCsr &A, int coarse_max, int max_levels, bool scale0 = false,
                                    bool build_g = true);
std::vector<double> dense_inverse(const HCsr &A);
// A (the unknowns being the cells of an nx-wide grid) with explicit zero entries added:
// every row within `margin` cells of a row that is not a 5-point row gets all the columns
// within `reach` cells of it, reach being the largest column offset of those rows, so a
// body's rows that move by up to about `margin` cells keep to this pattern. A is returned
// unchanged when every row is a 5-point row or a column lies more than 4 cells away.
HCsr pad_pattern(const HCsr &A, int nx, int margin);
// the sparsity pattern (row pointers, sorted columns) of A B, as amg_build's products form it
void multiply_pattern(const HCsr &A, const HCsr &B, std::vector<int> &rp, std::vector<int> &ci);

// ---- BiCGSTAB (krylov.hip)
struct SolveStat {
    int iterations = 0;
    int restarts = 0;  // breakdown restarts (0 in a recurrence that stays finite)
    double residual = 0, tolerance = 0;
    bool converged = false;
};

class Bicgstab {
  public:
    explicit Bicgstab(int n);
    // workspace in caller-owned memory: `ws` holds workspace_doubles(n) device doubles
    // (256-byte aligned), `host_sc` SC_COUNT host doubles; allocates nothing
    Bicgstab(int n, double *ws, double *host_sc);
    ~Bicgstab();
    static size_t workspace_doubles(int n) { return 8 * stride(n) + 2 * RED_GRID + SC_COUNT; }
    // cusp::krylov::bicgstab + cusp::monitor(b, maxit, rtol) semantics; x is the warm start
    SolveStat solve(const LinOp &A, double *x, const double *b, Preconditioner &M, double rtol, int maxit,
                    hipStream_t s);
    int n() const { return n_; }
    // exact launches: read the scalars back at every stopping test (twice per iteration)
    // and launch nothing past the end; otherwise iterations are queued in batches sized
    // from the observed convergence rate and the device stops them (kernel timing uses
    // the exact mode so that no launch returns early)
    void set_exact(bool on) { exact_ = on; }

  private:
    static constexpr int kMaxRestarts = 64;
    static constexpr int kVerifyIts = 100, kMaxChecks = 3;
    static constexpr double kDriftFactor = 1e3;
    // vector stride in the external workspace: a multiple of 32 doubles (256 bytes)
    static size_t stride(int n) { return ((size_t)n + 31) / 32 * 32; }
    int n_;
    bool exact_ = false;
    double rate_ = 0;  // mean residual reduction per iteration of the last solve
    DArray<double> own_;
    double *r_ = nullptr, *rs_ = nullptr, *p_ = nullptr, *Mp_ = nullptr, *AMp_ = nullptr, *s_ = nullptr,
           *Ms_ = nullptr, *AMs_ = nullptr, *part_ = nullptr, *sc_ = nullptr;
    double *host_sc_ = nullptr;
    bool pinned_ = false;  // host_sc_ is this object's pinned allocation
};

}  // namespace cuibm
