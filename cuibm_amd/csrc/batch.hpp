// Lockstep batches of independent Luo-external cases (no counterpart in the reference, which
// runs one case per process): every phase of the time step is one launch over all cases.
//
// A 2-D grid of 96^2..512^2 cells fills a few percent of the MI355X, and one solver per
// case on its own stream still issues every phase as a separate small launch per case.
// Here the cases' fields live in concatenated arrays (each case's segment padded to a
// multiple of 64 values so that vector loads stay aligned), and each batched kernel maps
// its blocks to (case, block within the case) through a block map; the case's geometry,
// operators and offsets come from a per-case table in device memory. The immersed-
// boundary stencils and the AMG levels below the finest are concatenated with their
// indices shifted, so the single-case kernels run on them unchanged; the finest AMG level
// stays in stencil form per case. Each case keeps its own BiCGSTAB scalars and stop code:
// its vector updates stop at its own convergence, and the host loop ends when every case
// has stopped, so each case's iterate is the one its own solve would stop at.
#pragma once
#include <memory>
#include <vector>

#include "ibm.hpp"
#include "kernels.hpp"
#include "krylov.hpp"
#include "stencil.hpp"

namespace cuibm {

constexpr int BATCH_PAD = 64;  // values: every case segment starts at a multiple of this
inline long batch_pad(long n) { return (n + BATCH_PAD - 1) / BATCH_PAD * BATCH_PAD; }

// device view of a block map: block b works on case cs[b], its loc[b]-th block
struct BMapDev {
    const int *cs, *loc;
    int nblk;
};

class BMap {
  public:
    // per_block items of case c per block, count[c] items in case c
    void build(const std::vector<long> &count, int per_block, hipStream_t s);
    BMapDev dev() const { return BMapDev{cs_.p(), loc_.p(), nblk_}; }
    int nblk() const { return nblk_; }
    // first block of each case and the total (ncases + 1 values, device)
    const int *first() const { return first_.p(); }

  private:
    DArray<int> cs_, loc_, first_;
    int nblk_ = 0;
};

// per-case view for the batched kernels
struct BCase {
    GridDev g;
    BcDev bc;  // bc.xp points into the batch's outflow values
    double nu, dt;
    long offUV, offP;  // segment offsets in the velocity and pressure arrays
    int offXP;         // segment offset in the outflow values (2 ny - 1 per case)
    int numUV, numP;
    V5Dev v5;          // LHS1 in stencil form
    P5Dev p5;          // LHS2 in stencil form; p5.stop -> the case's Poisson stop code
    double *scV, *scP; // the case's BiCGSTAB scalar blocks (velocity, Poisson)
    int fI, fJ, fncx, fncy;  // control-volume box
    const int *ghostUV;      // the case's own ghost tags (force)
};

// ---- batched kernels (k_ns.hip) ----
void launch_b_robin(const BCase *tab, const BMapDev &m, const double *u, double *xp, hipStream_t s);
void launch_b_rhs1(const BCase *tab, const BMapDev &m, const double *u, const double *Nold, double *N, double *rhs,
                   hipStream_t s);
void launch_b_rhs2(const BCase *tab, const BMapDev &m, const double *uhat, double *rhs2, hipStream_t s);
void launch_b_project(const BCase *tab, const BMapDev &m, double *u, const double *uhat, double *uold,
                      const double *p, hipStream_t s);

// ---- batched BiCGSTAB pieces (k_blas.hip, k_stencil.hip) ----
// which operator of the case table: the velocity (V5) or the Poisson (P5) stencil
enum { BOP_VEL = 0, BOP_POIS = 1 };
// r = b - A x, p = rs = r, partials of (r.r, b.b) per block
// (with mask: only the cases c with mask[c] != 0)
void launch_b_init(int op, const BCase *tab, const BMapDev &m, const double *x, const double *b, double *r, double *p,
                   double *rs, double *part, hipStream_t s, const int *mask = nullptr);
// y = A x with partials of (y.a) [and (y.y)] per block; returns at once for stopped cases
void launch_b_spmv_dot(int op, const BCase *tab, const BMapDev &m, const double *x, double *y, const double *a,
                       bool want_yy, double *part, hipStream_t s);
void launch_b_bicg_s(int op, const BCase *tab, const BMapDev &m, const double *r, const double *AMp, double *s_,
                     double *part, hipStream_t s);
void launch_b_bicg_xr(int op, const BCase *tab, const BMapDev &m, double *x, const double *Mp, const double *Ms,
                      double *r, const double *s_, const double *AMs, const double *rs, double *part,
                      hipStream_t s);
void launch_b_bicg_p(int op, const BCase *tab, const BMapDev &m, double *p, const double *r, const double *AMp,
                     hipStream_t s);
// per case: sum its blocks' partials and update its scalars (one block per case); FIN_INIT
// also sets the case's absolute tolerance rtol * ||b|| in SC_TOL
void launch_b_finish(int op, const BCase *tab, int ncases, const int *first, int nblk, const double *part, int fin,
                     double rtol, int maxit, hipStream_t s, const int *mask = nullptr);
// z = dinv r over the concatenated velocity arrays (stopped cases included)
void launch_b_jacobi(long n, const double *dinv, const double *r, double *z, hipStream_t s);

// ---- batched force (k_ibm.hip): out[6 c + k] as launch_force's out[k] ----
constexpr int BFORCE_BLOCKS = 8;  // blocks per case
void launch_b_force(const BCase *tab, int ncases, const double *u, const double *uold, const double *p, double *part,
                    double *out, hipStream_t s);

// ---- batched finest AMG level (k_stencil.hip) ----
// deflation: partial sums of w_i r_i per block (P map) and, per case, c = sum * inv_wa[c]
// into defl[c]
void launch_b_wdot(const BCase *tab, const BMapDev &m, const double *r, double *part, hipStream_t s);
void launch_b_wdot_finish(const BCase *tab, int ncases, const int *first, const double *part, const double *inv_wa,
                          double *defl, hipStream_t s);
// the tile kernels of the finest level per case (tile map: TX x TY tiles per case)
void launch_b_p5_pre(const BCase *tab, const BMapDev &tiles, float omega, const double *b, float *x2, float *r,
                     hipStream_t s);
void launch_b_p5_pre(const BCase *tab, const BMapDev &tiles, double omega, const double *b, double *x2, double *r,
                     hipStream_t s);
void launch_b_p5_sweep2(const BCase *tab, const BMapDev &tiles, float omega, const float *x, const double *b,
                        double *xo, hipStream_t s);
void launch_b_p5_sweep2(const BCase *tab, const BMapDev &tiles, double omega, const double *x, const double *b,
                        double *xo, hipStream_t s);
// tiles of one case's finest level
int p5_tiles(int nx, int ny);

// The AMG V-cycles of all cases as one preconditioner: the finest level per case in
// stencil form (the tile kernels above), the CSR levels concatenated block-diagonally, the
// dense coarsest solves per case. Built from the cases' own preconditioners, all of one
// precision (64 or 32); with 32 and a pinned operator the constant mode is deflated per case.
class BatchAmg {
  public:
    BatchAmg(const std::vector<AmgPrecond *> &amg, const std::vector<long> &offP, const BCase *tab_dev,
             const std::vector<BCase> &tab_host, const BMapDev &pmap, const int *pfirst, hipStream_t s);
    ~BatchAmg();
    void apply(const double *r, double *z, hipStream_t s);
    int levels() const;

  private:
    struct Impl;
    std::unique_ptr<Impl> im_;
};

class NavierStokesSolver;

// Lockstep time stepping of initialised LUO solvers (static bodies, stencil operators,
// the same AMG precision, constant body velocity 0): the cases' state is copied into the
// batch's arrays, stepped there, and can be copied back (copy_back).
class LockstepBatch {
  public:
    // graphs: run the solves' eight-iteration blocks as captured HIP graphs (one per
    // operator, captured at its first full block and replayed after) instead of launching
    // their kernels one by one
    LockstepBatch(const std::vector<NavierStokesSolver *> &cases, hipStream_t s, bool graphs = true);
    ~LockstepBatch();
    void step();
    // per case: 2 Fx, 2 Fy after the last step (input order)
    void coefficients(std::vector<double> &cd, std::vector<double> &cl);
    // the fields and counters back into the case solvers (u, uhat, uold, p, N, Nold, xp)
    void copy_back();
    // iterations of the last step's solves (largest over the cases) and their totals
    int last_iters(int which) const { return which == 0 ? it1_ : it2_; }
    long long cells() const { return cells_; }

  private:
    struct Impl;
    std::unique_ptr<Impl> im_;
    int it1_ = 0, it2_ = 0;
    long long cells_ = 0;
};

}  // namespace cuibm
