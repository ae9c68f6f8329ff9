// The pressure-Poisson operator LHS2 of the plain and Luo solvers as a 5-point
// stencil over four 1-D coefficient arrays, instead of a CSR matrix.
//
// LHS2 (NavierStokes/kernels/LHS2.cu:23-173, luoIBM/kernels/LHS2.cu:12-62) couples
// cell (I,J) to its east neighbour with -dt/(dx_I (dx_I + dx_{I+1})/2), and likewise
// west, north and south; the diagonal is the sum of the four magnitudes (walls drop
// the missing neighbour), doubled at cell 0 and, for the plain solver, at the centre
// cell. Every coefficient depends on I alone (east/west) or J alone (north/south),
// so O(nx + ny) doubles describe the whole matrix. Products are evaluated in the CSR
// row order S, W, C, E, N with the same expressions as the assembled matrix, so
// y = A x is bit-identical to the CSR product.
//
// The V-cycle's finest level runs on this form with two fused LDS-tiled kernels
// (pre-smoothing + residual; both post-smoothing sweeps after a row-parallel
// prolongation) instead of six CSR launches; each block owns a 64 x 16 tile of cells
// and computes the intermediate iterates on a halo, so the kernels read b (and x)
// once and write only their results. Blocks are assigned tiles in XCD-contiguous bands (the hardware deals
// workgroups to the 8 XCDs round-robin), so the halo rows a block needs sit in its
// own XCD's L2.
#pragma once
#include "kernels.hpp"
#include "krylov.hpp"

namespace cuibm {

class PoissonStencil : public LinOp {
  public:
    PoissonStencil(const GridDev &g, double dt, bool centre_pin, hipStream_t s);
    int n() const override { return nx_ * ny_; }
    void spmv_dot(const double *x, double *y, const double *a, bool want_yy, double *part,
                  hipStream_t s) const override;
    void init(const double *x, const double *b, double *r, double *p, double *rs, double *part,
              hipStream_t s) const override;
    double spmv_bytes() const override { return 16.0 * n() + 32.0 * (nx_ + ny_); }
    P5Dev dev() const {
        P5Dev d{nx_, ny_, centre_, cE_.p(), cW_.p(), cN_.p(), cS_.p(), dinv_.p(), stop_};
        d.dx = dx_;
        d.dy = dy_;
        return d;
    }
    // y = A x without dot partials (tests)
    void apply(const double *x, double *y, hipStream_t s) const;

  private:
    int nx_, ny_, centre_;
    const double *dx_, *dy_;
    DArray<double> cE_, cW_, cN_, cS_, dinv_;
};

// The velocity operator LHS1 = I - 0.5 dt nu Lap (NavierStokes/kernels/LHS1.cu:24-351) of
// the plain and Luo-external solvers in the same form: it is block diagonal in (u, v), and
// in each block the east/west couplings depend on I alone and the north/south ones on J
// alone. Per block and direction, `off` holds the coupling magnitude (the entry is its
// negative) and `dg` the term the row's diagonal adds for that direction, which differs
// from `off` at a wall (the reference's ghost-cell diagonal). The diagonal is
// ((1 + dgE[I]) + dgW[I]) + dgN[J]) + dgS[J], the assembly's order, so products equal
// the CSR product with LHS1 bit for bit.
struct V5Block {
    int ni, nj, off0;                   // nodes per row and per column, first packed index
    const double *oE, *oW, *oN, *oS;    // coupling magnitudes (o*[I] or o*[J])
    const double *dE, *dW, *dN, *dS;    // diagonal terms
};
struct V5Dev {
    V5Block u, v;
    const double *stop = nullptr;  // the kernels return at once while *stop != 0
};

class VelocityStencil : public LinOp {
  public:
    VelocityStencil(const GridDev &g, double nu, double dt, hipStream_t s);
    int n() const override { return nu_ + nv_; }
    void spmv_dot(const double *x, double *y, const double *a, bool want_yy, double *part,
                  hipStream_t s) const override;
    void init(const double *x, const double *b, double *r, double *p, double *rs, double *part,
              hipStream_t s) const override;
    double spmv_bytes() const override { return 16.0 * n() + 64.0 * (nx_ + ny_); }
    V5Dev dev() const;

  private:
    int nx_, ny_, nu_, nv_;
    DArray<double> c_;  // the 16 coefficient arrays, packed
};

// finest-level V-cycle pieces (omega-weighted Jacobi, D = diag(A)):
//   pre:    x1 = omega D^-1 b, x2 = x1 + omega D^-1 (b - A x1), r = b - A x2
//   sweep2: two sweeps x <- x + omega D^-1 (b - A x) (after the prolongation)
// with an enabled probe, the launch carries a start/stop event pair of the probe
void launch_p5_pre(const P5Dev &A, double omega, const double *b, double *x2, double *r, hipStream_t s,
                   KernelProbe *probe = nullptr);
void launch_p5_sweep2(const P5Dev &A, double omega, const double *x, const double *b, double *xo, hipStream_t s);
// the same in FP32 (the V-cycle with FP32 level vectors): b is the FP64 Krylov vector,
// x2 and r are FP32 level vectors, the result xo is written as FP64; D^-1 is A.dinv32
void launch_p5_pre(const P5Dev &A, float omega, const double *b, float *x2, float *r, hipStream_t s,
                   KernelProbe *probe = nullptr);
void launch_p5_sweep2(const P5Dev &A, float omega, const float *x, const double *b, double *xo, hipStream_t s);
// partial sums of w_i r_i (w = cell areas) over RED_GRID blocks, for the deflation
// constant that the pre-smoothing kernel forms (P5Dev::defl_part); returns at once while
// the Krylov solve's stop code is set
void launch_p5_wdot(const P5Dev &A, const double *r, double *part, hipStream_t s);
// one sweep xn = x + omega D^-1 (b - A x) (timing of the finest-level smoother)
void launch_p5_sweep(const P5Dev &A, double omega, const double *x, const double *b, double *xn, hipStream_t s);

}  // namespace cuibm
