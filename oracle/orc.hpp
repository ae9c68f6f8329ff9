// ORACLE — test infrastructure only.
//
// This is the CPU (FP64, OpenMP) restatement of cuIBM's per-time-step hot path
// that the parity tests and the bench's cpu_baseline leg compare against.
// Nothing in the product (cuibm_amd/) links, loads or calls it.  Every function
// cites the reference file:line it restates; the reference itself (CUDA + CUSP
// 0.5.1) cannot be compiled here, so this restatement is the oracle.
//
// Pinning (what checks each part, beyond re-typing the formula):
//   grid generator      reference-held grid sizes (scripts/validation/performance.py:22,25)
//                       + closed forms (sum dx = end-start per block, geometric ratio)
//   N, L, bc1, rhs1     closed forms (L of linear fields = 0 on uniform interiors,
//                       L of quadratics exact, N(const) = 0), mirror symmetry
//   LHS1, LHS2          row-sum invariants, nnz formulas of NavierStokes/cast.cu:23-24,
//                       equality with a Kronecker-product Laplacian on tiny grids
//   RHS2 + projection   div(u) = 0 after an exact solve, except at pinned rows
//   BiCGSTAB            dense solve on small nonsymmetric systems
//   SA-AMG              Galerkin identity A_c = P^T A P, V-cycle contraction factor
//   tagging             brute-force point-in-polygon classification invariants
//   interpolation       exact reproduction of bilinear fields, Neumann rows
//   CV force            uniform flow gives F = 0; mirror symmetry gives Fy = 0
// The racy neighbour writes of the tagging kernels have no unique reference
// answer; the oracle fixes them by the rule stated in orc_ibm.cpp.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace orc {

// boundary locations, reference order (src/boundaryTypes.h:8-13)
enum { XMINUS = 0, XPLUS = 1, YMINUS = 2, YPLUS = 3 };
// solver types, reference order (src/types.h:4-11)
enum { NAVIERSTOKES = 0, FADLUN = 1, OSC = 2, FSI = 3, LUO = 4 };
// preconditioners, reference order (src/preconditioner.h:23-29)
enum { P_NONE = 0, P_DIAGONAL = 1, P_SMOOTHED_AGGREGATION = 2, P_AINV = 3 };

struct Grid {
    int nx = 0, ny = 0;
    double mid_h = 0;
    std::vector<double> x, dx, y, dy, xu, yv;  // xv = x, yu = y
};

struct Csr {
    int n = 0;
    std::vector<int> rp, ci;
    std::vector<double> v;
    int nnz() const { return (int)ci.size(); }
};

// y = A x
void spmv(const Csr &A, const double *x, double *y);

struct AmgLevel {
    Csr A, P, R;
    std::vector<double> dinv;
};
struct Amg {
    std::vector<AmgLevel> lv;
    int nc = 0;
    std::vector<double> coarse_inv;  // dense row-major inverse of the coarsest A
    double omega = 0.8;
    int npre = 2, npost = 2;
    // restrict the finest level with P^T D^-1 (the hierarchy of the row-scaled D^-1 A);
    // set for matrices whose rows mix scales (the embedded method's Poisson matrix)
    bool scale0 = false;
    // coarse_max <= 0: amg_default_coarse_max (2500; the tests lower it to reach
    // several levels on small matrices)
    void setup(const Csr &A, int coarse_max = 0, int max_levels = 25);
    void apply(const double *b, double *x);  // one V-cycle, zero initial guess
};

extern int amg_default_coarse_max;

struct Precond {
    int type = P_NONE;
    std::vector<double> dinv;
    Amg amg;
    void build(const Csr &A, int type, bool amg_scale0 = false, int amg_coarse_max = 0);
    void apply(const double *r, double *z);
};

// cusp::krylov::bicgstab + cusp::monitor semantics; returns iteration count
struct SolveStat { int iters = 0; double resid = 0; bool converged = false; };
SolveStat bicgstab(const Csr &A, double *x, const double *b, Precond &M, double rtol, int maxit);

// grid
int axis_cells(double start, double mid_h, int nsub, const double *ends, const double *ratios);
void axis_fill(double start, double mid_h, int nsub, const double *ends, const double *ratios,
               double *x, double *dx);

struct Params {
    int solver_type = NAVIERSTOKES;
    double nu = 0.01, dt = 0.02;
    int nt = 100;
    double uInitial = 1, vInitial = 0, uPerturb = 0, vPerturb = 0;
    double bc_u[4] = {0, 0, 0, 0}, bc_v[4] = {0, 0, 0, 0};
    int precond1 = P_DIAGONAL, precond2 = P_DIAGONAL;
    double tol1 = 1e-5, tol2 = 1e-5;
    int maxit1 = 10000, maxit2 = 20000;
    double scaleCV = 2.0;
    // prescribed in-line oscillation of the body (src/io/parseBodiesFile.cu:60-104)
    double xCoeff = 0, uCoeff = 0, xPhase = 0, uPhase = 0, xfreq = 0;
    // flow-induced transverse motion (src/io/parseSimulationFile.cu:80-142): 0 none,
    // 1 loosely coupled, 2 strongly coupled; reduced velocity (an integer there)
    int viv = 0, ured = 3;
};

struct Body {
    int n = 0;
    std::vector<double> x, y, uB, vB, uB0, vB0, uBk, vBk;
    std::vector<double> dx, dy;  // point offsets from the body centre
    bool moving = false;
    double midX = 0, midY = 0, midY0 = 0;
    double centerV = 0, centerV0 = 0;  // transverse velocity of the body centre (VIV)
    int startI = 0, startJ = 0, numCellsX = 0, numCellsY = 0;
    double forceX = 0, forceY = 0;
};

struct Solver {
    Params P;
    Grid G;
    Body B;
    int nx, ny, numU, numV, numUV, numP;
    int timeStep = 0, iter1 = 0, iter2 = 0;
    double res1 = 0, res2 = 0;
    double fxx = 0, fxy = 0, fxu = 0;
    std::vector<double> u, uhat, uold, p, p_old, N, Nold, L, bc1, rhs1, rhs2, bc[4];
    Csr LHS1, LHS2;
    Precond PC1, PC2;
    // IBM state (luo_base/cast.cu:7-82)
    std::vector<int> ghostTagsUV, hybridTagsUV, hybridTagsUV2, ghostTagsP, hybridTagsP;
    std::vector<double> bi_x, bi_y, ip_x, ip_y, bip_x, bip_y, ipp_x, ipp_y;
    std::vector<double> dist_int, dist_between, dist_u_body, dist_v_body, uv;
    std::vector<double> ustar, pstar, image_point_u, dpdn;
    std::vector<double> qcoef[4];
    std::vector<int> qidx[4];
    // corner values of the last velocity interpolation (body-intercept corners hold uB)
    std::vector<double> qval[4];
    // pressure interpolation per node: corner indices, weights at the node (hybrid) or
    // image point (ghost), corner values (Neumann corners hold their normal-derivative value)
    std::vector<double> pqcoef[4], pqval[4];
    std::vector<int> pqidx[4];
    // Luo embedded (luo_iter): blending weights and the row scale / added right-hand side
    std::vector<double> alpha1, alpha2, ns_rhs1, interp_rhs1, ns_rhs2, interp_rhs2;
    // slot accounting of the embedded assembly, summed over all steps:
    // [0] 5-point slots a ghost row leaves unwritten, [1] hybrid tail slots never written,
    // [2] hybrid tail slots written more than once (last write kept), [3] interpolation
    // corners outside the 3x3 stencil (dropped); LHS1 in [0..3], LHS2 in [4..7]
    long emb_stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    bool has_body = false;

    void init(const Params &P, const Grid &G, const std::vector<double> &bx,
              const std::vector<double> &by);
    // plain NS pieces
    void generateN();
    void generateL();
    void generateBC1();
    void generateRHS();
    void buildLHS1();
    void buildLHS2();
    void rhs2_nobody();
    void project_nobody();
    void solve1();
    void solve2();
    // IBM pieces
    void boundingBox();
    void tagPoints();
    void robin();
    void gn_velocity(std::vector<double> &f, bool set);
    void set_inside(std::vector<double> &f);
    // per-point body velocities (no reference counterpart): per velocity node, the body
    // value its ghost / hybrid interpolation and the solid-node update use; empty = uB[0]
    std::vector<double> node_ub;
    void set_body_velocity(const double *ub, const double *vb, double omega);
    void hn_velocity(const std::vector<double> &f);
    void weight_uhat();
    void hn_pressure();
    void weight_p();
    void gn_pressure(bool set);
    void fadlun_rhs1();
    void fadlun_rhs2();
    void fadlun_project();
    void force();
    // steps
    void step();
    void step_ns();
    void step_luo();
    void move_body();
    // luo_base::stepTime (luo_base.cu:208-285): pre-step, one pass (or strongly coupled
    // passes until the body position settles), step counter
    void luo_pre_step();
    void luo_pass();
    void update_body();
    void viv_move(bool strong);
    int sc_passes = 0;      // passes of the last step
    double sc_tol = 0;
    void step_fadlun();
    // Luo embedded
    void step_embedded();
    void embedded_pass();
    void alpha_uv();
    void alpha_pressure();
    void buildLHS1_embedded();
    void buildLHS2_embedded();
    void update_rhs1_embedded();
    void update_rhs2_embedded();
    std::vector<double> *field(const std::string &name);
    std::vector<int> *ifield(const std::string &name);
};

}  // namespace orc

// one transverse spring-mass step of a VIV body (luo_base/move.cu:38-93)
extern "C" void orc_viv_update(double yold, double vold, double Cy, int Ured, double dt, int strong, double ycur,
                               double *vnew, double *ynew);
