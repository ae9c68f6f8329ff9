// Solver classes mirroring cuIBM's NavierStokesSolver / luo_base / luoIBM /
// fadlunModified API (src/solvers/NavierStokes/NavierStokesSolver.h:145-186,
// luo_base.h:216-243), driving the HIP kernels of this library.
#pragma once
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "case_io.hpp"
#include "embedded.hpp"
#include "ibm.hpp"
#include "kernels.hpp"
#include "krylov.hpp"
#include "logger.hpp"

namespace cuibm {

struct SolverOptions {
    bool output = true;          // write the reference's files into the output folder
    std::string outputFolder;    // default: the case folder
    bool keepDiagnostics = true; // store L and bc1 (not needed by the step itself)
};

class NavierStokesSolver {
  public:
    NavierStokesSolver(ParameterDB *pDB, Domain *dInfo, const SolverOptions &opt = SolverOptions());
    virtual ~NavierStokesSolver();

    virtual void initialise();
    virtual void initialiseLHS();
    virtual void stepTime();
    virtual void writeCommon();
    virtual void writeData();
    bool finished() const { return timeStep >= db->nt; }
    virtual void shutDown();

    virtual void generateRHS1();
    virtual void generateLHS1();
    virtual void generateRHS2();
    virtual void generateLHS2();
    virtual void velocityProjection();
    void solveIntermediateVelocity();
    void solvePoisson();
    virtual std::string name() const { return "Navier-Stokes"; }

    // phase entry points used by the C ABI (reference method names where one exists)
    virtual int phase(const std::string &name);

    // field access in reference layouts; returns the element count or -1
    virtual long fieldCount(const std::string &name) const;
    virtual long getField(const std::string &name, void *out, long n);
    virtual long setField(const std::string &name, const void *in, long n);
    virtual void info(double *out, int n) const;

    hipStream_t stream() const { return stream_; }
    const SolveStat &lastSolve1() const { return stat1; }
    const SolveStat &lastSolve2() const { return stat2; }
    int step() const { return timeStep; }
    long info_cells() const { return (long)nx * ny; }
    // algorithmic HBM bytes of the last step's Poisson solve and whole step
    double poissonBytes() const;
    double velocityBytes() const;
    double stepBytes() const;
    std::string preconditionerName(int which) const;
    // mean time of `reps` launches of one kernel on the live state (HIP events)
    int timeKernel(const std::string &kernel, int reps, double *ms, double *bytes);
    // algorithmic bytes of each launch of one Poisson BiCGSTAB iteration, in launch order
    // (empty unless the Poisson preconditioner is the AMG V-cycle)
    std::vector<double> poissonIterationLaunches() const;
    // LHS2 rows follow the plain formula (no IBM row overrides): the Poisson solve
    // can use the stencil form of the operator
    virtual bool stencilLHS2() const { return !db->csrOperators; }
    AmgPrecond *amg2() const;  // the Poisson AMG preconditioner (throws if another)
    AmgHierarchyInfo amgInfo() const;
    // NavierStokesSolver::CFL (NavierStokes/CFL.cu:10-35): track the largest CFL number
    void CFL();
    const Logger &timers() const { return logger; }

  protected:
    friend class LockstepBatch;
    ParameterDB *db;
    Domain *dom;
    SolverOptions opt;
    hipStream_t stream_ = nullptr;
    int nx = 0, ny = 0, numU = 0, numV = 0, numUV = 0, numP = 0;
    double nu = 0, dt = 0;
    int timeStep = 0, iterationCount1 = 0, iterationCount2 = 0;
    // running totals over all solves since initialise (for throughput accounting)
    double itTotal1 = 0, itTotal2 = 0, bytesTotal1 = 0, bytesTotal2 = 0, solvesTotal2 = 0, restartsTotal = 0;
    SolveStat stat1, stat2;
    Logger logger;  // the reference's named phase timers
    KernelProbe poissonProbe;  // events around each Poisson solve while enabled (live:poisson)
    double cfl_max = 0;
    int cfl_I = 0, cfl_J = 0, cfl_ts = 0;
    DArray<double> cflPart, cflOut;

    DArray<double> gx, gdx, gy, gdy, gxu, gyv;
    DArray<double> u, uhat, uold, pressure, N, Nold, L, bc1, rhs1, rhs2, bc[4];
    DCsr<double> LHS1, LHS2;
    std::unique_ptr<LinOp> op1, op2;  // the velocity and Poisson operators of the two solves
    std::unique_ptr<Preconditioner> PC1, PC2;
    std::unique_ptr<Bicgstab> ks1, ks2;
    std::ofstream iterationsFile;

    GridDev grid() const;
    BcDev bcs() const;
    std::string outFolder() const;
    void initialiseNoBody();
    void buildPreconditioners();
    void projectPlain();
    virtual FadlunRhsDev fadlunRhs() const { return FadlunRhsDev{nullptr, nullptr, nullptr, nullptr, nullptr}; }
    virtual FadlunP2Dev fadlunP2() const { return FadlunP2Dev{nullptr, nullptr, nullptr, nullptr}; }
    virtual FadlunLhsDev fadlunLhs() const {
        return FadlunLhsDev{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    }
    virtual bool centrePin() const { return true; }
    // LHS2 rows of different scales (unit-diagonal interpolation rows beside Poisson rows):
    // the AMG hierarchy then restricts the finest level with P^T D^-1
    virtual bool mixedScaleLHS2() const { return false; }
    // largest coarsest AMG level of the Poisson hierarchy (solved with a dense inverse,
    // whose set-up costs 2 n^3)
    virtual int amgCoarseMax() const { return db->PoissonSolve.amgCoarseMax > 0 ? db->PoissonSolve.amgCoarseMax : 2500; }
    // the Poisson hierarchy is rebuilt every step (its set-up then skips the fused
    // restriction's extra product)
    virtual bool amgRebuiltEachStep() const { return false; }
    DArray<double> *dfield(const std::string &name);
    const DArray<double> *dfield(const std::string &name) const;
    virtual DArray<int> *ifield(const std::string &name) { return nullptr; }
};

// common immersed-boundary layer (luo_base.h:13-244)
class luo_base : public NavierStokesSolver {
  public:
    luo_base(ParameterDB *pDB, Domain *dInfo, const SolverOptions &opt = SolverOptions());
    void initialise() override;
    void initialiseLHS() override {}
    void writeData() override;
    void writeCommon() override;
    void shutDown() override;
    void stepTime() override;
    virtual void _pre_step();
    virtual void _intermediate_velocity() {}
    virtual void _pressure() {}
    virtual void _project_velocity();
    virtual void _update_body();
    virtual void _post_step();
    void moveBody();
    void vivMovement(bool strong);
    virtual void updateSolver();
    void tagPoints();
    void calculateForce();
    void updateRobinBoundary();
    // per-point body velocities ub[k], vb[k] of the body points and the angular velocity
    // omega of the rigid rotation given to the solid nodes (a build extension: the
    // reference gives every point the velocity uB[0]); static bodies only, after
    // initialise()
    void setBodyVelocity(const double *ub, const double *vb, int n, double omega);

    int phase(const std::string &name) override;
    long fieldCount(const std::string &name) const override;
    long getField(const std::string &name, void *out, long n) override;
    void info(double *out, int n) const override;

  protected:
    friend class LockstepBatch;
    // body state (src/bodies.h)
    std::vector<double> bX, bY, bUB, bVB;  // host copies of the body points / velocities
    std::vector<double> bDX, bDY;          // point offsets from the body centre
    bool moving = false;                   // prescribed oscillation or VIV
    double midX = 0, midY = 0, uB0 = 0, vB0 = 0, uBk = 0, vBk = 0;
    double midY0 = 0, centerV = 0, centerV0 = 0;  // transverse motion (VIV)
    double scTol = 0;                      // strongly coupled VIV: relative position change
    int scPasses = 0;                      // passes of the last step
    int startI = 0, startJ = 0, numCellsX = 0, numCellsY = 0;
    double forceX = 0, forceY = 0, fxx = 0, fxy = 0, fxu = 0;
    DArray<double> dbx, dby, dbu, dbv;
    // tags and intercept data (luo_base/cast.cu:7-82)
    DArray<int> ghostTagsUV, hybridTagsUV, hybridTagsUV2, ghostTagsP, hybridTagsP;
    DArray<double> bi_x, bi_y, ip_x, ip_y, bip_x, bip_y, ipp_x, ipp_y;
    DArray<double> dist_int, dist_between, dist_u_body, dist_v_body, uvb;
    DArray<double> ustar, pstar;
    DArray<double> forcePart, forceOut;
    TagRecords rec;
    Stencils gnV, hnV, hnP, gnP;
    DArray<int> solid;
    DArray<int> nlPart, nlCount;  // node-list chunk counts and list lengths
    int nsolid = 0;
    // per-point body velocities (setBodyVelocity): the velocity at each ghost / hybrid
    // node's body intercept, linear along the body segment that holds it, and rigid
    // rotation at the solid nodes; the segment and its weight per node are kept for the
    // adjoint's parameter gradient
    DArray<double> solidUB;
    std::vector<int> gnSeg, hnSeg;      // first point of the segment holding each node's intercept
    std::vector<double> gnW, hnW;       // the intercept's weight on the segment's second point
    std::vector<int> gnNode, hnNode, solidNode;
    bool perPointUB = false;
    double bodyOmega = 0;
    std::ofstream forceFile, midPositionFile;
    std::ofstream forceFullFile;  // the forces at 17 significant digits (no reference counterpart)
    bool tagged = false;

    void boundingBox();
    void buildStencils();
    void readForces();
    DArray<int> *ifield(const std::string &name) override;
    const DArray<int> *ifieldc(const std::string &name) const;
    DArray<double> *ibmField(const std::string &name);
};

// Luo et al. external ghost-node method (luoIBM.h)
class luoIBM : public luo_base {
  public:
    using luo_base::luo_base;
    void initialise() override;
    void initialiseLHS() override;
    void _intermediate_velocity() override;
    void _pressure() override;
    void generateRHS1() override;
    void generateRHS2() override;
    void weightUhat();
    void weightPressure();
    void rhs1GNInterpolation();
    void rhs1HNInterpolation();
    void preRHS2Interpolation();
    void zeroVelocity();
    std::string name() const override { return "LUO"; }
    int phase(const std::string &name) override;

    // Discrete adjoint of J = mean over steps n0..n1 (1-based) of C_D = 2 Fx, for a run of
    // `nsteps` steps from the current state (time step 0, after initialise and
    // setBodyVelocity): the forward run stores checkpoints every `every` steps, then
    // segments are recomputed from their checkpoint and stepped backwards (adjoint.hip).
    // Writes J and, when backward, dJ/dub[k], dJ/dvb[k] of the per-point body velocities
    // and dJ/domega of the solid nodes' rotation; stats[0..5] = forward and backward
    // seconds, forward Poisson and velocity iterations, adjoint Poisson and velocity
    // iterations.
    void adjointRun(int nsteps, int n0, int n1, int every, bool backward, double *J, double *gub, double *gvb,
                    double *gomega, double *stats);
    ~luoIBM() override;

  protected:
    bool centrePin() const override { return false; }
    struct Adjoint;
    std::shared_ptr<Adjoint> adj_;  // built on the first backward run (adjoint.hip)
};

// Luo et al. embedded method (luo_iter.h): the interpolation lives in the matrices
class luoIter : public luo_base {
  public:
    using luo_base::luo_base;
    void initialise() override;
    void initialiseLHS() override;
    void _intermediate_velocity() override;
    void _pressure() override;
    void updateSolver() override;
    std::string name() const override { return "OSC_CYLINDER"; }
    int phase(const std::string &name) override;
    void info(double *out, int n) const override;

  protected:
    bool centrePin() const override { return false; }
    bool stencilLHS2() const override { return false; }  // interpolation rows
    bool mixedScaleLHS2() const override { return true; }
    // a moving body sets the hierarchy up again every step: a coarsest level of at most
    // 60 unknowns is inverted in one workgroup on the device (against up to 2500)
    int amgCoarseMax() const override {
        return db->PoissonSolve.amgCoarseMax > 0 ? db->PoissonSolve.amgCoarseMax : moving ? 60 : 2500;
    }
    bool amgRebuiltEachStep() const override { return moving; }
    long refreshCount = 0;  // Poisson hierarchies set up again on the device (kept pattern)
    // row maps (-1 plain, t hybrid slot, -(t + 2) ghost slot), row buffers of each system,
    // row scales, the systems' nonzero counts, and the slot-anomaly counts summed over
    // every assembly (velocity [0..3], pressure [4..7])
    DArray<int> map1, map2, cnt, rowc, cnt2, rowc2, stats, scanPart, nnzOut;
    DArray<double> rowv, rowv2, ns1, ns2;
    // the plain matrices whose rows the embedded assembly keeps away from the body
    DCsr<double> plain1, plain2;
    bool plainReady = false;
    EmbAsm asm1() const;
    EmbAsm asm2() const;
    void assemble();
};

// modified Fadlun direct forcing (fadlunModified.h)
class fadlunModified : public luo_base {
  public:
    using luo_base::luo_base;
    void initialise() override;
    void initialiseLHS() override;
    void writeData() override;
    void writeCommon() override;
    void _pre_step() override;
    void _intermediate_velocity() override;
    void _pressure() override;
    void _project_velocity() override;
    void generateRHS1() override;
    void generateRHS2() override;
    std::string name() const override { return "FADLUN"; }
    int phase(const std::string &name) override;

  protected:
    FadlunRhsDev fadlunRhs() const override;
    FadlunP2Dev fadlunP2() const override;
    FadlunLhsDev fadlunLhs() const override;
    bool stencilLHS2() const override { return false; }  // cut-cell and ghost rows differ
};

// factory, the switch of src/cuIBM.cu:42-60
std::unique_ptr<NavierStokesSolver> createSolver(ParameterDB *db, Domain *dom, const SolverOptions &opt);

}  // namespace cuibm
