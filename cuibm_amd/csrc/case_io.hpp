// Case front end: parameter database, stretched grid, body points and the on-disk
// output formats of cuIBM, restated for the C++ host side of the library.
#pragma once
#include <string>
#include <vector>

namespace cuibm {

// reference enum orders: src/types.h:4-11, src/preconditioner.h:23-29,
// src/bcTypes.h:8-15, src/boundaryTypes.h:8-13
enum SolverType { NAVIERSTOKES = 0, FADLUN = 1, OSC = 2, FSI = 3, LUO = 4 };
enum PreconditionerType { NONE = 0, DIAGONAL = 1, SMOOTHED_AGGREGATION = 2, AINV = 3 };
enum BcType { DIRICHLET = 0, NEUMANN = 1, CONVECTIVE = 2, PERIODIC = 3, SPECIAL = 4 };
enum Boundary { XMINUS = 0, XPLUS = 1, YMINUS = 2, YPLUS = 3 };

struct BoundaryCondition {
    BcType type = DIRICHLET;
    double value = 0.0;
};

// one immersed body (src/body.h:14-114); coefficients absent from bodies.yaml are 0
struct BodyDesc {
    int numPoints = 0;
    std::vector<double> X, Y;
    double xCoefficient = 0, yCoefficient = 0, uCoefficient = 0, vCoefficient = 0;
    double xfrequency = 0, yfrequency = 0, xPhase = 0, yPhase = 0, uPhase = 0, vPhase = 0;
};

struct LinearSolverParams {
    std::string solver = "CG";  // parsed, ignored: the solve is always BiCGSTAB
    PreconditionerType preconditioner = DIAGONAL;
    double tolerance = 1e-5;
    int maxIterations = 10000;
    // 64: FP64 smoothed-aggregation hierarchy; 32: V-cycle in FP32 (matrix values, level
    // vectors, smoother arithmetic); 3264: FP32 matrix values with FP64 vectors
    int amgPrecision = 64;
    // largest coarsest AMG level (a dense inverse); 0: the solver's default
    int amgCoarseMax = 0;
    // BiCGSTAB restarts from the true residual each time its residual has fallen by this
    // factor (reliable updates); < 0: the preconditioner's default (1e-3 for the FP32
    // V-cycle, never otherwise), 0: never
    double restartReduction = -1.0;
    // a moving body's hierarchy whose matrix kept its pattern: set up again numerically on
    // the device (true) or built anew on the host (false)
    bool amgRefresh = true;
};

// typed restatement of the parameterDB filled by src/io/io.cu:145-197 and the
// parse*File functions
struct ParameterDB {
    std::string caseFolder = ".";
    int deviceNumber = 0;
    // -operators csr: stored CSR operators for both solves and every AMG level, no stencil
    // forms (no reference counterpart; the straightforward form, for comparison)
    bool csrOperators = false;
    // flow
    double nu = 0.01, uInitial = 1.0, vInitial = 0.0, uPerturb = 0.0, vPerturb = 0.0;
    BoundaryCondition bc[4][2];
    std::vector<BodyDesc> bodies;
    // simulation
    double dt = 0.02, scaleCV = 2.0;
    int nt = 100, nsave = 100, startStep = 0, VIV = 0, Ured = 3;
    bool restart = false;
    SolverType solverType = NAVIERSTOKES;
    LinearSolverParams velocitySolve, PoissonSolve;
};

// computational grid (src/domain.h:17-34); xv == x and yu == y are kept as copies
struct Domain {
    int nx = 0, ny = 0;
    double mid_h = 0;
    std::vector<double> x, dx, xv, xu, y, dy, yu, yv;
};

// src/io/io.cu:108-138: defaults, -caseFolder/-deviceNumber, the four case files,
// then the command-line overrides (-nu -uPerturb -vPerturb -scaleCV -nsave -nt -dt
// -velocityTol -poissonTol, plus -amgPrecision)
void readInputs(const std::string &caseFolder, int argc, const char *const *argv, ParameterDB &DB, Domain &D);

// one direction of the stretched grid (src/io/parseDomainFile.cu:28-222)
int gridCells(double start, double mid_h, const std::vector<double> &ends, const std::vector<double> &ratios);
void gridFill(double start, double mid_h, const std::vector<double> &ends, const std::vector<double> &ratios,
              std::vector<double> &x, std::vector<double> &dx);

// circle points (src/io/parseBodiesFile.cu:145-178)
void circlePoints(double cx, double cy, double R, int numPoints, bool selfSize, double mid_h, BodyDesc &b);

// output formats (src/io/io.cu:406-510)
void makeDirectory(const std::string &path);
void writeGrid(const std::string &caseFolder, const Domain &D);
void writeData(const std::string &caseFolder, int n, const std::vector<double> &q, const std::vector<double> &lambda);
void writeInfoFile(const ParameterDB &DB, const Domain &D, const std::string &folder);
void writeBodies(const std::string &caseFolder, int timeStep, const std::vector<double> &bx, const std::vector<double> &by);

const char *solverTypeName(SolverType t);

}  // namespace cuibm
