#include "case_io.hpp"

#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

#include "yaml_lite.hpp"

namespace cuibm {

namespace {

bool file_exists(const std::string &p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}

// src/io/parseSimulationFile.cu:33-65
PreconditionerType preconditionerFromString(const std::string &s) {
    if (s == "NONE") return NONE;
    if (s == "DIAGONAL") return DIAGONAL;
    if (s == "SMOOTHED_AGGREGATION") return SMOOTHED_AGGREGATION;
    if (s == "AINV") return AINV;
    return NONE;
}

SolverType solverTypeFromString(const std::string &s) {
    if (s == "NAVIER_STOKES_SOLVER") return NAVIERSTOKES;
    if (s == "FADLUN") return FADLUN;
    if (s == "OSC_CYLINDER") return OSC;
    if (s == "FSI") return FSI;
    if (s == "LUO") return LUO;
    std::fprintf(stderr, "\nWARNING: unrecognized solverType\nSeting to default: NavierStokes\n\n");
    return NAVIERSTOKES;
}

Boundary boundaryFromString(const std::string &s) {
    if (s == "xMinus") return XMINUS;
    if (s == "xPlus") return XPLUS;
    if (s == "yMinus") return YMINUS;
    if (s == "yPlus") return YPLUS;
    throw std::runtime_error("invalid boundary condition location '" + s + "'");
}

BcType bcTypeFromString(const std::string &s) {
    if (s == "DIRICHLET") return DIRICHLET;
    if (s == "NEUMANN") return NEUMANN;
    if (s == "CONVECTIVE") return CONVECTIVE;
    if (s == "PERIODIC") return PERIODIC;
    if (s == "SPECIAL") return SPECIAL;
    throw std::runtime_error("invalid boundary condition type '" + s + "'");
}

// src/io/parseSimulationFile.cu:71-182; absent optional keys keep their defaults,
// and preconditioner/tolerance/maxIterations carry over between solver entries
void parseSimulation(const YamlNode &node, ParameterDB &DB) {
    DB.dt = node["dt"].as_double();
    DB.nsave = node["nsave"].as_int();
    DB.nt = node["nt"].as_int();
    if (node.has("SolverType")) DB.solverType = solverTypeFromString(node["SolverType"].as_string());
    if (node.has("restart")) DB.restart = node["restart"].as_bool();
    if (node.has("startStep")) DB.startStep = node["startStep"].as_int();
    if (node.has("scaleCV")) DB.scaleCV = node["scaleCV"].as_double();
    if (node.has("VIV")) DB.VIV = node["VIV"].as_int();
    if (node.has("Ured")) DB.Ured = node["Ured"].as_int();
    std::string prec = "DIAGONAL";
    double tol = 1e-5;
    int maxIter = 10000, amgPrec = 64;
    if (!node.has("linearSolvers")) return;
    const YamlNode &solvers = node["linearSolvers"];
    for (size_t i = 0; i < solvers.size(); i++) {
        const YamlNode &s = solvers[i];
        const std::string system = s["system"].as_string();
        LinearSolverParams p;
        p.solver = s["solver"].as_string();
        if (s.has("preconditioner")) prec = s["preconditioner"].as_string();
        if (s.has("tolerance")) tol = s["tolerance"].as_double();
        if (s.has("maxIterations")) maxIter = s["maxIterations"].as_int();
        if (s.has("amgPrecision")) {
            const std::string a = s["amgPrecision"].as_string();
            if (a == "FP32" || a == "32" || a == "float") amgPrec = 32;
            else if (a == "FP32VALUES" || a == "3264") amgPrec = 3264;
            else amgPrec = 64;
        }
        p.preconditioner = preconditionerFromString(prec);
        p.tolerance = tol;
        p.maxIterations = maxIter;
        p.amgPrecision = amgPrec;
        if (system == "velocity") DB.velocitySolve = p;
        else if (system == "Poisson") DB.PoissonSolve = p;
    }
}

// src/io/parseFlowFile.cu:79-128
void parseFlow(const YamlNode &node, ParameterDB &DB) {
    DB.nu = node["nu"].as_double();
    DB.uInitial = node["initialVelocity"][0].as_double();
    DB.vInitial = node["initialVelocity"][1].as_double();
    if (node.has("initialPerturbation")) {
        DB.uPerturb = node["initialPerturbation"][0].as_double();
        DB.vPerturb = node["initialPerturbation"][1].as_double();
    }
    if (!node.has("boundaryConditions")) return;
    const YamlNode &bcs = node["boundaryConditions"];
    for (size_t i = 0; i < bcs.size(); i++) {
        const YamlNode &b = bcs[i];
        const Boundary loc = boundaryFromString(b["location"].as_string());
        DB.bc[loc][0] = {bcTypeFromString(b["u"][0].as_string()), b["u"][1].as_double()};
        DB.bc[loc][1] = {bcTypeFromString(b["v"][0].as_string()), b["v"][1].as_double()};
    }
}

// src/io/parseDomainFile.cu:28-110 + 355-378
void parseDomain(const YamlNode &doc, Domain &D) {
    for (size_t i = 0; i < doc.size(); i++) {
        const YamlNode &node = doc[i];
        const std::string dir = node["direction"].as_string();
        const double start = node["start"].as_double();
        const double mid_h = node["mid_h"].as_double();
        D.mid_h = mid_h;
        std::vector<double> ends, ratios;
        const YamlNode &sd = node["subDomains"];
        for (size_t k = 0; k < sd.size(); k++) {
            ends.push_back(sd[k]["end"].as_double());
            ratios.push_back(sd[k]["stretchRatio"].as_double());
        }
        if (dir == "x") {
            gridFill(start, mid_h, ends, ratios, D.x, D.dx);
            D.nx = (int)D.x.size();
            D.xv = D.x;
        } else if (dir == "y") {
            gridFill(start, mid_h, ends, ratios, D.y, D.dy);
            D.ny = (int)D.y.size();
            D.yu = D.y;
        }
    }
    D.xu.resize(D.nx - 1);
    D.yv.resize(D.ny - 1);
    for (int i = 0; i < D.nx - 1; i++) D.xu[i] = D.x[i] + D.dx[i] / 2;
    for (int j = 0; j < D.ny - 1; j++) D.yv[j] = D.y[j] + D.dy[j] / 2;
}

// src/io/parseBodiesFile.cu:30-178 (point files and circles)
void parseBodies(const YamlNode &doc, const std::string &folder, ParameterDB &DB) {
    for (size_t i = 0; i < doc.size(); i++) {
        const YamlNode &node = doc[i];
        BodyDesc b;
        auto opt = [&](const char *k, double &v) {
            if (node.has(k)) v = node[k].as_double();
        };
        opt("xCoefficient", b.xCoefficient);
        opt("xPhase", b.xPhase);
        opt("yCoefficient", b.yCoefficient);
        opt("yPhase", b.yPhase);
        opt("uCoefficient", b.uCoefficient);
        opt("uPhase", b.uPhase);
        opt("vCoefficient", b.vCoefficient);
        opt("vPhase", b.vPhase);
        opt("xfrequency", b.xfrequency);
        opt("yfrequency", b.yfrequency);
        const std::string type = node["type"].as_string();
        if (type == "points") {
            const std::string fname = folder + "/bodyFiles/" + node["pointsFile"].as_string();
            std::ifstream f(fname);
            if (!f) throw YamlError("cannot open body points file " + fname);
            f >> b.numPoints;
            b.X.resize(b.numPoints);
            b.Y.resize(b.numPoints);
            for (int k = 0; k < b.numPoints; k++) f >> b.X[k] >> b.Y[k];
        } else if (type == "circle") {
            const YamlNode &o = node["circleOptions"];
            const bool selfSize = o.size() > 4 ? o[4].as_bool() : false;
            const double mid_h = o.size() > 5 ? o[5].as_double() : 0.0;
            circlePoints(o[0].as_double(), o[1].as_double(), o[2].as_double(), o[3].as_int(), selfSize, mid_h, b);
        } else {
            throw std::runtime_error("unknown body type '" + type + "'");
        }
        DB.bodies.push_back(b);
    }
}

int cellsInBlock(double start, double end, double mid_h, double r) {
    int numCells;
    if (r < 1) {
        const double a = mid_h / ((start - end) * (r - 1));
        numCells = (int)(std::log(1 / a + 1) / std::log(r));
    } else if (r == 1) {
        numCells = (int)((end - start) / mid_h);
    } else {
        const double a = mid_h / ((end - start) * (r - 1));
        numCells = (int)(std::log(1 / a + 1) / std::log(r));
    }
    return numCells;
}

}  // namespace

int gridCells(double start, double mid_h, const std::vector<double> &ends, const std::vector<double> &ratios) {
    int count = 0;
    for (size_t i = 0; i < ends.size(); i++) {
        count += cellsInBlock(start, ends[i], mid_h, ratios[i]);
        start = ends[i];
    }
    return count;
}

void gridFill(double start, double mid_h, const std::vector<double> &ends, const std::vector<double> &ratios,
              std::vector<double> &x, std::vector<double> &dx) {
    const int total = gridCells(start, mid_h, ends, ratios);
    x.assign(total, 0.0);
    dx.assign(total, 0.0);
    int beg = 0;
    for (size_t b = 0; b < ends.size(); b++) {
        const double end = ends[b], r = ratios[b];
        const int n = cellsInBlock(start, end, mid_h, r);
        if (std::fabs(r - 1.0) < 1.0e-6) {
            const double h = (end - start) / n;
            for (int j = beg; j < beg + n; j++) {
                dx[j] = h;
                x[j] = (j == beg) ? start + h / 2 : x[j - 1] + 0.5 * dx[j] + 0.5 * dx[j - 1];
            }
        } else if (end < 0) {
            // cells shrink towards `end`: fill backwards from the core
            const double h = (end - start) * (r - 1) / (std::pow(r, n) - 1);
            for (int m = 0; m < n; m++) {
                const int j = beg + n - 1 - m;
                if (m == 0) {
                    x[j] = end - h / 2.0;
                    dx[j] = h;
                } else {
                    dx[j] = h * std::pow(r, m);
                    x[j] = x[j + 1] - 0.5 * dx[j] - 0.5 * dx[j + 1];
                }
            }
        } else {
            const double h = (end - start) * (r - 1) / (std::pow(r, n) - 1);
            for (int j = beg; j < beg + n; j++) {
                dx[j] = h * std::pow(r, j - beg);
                x[j] = (j == beg) ? start + h * std::pow(r, j - beg) / 2 : x[j - 1] + 0.5 * dx[j] + 0.5 * dx[j - 1];
            }
        }
        beg += n;
        start = end;
    }
}

void circlePoints(double cx, double cy, double R, int numPoints, bool selfSize, double mid_h, BodyDesc &b) {
    if (selfSize) {
        numPoints = (int)(2.0 * R * M_PI / mid_h);
        if (numPoints % 2 != 0) numPoints++;
    }
    b.numPoints = numPoints;
    b.X.resize(numPoints);
    b.Y.resize(numPoints);
    // two loops: a fused sincos call can differ from cos alone in the last bit
    for (int i = 0; i < numPoints; i++) b.X[i] = cx + R * std::cos(i * 2 * M_PI / numPoints);
    for (int i = 0; i < numPoints; i++) b.Y[i] = cy + R * std::sin(i * 2 * M_PI / numPoints);
}

void readInputs(const std::string &caseFolder, int argc, const char *const *argv, ParameterDB &DB, Domain &D) {
    DB = ParameterDB();
    D = Domain();
    DB.caseFolder = caseFolder;
    for (int i = 0; i + 1 < argc; i++) {
        if (std::strcmp(argv[i], "-caseFolder") == 0) DB.caseFolder = argv[++i];
        else if (std::strcmp(argv[i], "-deviceNumber") == 0) DB.deviceNumber = std::atoi(argv[++i]);
    }
    const std::string f = DB.caseFolder;
    {
        const YamlNode doc = yaml_load(f + "/simParams.yaml");
        for (size_t i = 0; i < doc.size(); i++) parseSimulation(doc[i], DB);
    }
    {
        const YamlNode doc = yaml_load(f + "/flow.yaml");
        for (size_t i = 0; i < doc.size(); i++) parseFlow(doc[i], DB);
    }
    parseDomain(yaml_load(f + "/domain.yaml"), D);
    if (file_exists(f + "/bodies.yaml")) parseBodies(yaml_load(f + "/bodies.yaml"), f, DB);
    // src/io/io.cu:234-299
    for (int i = 0; i + 1 < argc; i++) {
        const std::string a = argv[i];
        const char *v = argv[i + 1];
        if (a == "-nu") { DB.nu = std::atof(v); i++; }
        else if (a == "-uPerturb") { DB.uPerturb = std::atof(v); i++; }
        else if (a == "-vPerturb") { DB.vPerturb = std::atof(v); i++; }
        else if (a == "-scaleCV") { DB.scaleCV = std::atof(v); i++; }
        else if (a == "-nsave") { DB.nsave = std::atoi(v); i++; }
        else if (a == "-nt") { DB.nt = std::atoi(v); i++; }
        else if (a == "-dt") { DB.dt = std::atof(v); i++; }
        else if (a == "-velocityTol") { DB.velocitySolve.tolerance = std::atof(v); i++; }
        else if (a == "-poissonTol") { DB.PoissonSolve.tolerance = std::atof(v); i++; }
        else if (a == "-amgPrecision") {
            const int prec = std::atoi(v);
            DB.PoissonSolve.amgPrecision = (prec == 32 || prec == 3264) ? prec : 64;
            i++;
        }
        else if (a == "-amgCoarseMax") {
            const int m = std::atoi(v);
            if (m < 16 || m > 6144) throw YamlError("-amgCoarseMax takes 16..6144");
            DB.PoissonSolve.amgCoarseMax = m;
            i++;
        }
        else if (a == "-poissonRestart") {
            char *end = nullptr;
            const double f = std::strtod(v, &end);
            if (end == v || *end != '\0' || !(f >= 0.0 && f < 1.0))
                throw YamlError("-poissonRestart takes a factor in [0, 1)");
            DB.PoissonSolve.restartReduction = f;
            i++;
        }
        else if (a == "-amgRefresh") {
            const std::string o = v;
            if (o != "0" && o != "1") throw YamlError("-amgRefresh takes 0 or 1, not '" + o + "'");
            DB.PoissonSolve.amgRefresh = o == "1";
            i++;
        }
        else if (a == "-poissonPreconditioner") { DB.PoissonSolve.preconditioner = preconditionerFromString(v); i++; }
        else if (a == "-operators") {
            const std::string o = v;
            if (o != "csr" && o != "stencil") throw YamlError("-operators takes csr or stencil, not '" + o + "'");
            DB.csrOperators = o == "csr";
            i++;
        }
    }
    if (D.nx < 3 || D.ny < 3) throw std::runtime_error("grid must have at least 3 cells in each direction");
}

void makeDirectory(const std::string &path) { mkdir(path.c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH); }

// src/io/io.cu:428-444: int32 n, then n+1 doubles. The reference copies the
// centres into an n-element buffer and writes n+1 of them; the extra value here is
// the last cell's outer face, so the file keeps the reference size and layout.
void writeGrid(const std::string &caseFolder, const Domain &D) {
    std::ofstream file(caseFolder + "/grid", std::ios::binary);
    std::vector<double> x(D.x), y(D.y);
    x.push_back(D.x[D.nx - 1] + 0.5 * D.dx[D.nx - 1]);
    y.push_back(D.y[D.ny - 1] + 0.5 * D.dy[D.ny - 1]);
    file.write((const char *)&D.nx, sizeof(int));
    file.write((const char *)x.data(), (D.nx + 1) * sizeof(double));
    file.write((const char *)&D.ny, sizeof(int));
    file.write((const char *)y.data(), (D.ny + 1) * sizeof(double));
}

// src/io/io.cu:460-488: <case>/%07d/q and lambda, int32 N then N doubles
void writeData(const std::string &caseFolder, int n, const std::vector<double> &q, const std::vector<double> &lambda) {
    std::ostringstream out;
    out << caseFolder << '/' << std::setfill('0') << std::setw(7) << n;
    const std::string path = out.str();
    makeDirectory(path);
    {
        std::ofstream file(path + "/q", std::ios::binary);
        const int N = (int)q.size();
        file.write((const char *)&N, sizeof(int));
        file.write((const char *)q.data(), N * sizeof(double));
    }
    {
        std::ofstream file(path + "/lambda", std::ios::binary);
        const int N = (int)lambda.size();
        file.write((const char *)&N, sizeof(int));
        file.write((const char *)lambda.data(), N * sizeof(double));
    }
}

// src/io/io.cu:406-420
void writeInfoFile(const ParameterDB &DB, const Domain &D, const std::string &folder) {
    std::ofstream f(folder + "/run.info");
    f << std::setw(20) << "--nx" << "\t" << D.nx << '\n';
    f << std::setw(20) << "--ny" << "\t" << D.ny << '\n';
    f << std::setw(20) << "--startStep" << "\t" << DB.startStep << '\n';
    f << std::setw(20) << "--nt" << "\t" << DB.nt << '\n';
    f << std::setw(20) << "--nsave" << "\t" << DB.nsave << '\n';
    f << std::setw(20) << "--dt" << "\t" << DB.dt << '\n';
    f << std::setw(20) << "--vortlim" << "\t" << 15 << '\n';
    f << std::setw(20) << "--folder" << "\t" << DB.caseFolder << '\n';
    f << std::setw(20) << "--nu" << "\t" << DB.nu << '\n';
}

// src/bodies.cu:459-471
void writeBodies(const std::string &caseFolder, int timeStep, const std::vector<double> &bx,
                 const std::vector<double> &by) {
    std::ostringstream out;
    out << caseFolder << '/' << std::setfill('0') << std::setw(7) << timeStep;
    makeDirectory(out.str());
    std::ofstream file(out.str() + "/bodies");
    file << '#' << std::setw(19) << "x-coordinate" << std::setw(20) << "y-coordinate" << std::endl;
    for (size_t l = 0; l < bx.size(); l++) file << bx[l] << '\t' << by[l] << '\n';
}

const char *solverTypeName(SolverType t) {
    switch (t) {
        case NAVIERSTOKES: return "Navier-Stokes";
        case FADLUN: return "FADLUN";
        case OSC: return "OSC_CYLINDER";
        case FSI: return "FSI";
        case LUO: return "LUO";
    }
    return "?";
}

}  // namespace cuibm
