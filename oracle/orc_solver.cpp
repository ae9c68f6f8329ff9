// ORACLE — test infrastructure only (see orc.hpp).
// Initialisation and step orchestration, restating (paths under src/solvers/NavierStokes/):
//   NavierStokesSolver.cu:45-154   initialiseNoBody (fields, boundary arrays), initialiseLHS
//   NavierStokes/kernels/initialise.cu:26-63   initial velocity with perturbation
//   NavierStokesSolver.cu:163-188  plain step (NavierStokes/intermediateVelocity.cu:14-38)
//   src/bodies.cu:20-185, luo_base.cu:28-107   body set-up and re-centring
//   luo_base.cu:208-285, luoIBM.cu:91-102, luoIBM/*.cu       LUO step
//   fadlunModified.cu:97-124, FadlunModified/*.cu             FADLUN step
// plus the C entry points the tests call through ctypes.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include "orc.hpp"

namespace orc {

void Solver::init(const Params &Pin, const Grid &Gin, const std::vector<double> &bx,
                  const std::vector<double> &by) {
    P = Pin;
    G = Gin;
    nx = G.nx;
    ny = G.ny;
    numU = (nx - 1) * ny;
    numV = nx * (ny - 1);
    numUV = numU + numV;
    numP = nx * ny;
    for (auto *v : {&u, &uhat, &uold, &N, &Nold, &L, &bc1, &rhs1, &ustar, &image_point_u})
        v->assign(numUV, 0.0);
    for (auto *v : {&p, &p_old, &rhs2, &pstar, &dpdn}) v->assign(numP, 0.0);
    bc[XMINUS].assign(2 * ny - 1, 0.0);
    bc[XPLUS].assign(2 * ny - 1, 0.0);
    bc[YMINUS].assign(2 * nx - 1, 0.0);
    bc[YPLUS].assign(2 * nx - 1, 0.0);
    timeStep = 0;
    // initialise.cu:26-63
    const double xmin = G.x[0], xmax = G.x[nx - 1], ymin = G.y[0], ymax = G.y[ny - 1], pi = M_PI;
    for (int idx = 0; idx < numU; idx++) {
        const int i = idx % (nx - 1), j = idx / (nx - 1);
        u[idx] = P.uInitial + P.uPerturb * std::cos(0.5 * pi * (2 * G.xu[i] - xmax - xmin) / (xmax - xmin)) *
                                  std::sin(pi * (2 * G.y[j] - ymax - ymin) / (ymax - ymin));
    }
    for (int idx = 0; idx < numV; idx++) {
        const int i = idx % nx, j = idx / nx;
        u[idx + numU] = P.vInitial + P.vPerturb * std::cos(0.5 * pi * (2 * G.yv[j] - ymax - ymin) / (ymax - ymin)) *
                                         std::sin(pi * (2 * G.x[i] - xmax - xmin) / (xmax - xmin));
    }
    uhat = u;
    // boundary arrays, NavierStokesSolver.cu:108-128
    for (int i = 0; i < nx - 1; i++) {
        bc[YMINUS][i] = P.bc_u[YMINUS];
        bc[YPLUS][i] = P.bc_u[YPLUS];
        bc[YMINUS][i + nx - 1] = P.bc_v[YMINUS];
        bc[YPLUS][i + nx - 1] = P.bc_v[YPLUS];
    }
    bc[YMINUS][2 * nx - 2] = P.bc_v[YMINUS];
    bc[YPLUS][2 * nx - 2] = P.bc_v[YPLUS];
    for (int i = 0; i < ny - 1; i++) {
        bc[XMINUS][i] = P.bc_u[XMINUS];
        bc[XPLUS][i] = P.bc_u[XPLUS];
        bc[XMINUS][i + ny] = P.bc_v[XMINUS];
        bc[XPLUS][i + ny] = P.bc_v[XPLUS];
    }
    bc[XMINUS][ny - 1] = P.bc_u[XMINUS];
    bc[XPLUS][ny - 1] = P.bc_u[XPLUS];

    has_body = (P.solver_type == LUO || P.solver_type == FADLUN || P.solver_type == OSC) && !bx.empty();
    if (has_body) {
        for (auto *v : {&ghostTagsUV, &hybridTagsUV, &hybridTagsUV2}) v->assign(numUV, -1);
        for (auto *v : {&ghostTagsP, &hybridTagsP}) v->assign(numP, -1);
        for (auto *v : {&bi_x, &bi_y, &ip_x, &ip_y, &uv}) v->assign(numUV, 0.0);
        for (auto *v : {&dist_int, &dist_between}) v->assign(numUV, 1.0);
        for (auto *v : {&bip_x, &bip_y, &ipp_x, &ipp_y, &dist_u_body, &dist_v_body}) v->assign(numP, 0.0);
        for (int l = 0; l < 4; l++) {
            qcoef[l].assign(numUV, 0.0);
            qidx[l].assign(numUV, 0);
            qval[l].assign(numUV, 0.0);
            pqcoef[l].assign(numP, 0.0);
            pqidx[l].assign(numP, 0);
            pqval[l].assign(numP, 0.0);
        }
        for (auto *v : {&alpha1, &ns_rhs1, &interp_rhs1}) v->assign(numUV, 0.0);
        for (auto *v : {&alpha2, &ns_rhs2, &interp_rhs2}) v->assign(numP, 0.0);
        // bodies.cu:20-185: points, bounding box from the un-shifted points, centroid
        B.n = (int)bx.size();
        B.x = bx;
        B.y = by;
        for (auto *v : {&B.uB, &B.vB, &B.uB0, &B.vB0, &B.uBk, &B.vBk}) v->assign(B.n, 0.0);
        boundingBox();
        double mx = 0, my = 0;
        for (int i = 0; i < B.n; i++) { mx += B.x[i]; my += B.y[i]; }
        mx /= B.n;
        my /= B.n;
        // luo_base.cu:44-85: position and velocity of the prescribed oscillation at t = dt * timeStep,
        // the previous velocity at t - dt; the points keep their offsets from the centroid
        B.moving = P.xCoeff != 0 || P.uCoeff != 0 || P.viv != 0;
        const double t = P.dt * timeStep;
        const double xnew = P.xCoeff * std::sin(2 * M_PI * P.xfreq * t + P.xPhase);
        const double unew = P.uCoeff * std::cos(2 * M_PI * P.xfreq * t + P.uPhase);
        const double uoldd = P.uCoeff * std::cos(2 * M_PI * P.xfreq * (t - P.dt) + P.uPhase);
        B.midX = xnew;
        B.midY = my;
        B.dx.resize(B.n);
        B.dy.resize(B.n);
        for (int i = 0; i < B.n; i++) {
            B.dx[i] = B.x[i] - mx;
            B.dy[i] = B.y[i] - my;
            B.x[i] = B.midX + B.dx[i];
        }
        B.uB.assign(B.n, unew);
        B.uBk.assign(B.n, uoldd);
        tagPoints();
        if (P.solver_type == LUO) set_inside(u);  // luoIBM.cu:42
    }
    // luo_base::initialiseLHS is empty: the embedded method assembles both systems and
    // builds both preconditioners inside every step (luo_iter/intermediateVelocity.cu:64-71)
    if (P.solver_type == OSC && has_body) return;
    buildLHS1();
    buildLHS2();
    PC1.build(LHS1, P.precond1);
    PC2.build(LHS2, P.precond2);
}

void Solver::solve1() {
    SolveStat s = bicgstab(LHS1, uhat.data(), rhs1.data(), PC1, P.tol1, P.maxit1);
    iter1 = s.iters;
    res1 = s.resid;
}

void Solver::solve2() {
    SolveStat s = bicgstab(LHS2, p.data(), rhs2.data(), PC2, P.tol2, P.maxit2);
    iter2 = s.iters;
    res2 = s.resid;
}

// NavierStokesSolver.cu:163-188
void Solver::step_ns() {
    Nold = N;
    generateN();
    generateL();
    generateBC1();
    generateRHS();
    solve1();
    rhs2_nobody();
    solve2();
    project_nobody();
    timeStep++;
}

// luo_base.cu:236-248
void Solver::luo_pre_step() {
    p_old = p;
    uold = u;
    Nold = N;
    B.uB0 = B.uB; B.vB0 = B.vB; B.uBk = B.uB; B.vBk = B.vB;
    B.midY0 = B.midY;
    B.centerV0 = B.centerV;
}

// one pass of luo_base::stepTime for the external method (luoIBM.cu:91-102)
void Solver::luo_pass() {
    // luoIBM/intermediateVelocity.cu:18-57
    u = uold;
    robin();
    gn_velocity(u, true);
    set_inside(u);
    hn_velocity(u);
    generateN();
    generateL();
    generateBC1();
    generateRHS();
    solve1();
    weight_uhat();
    // luoIBM/intermediatePressure.cu:15-29, 44-77
    gn_velocity(uhat, true);
    rhs2_nobody();
    solve2();
    std::fill(pstar.begin(), pstar.end(), 0.0);
    hn_pressure();
    weight_p();
    gn_pressure(true);
    // luo_base/projectVelocity.cu:11-53
    project_nobody();
    gn_velocity(u, true);
    set_inside(u);
    update_body();
}

// luo_base.cu:154-174 and 137-147: force, the body's motion, bounding box and re-tag
// (for a static body the re-tag reproduces the tags and is skipped)
void Solver::update_body() {
    force();
    if (!B.moving) return;
    if (P.viv == 0) move_body();
    else viv_move(P.viv == 2);
    boundingBox();
    tagPoints();
}

// luo_base.cu:208-233 (the strongly coupled loop stops after 200 passes here; the
// reference loops until the position settles)
void Solver::step_luo() {
    luo_pre_step();
    if (P.viv == 2) {
        sc_tol = 0.0001;
        sc_passes = 0;
        while (std::fabs(sc_tol) > 0.001 || sc_passes < 2) {
            if (sc_passes >= 200) throw std::runtime_error("strongly coupled VIV passes did not settle");
            P.solver_type == OSC ? embedded_pass() : luo_pass();
            sc_passes++;
        }
    } else {
        sc_passes = 1;
        P.solver_type == OSC ? embedded_pass() : luo_pass();
    }
    timeStep++;
}

// luo_base/move.cu:11-36 (set_movement) at t = dt * timeStep, before the step counter advances
void Solver::move_body() {
    const double t = P.dt * timeStep;
    const double xnew = P.xCoeff * std::sin(2 * M_PI * P.xfreq * t + P.xPhase);
    const double unew = P.uCoeff * std::cos(2 * M_PI * P.xfreq * t + P.uPhase);
    B.midX = xnew;
    for (int i = 0; i < B.n; i++) {
        B.uB[i] = unew;
        B.x[i] = B.midX + B.dx[i];
    }
}

// luo_base/move.cu:38-93 (viv_movement_LC / viv_movement_SC): transverse spring-mass
// response to the lift, reduced mass 2; the strongly coupled form relaxes the position
// and reports its relative change in sc_tol
void Solver::viv_move(bool strong) {
    double vnew, ynew;
    sc_tol = B.midY;
    orc_viv_update(B.midY0, B.centerV0, B.forceY * 2.0, P.ured, P.dt, strong ? 1 : 0, B.midY, &vnew, &ynew);
    B.centerV = vnew;
    B.midY = ynew;
    for (int i = 0; i < B.n; i++) {
        B.vB[i] = vnew;
        B.y[i] = B.midY + B.dy[i];
    }
    if (strong) sc_tol = (sc_tol - ynew) / sc_tol;
}

// luo_base.cu:208-285 with fadlunModified.cu:97-124
void Solver::step_fadlun() {
    p_old = p;
    uold = u;
    Nold = N;
    robin();
    generateN();
    generateL();
    generateBC1();
    generateRHS();
    fadlun_rhs1();
    solve1();
    fadlun_rhs2();
    p_old = p;
    solve2();
    fadlun_project();
    force();
    timeStep++;
}

void Solver::step() {
    if (P.solver_type == LUO && has_body) step_luo();
    else if (P.solver_type == OSC && has_body) step_luo();
    else if (P.solver_type == FADLUN && has_body) step_fadlun();
    else step_ns();
}

std::vector<double> *Solver::field(const std::string &n) {
    if (n == "u") return &u;
    if (n == "uhat") return &uhat;
    if (n == "uold") return &uold;
    if (n == "p" || n == "pressure") return &p;
    if (n == "N") return &N;
    if (n == "Nold") return &Nold;
    if (n == "L") return &L;
    if (n == "bc1") return &bc1;
    if (n == "rhs1") return &rhs1;
    if (n == "rhs2") return &rhs2;
    if (n == "ustar") return &ustar;
    if (n == "pstar") return &pstar;
    if (n == "xm") return &bc[XMINUS];
    if (n == "xp") return &bc[XPLUS];
    if (n == "ym") return &bc[YMINUS];
    if (n == "yp") return &bc[YPLUS];
    if (n == "body_intercept_x") return &bi_x;
    if (n == "body_intercept_y") return &bi_y;
    if (n == "image_point_x") return &ip_x;
    if (n == "image_point_y") return &ip_y;
    if (n == "body_intercept_p_x") return &bip_x;
    if (n == "body_intercept_p_y") return &bip_y;
    if (n == "image_point_p_x") return &ipp_x;
    if (n == "image_point_p_y") return &ipp_y;
    if (n == "distance_from_intersection_to_node") return &dist_int;
    if (n == "distance_between_nodes_at_IB") return &dist_between;
    if (n == "distance_from_u_to_body") return &dist_u_body;
    if (n == "distance_from_v_to_body") return &dist_v_body;
    if (n == "uv") return &uv;
    if (n == "x") return &G.x;
    if (n == "dx") return &G.dx;
    if (n == "y") return &G.y;
    if (n == "dy") return &G.dy;
    if (n == "xu") return &G.xu;
    if (n == "yv") return &G.yv;
    if (n == "body_x") return &B.x;
    if (n == "body_y") return &B.y;
    if (n == "LHS1_val") return &LHS1.v;
    if (n == "LHS2_val") return &LHS2.v;
    // interpolation records and the embedded method's per-row data
    if (n == "alpha1") return &alpha1;
    if (n == "alpha2") return &alpha2;
    if (n == "ns_rhs1") return &ns_rhs1;
    if (n == "interp_rhs1") return &interp_rhs1;
    if (n == "ns_rhs2") return &ns_rhs2;
    if (n == "interp_rhs2") return &interp_rhs2;
    if (n == "dpdn") return &dpdn;
    for (int l = 0; l < 4; l++) {
        const std::string k = std::to_string(l);
        if (n == "qcoef" + k) return &qcoef[l];
        if (n == "qval" + k) return &qval[l];
        if (n == "pqcoef" + k) return &pqcoef[l];
        if (n == "pqval" + k) return &pqval[l];
    }
    return nullptr;
}

std::vector<int> *Solver::ifield(const std::string &n) {
    for (int l = 0; l < 4; l++) {
        const std::string k = std::to_string(l);
        if (n == "qidx" + k) return &qidx[l];
        if (n == "pqidx" + k) return &pqidx[l];
    }
    if (n == "ghostTagsUV") return &ghostTagsUV;
    if (n == "hybridTagsUV") return &hybridTagsUV;
    if (n == "hybridTagsUV2") return &hybridTagsUV2;
    if (n == "ghostTagsP") return &ghostTagsP;
    if (n == "hybridTagsP") return &hybridTagsP;
    if (n == "LHS1_rowptr") return &LHS1.rp;
    if (n == "LHS1_col") return &LHS1.ci;
    if (n == "LHS2_rowptr") return &LHS2.rp;
    if (n == "LHS2_col") return &LHS2.ci;
    return nullptr;
}

}  // namespace orc

extern "C" {

// flat parameter block filled by the test harness (ctypes)
struct orc_params {
    int solver_type;
    int nx, ny;
    const double *x, *dx, *y, *dy;
    double mid_h;
    double nu, dt;
    int nt;
    double uInitial, vInitial, uPerturb, vPerturb;
    double bc_u[4], bc_v[4];
    int precond1, precond2;
    double tol1, tol2;
    int maxit1, maxit2;
    double scaleCV;
    int nbody;
    const double *bx, *by;
    double xCoeff, uCoeff, xPhase, uPhase, xfreq;
    int viv, ured;
};

void *orc_create(const orc_params *q) {
    orc::Params P;
    P.solver_type = q->solver_type;
    P.nu = q->nu;
    P.dt = q->dt;
    P.nt = q->nt;
    P.uInitial = q->uInitial;
    P.vInitial = q->vInitial;
    P.uPerturb = q->uPerturb;
    P.vPerturb = q->vPerturb;
    for (int i = 0; i < 4; i++) { P.bc_u[i] = q->bc_u[i]; P.bc_v[i] = q->bc_v[i]; }
    P.precond1 = q->precond1;
    P.precond2 = q->precond2;
    P.tol1 = q->tol1;
    P.tol2 = q->tol2;
    P.maxit1 = q->maxit1;
    P.maxit2 = q->maxit2;
    P.scaleCV = q->scaleCV;
    P.xCoeff = q->xCoeff;
    P.uCoeff = q->uCoeff;
    P.xPhase = q->xPhase;
    P.uPhase = q->uPhase;
    P.xfreq = q->xfreq;
    P.viv = q->viv;
    P.ured = q->ured;
    orc::Grid G;
    G.nx = q->nx;
    G.ny = q->ny;
    G.mid_h = q->mid_h;
    G.x.assign(q->x, q->x + q->nx);
    G.dx.assign(q->dx, q->dx + q->nx);
    G.y.assign(q->y, q->y + q->ny);
    G.dy.assign(q->dy, q->dy + q->ny);
    G.xu.resize(G.nx - 1);
    G.yv.resize(G.ny - 1);
    for (int i = 0; i < G.nx - 1; i++) G.xu[i] = G.x[i] + G.dx[i] / 2;  // parseDomainFile.cu:370-377
    for (int j = 0; j < G.ny - 1; j++) G.yv[j] = G.y[j] + G.dy[j] / 2;
    std::vector<double> bx, by;
    if (q->nbody > 0) {
        bx.assign(q->bx, q->bx + q->nbody);
        by.assign(q->by, q->by + q->nbody);
    }
    auto *S = new orc::Solver();
    S->init(P, G, bx, by);
    return S;
}

void orc_destroy(void *h) { delete (orc::Solver *)h; }

int orc_step(void *h, int nsteps) {
    auto *S = (orc::Solver *)h;
    for (int k = 0; k < nsteps; k++) S->step();
    return 0;
}

// phase entry points (same names as the product's C ABI phases)
int orc_phase(void *h, const char *name) {
    auto *S = (orc::Solver *)h;
    const std::string n(name);
    if (n == "generateN") S->generateN();
    else if (n == "generateL") S->generateL();
    else if (n == "generateBC1") S->generateBC1();
    else if (n == "generateRHS") S->generateRHS();
    else if (n == "rhs2") S->rhs2_nobody();
    else if (n == "project") S->project_nobody();
    else if (n == "solve1") S->solve1();
    else if (n == "solve2") S->solve2();
    else if (n == "robin") S->robin();
    else if (n == "gn_u") S->gn_velocity(S->u, true);
    else if (n == "gn_uhat") S->gn_velocity(S->uhat, true);
    else if (n == "set_inside") S->set_inside(S->u);
    else if (n == "hn_u") S->hn_velocity(S->u);
    else if (n == "weight_uhat") S->weight_uhat();
    else if (n == "hn_p") S->hn_pressure();
    else if (n == "weight_p") S->weight_p();
    else if (n == "gn_p") S->gn_pressure(true);
    else if (n == "fadlun_rhs1") S->fadlun_rhs1();
    else if (n == "fadlun_rhs2") S->fadlun_rhs2();
    else if (n == "fadlun_project") S->fadlun_project();
    else if (n == "force") S->force();
    else if (n == "tag") S->tagPoints();
    else return -1;
    return 0;
}

// per-point body velocities of a LUO case after its set-up (orc::Solver::set_body_velocity)
int orc_set_body_velocity(void *h, const double *ub, const double *vb, int n, double omega) {
    auto *S = (orc::Solver *)h;
    if (n != S->B.n || S->P.solver_type != orc::LUO) return -1;
    S->set_body_velocity(ub, vb, omega);
    return 0;
}

long orc_get(void *h, const char *name, void *out, long n) {
    auto *S = (orc::Solver *)h;
    if (auto *f = S->field(name)) {
        if (out) std::memcpy(out, f->data(), sizeof(double) * std::min<long>(n, (long)f->size()));
        return (long)f->size();
    }
    if (auto *f = S->ifield(name)) {
        if (out) std::memcpy(out, f->data(), sizeof(int) * std::min<long>(n, (long)f->size()));
        return (long)f->size();
    }
    return -1;
}

long orc_set(void *h, const char *name, const void *in, long n) {
    auto *S = (orc::Solver *)h;
    if (auto *f = S->field(name)) {
        std::memcpy(f->data(), in, sizeof(double) * std::min<long>(n, (long)f->size()));
        return (long)f->size();
    }
    if (auto *f = S->ifield(name)) {
        std::memcpy(f->data(), in, sizeof(int) * std::min<long>(n, (long)f->size()));
        return (long)f->size();
    }
    return -1;
}

// scalars: iterations, residuals, forces, box
void orc_info(void *h, double *out) {
    auto *S = (orc::Solver *)h;
    out[0] = S->timeStep;
    out[1] = S->iter1;
    out[2] = S->iter2;
    out[3] = S->res1;
    out[4] = S->res2;
    out[5] = S->B.forceX;
    out[6] = S->B.forceY;
    out[7] = S->fxx;
    out[8] = S->fxy;
    out[9] = S->fxu;
    out[10] = S->B.startI;
    out[11] = S->B.startJ;
    out[12] = S->B.numCellsX;
    out[13] = S->B.numCellsY;
    out[14] = S->B.midX;
    out[15] = S->B.midY;
    out[16] = S->PC2.type == orc::P_SMOOTHED_AGGREGATION ? (double)S->PC2.amg.lv.size() : 0;
    out[17] = S->sc_passes;
    out[18] = S->B.centerV;
}

// standalone linear-algebra entry points used by the solver tests
// coarsest size of the AMG hierarchies set up from now on (<= 0 restores 2500)
void orc_set_amg_coarse_max(int n) { orc::amg_default_coarse_max = n > 0 ? n : 2500; }
// hierarchies built by orc_bicgstab and the inspection entry points below restrict the
// finest level with P^T D^-1 while this is set (the embedded method's hierarchy)
static bool g_scale0 = false;
void orc_amg_scale0(int on) { g_scale0 = on != 0; }

void *orc_csr(int n, const int *rp, const int *ci, const double *v) {
    auto *A = new orc::Csr();
    A->n = n;
    A->rp.assign(rp, rp + n + 1);
    A->ci.assign(ci, ci + rp[n]);
    A->v.assign(v, v + rp[n]);
    return A;
}
void orc_csr_free(void *A) { delete (orc::Csr *)A; }

int orc_bicgstab(void *A, double *x, const double *b, int precond, double rtol, int maxit, double *resid) {
    auto *M = (orc::Csr *)A;
    orc::Precond P;
    P.build(*M, precond, g_scale0);
    orc::SolveStat s = orc::bicgstab(*M, x, b, P, rtol, maxit);
    if (resid) *resid = s.resid;
    return s.converged ? s.iters : -1 - s.iters;
}

// AMG hierarchy inspection: returns the number of levels; sizes[l] = n, nnzs[l] = nnz(A_l)
int orc_amg_levels(void *A, int *sizes, int *nnzs, int cap) {
    orc::Amg G;
    G.scale0 = g_scale0;
    G.setup(*(orc::Csr *)A);
    const int L = (int)G.lv.size();
    for (int l = 0; l < L && l < cap; l++) {
        sizes[l] = G.lv[l].A.n;
        nnzs[l] = G.lv[l].A.nnz();
    }
    return L;
}

// one V-cycle x = M^-1 b (zero start)
int orc_amg_vcycle(void *A, const double *b, double *x) {
    orc::Amg G;
    G.scale0 = g_scale0;
    G.setup(*(orc::Csr *)A);
    G.apply(b, x);
    return (int)G.lv.size();
}

// Galerkin product of the first level: writes A_1 = R A P as dense (n1 x n1); returns n1
int orc_amg_coarse_dense(void *A, int level, double *out, int cap) {
    orc::Amg G;
    G.scale0 = g_scale0;
    G.setup(*(orc::Csr *)A);
    if (level < 0) level = (int)G.lv.size() - 1;
    if (level >= (int)G.lv.size()) return -1;
    const orc::Csr &Ac = G.lv[level].A;
    if ((long)Ac.n * Ac.n > cap) return Ac.n;
    std::memset(out, 0, sizeof(double) * (size_t)Ac.n * Ac.n);
    for (int i = 0; i < Ac.n; i++)
        for (int k = Ac.rp[i]; k < Ac.rp[i + 1]; k++) out[(size_t)i * Ac.n + Ac.ci[k]] += Ac.v[k];
    return Ac.n;
}
// level `level` matrix as CSR (which = 0: A_l, 1: P_l, 2: R_l): returns the row count and
// writes the entry count to *nnz; the arrays are filled only when they hold cap entries
int orc_amg_level_csr(void *A, int level, int which, int *rp, int *ci, double *v, int cap, int *nnz) {
    orc::Amg G;
    G.scale0 = g_scale0;
    G.setup(*(orc::Csr *)A);
    if (level < 0 || level >= (int)G.lv.size()) return -1;
    const orc::Csr &M = which == 0 ? G.lv[level].A : which == 1 ? G.lv[level].P : G.lv[level].R;
    *nnz = M.nnz();
    if (M.nnz() > cap || M.n + 1 > cap) return M.n;
    std::copy(M.rp.begin(), M.rp.end(), rp);
    std::copy(M.ci.begin(), M.ci.end(), ci);
    std::copy(M.v.begin(), M.v.end(), v);
    return M.n;
}
// dense P of the first level (n0 x n1)
int orc_amg_prolongator_dense(void *A, double *out, int cap) {
    orc::Amg G;
    G.scale0 = g_scale0;
    G.setup(*(orc::Csr *)A);
    if (G.lv.size() < 2) return -1;
    const orc::Csr &P = G.lv[0].P;
    const int n1 = G.lv[1].A.n;
    if ((long)P.n * n1 > cap) return -2;
    std::memset(out, 0, sizeof(double) * (size_t)P.n * n1);
    for (int i = 0; i < P.n; i++)
        for (int k = P.rp[i]; k < P.rp[i + 1]; k++) out[(size_t)i * n1 + P.ci[k]] += P.v[k];
    return n1;
}
}

// luo_base/move.cu:38-93 (viv_movement_LC / viv_movement_SC): one step of the transverse
// spring-mass response to the lift coefficient Cy, reduced mass 2 and integer reduced
// velocity Ured: the velocity from the spring force at the new position, the position by
// the trapezoidal rule; the strongly coupled form relaxes the position by 0.9 towards
// ycur, the position of the previous pass
extern "C" void orc_viv_update(double yold, double vold, double Cy, int Ured, double dt, int strong, double ycur,
                               double *vnew, double *ynew) {
    const double Mred = 2.0;
    const double a = dt * M_PI * M_PI * 4 / (Ured * Ured), b = dt * dt * 2 * M_PI * M_PI / (Ured * Ured);
    const double v = (vold - a * (yold + dt / 2 * vold) + dt * Cy / 2 / Mred) / (1 + b);
    *vnew = v;
    if (strong) {
        const double relax = 0.9;
        *ynew = relax * (yold + dt / 2 * (v + vold)) + (1 - relax) * ycur;
    } else {
        *ynew = yold + dt / 2 * (v + vold);
    }
}
