// Lockstep batches of independent Luo-external cases: set-up from the initialised case
// solvers, the batched time step and its BiCGSTAB driver (no counterpart in the
// reference). The step is the luoIBM step (luoIBM.cu, luo_base.cu:208-264) phase by phase:
//   outflow values, ghost and solid velocity, hybrid interpolation, rhs1,
//   velocity solve, hybrid blending of uhat, ghost values of uhat, rhs2,
//   Poisson solve, hybrid blending and ghost values of p, projection, ghost and solid
//   velocity, control-volume force,
// each phase one launch (or one launch per BiCGSTAB kernel) over every case.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <string>

#include "batch.hpp"
#include "solver.hpp"

namespace cuibm {

void BMap::build(const std::vector<long> &count, int per_block, hipStream_t s) {
    std::vector<int> cs, loc, first{0};
    for (size_t c = 0; c < count.size(); c++) {
        const long nb = (count[c] + per_block - 1) / per_block;
        for (long b = 0; b < nb; b++) {
            cs.push_back((int)c);
            loc.push_back((int)b);
        }
        first.push_back((int)cs.size());
    }
    nblk_ = (int)cs.size();
    if (cs.empty()) {  // keep the pointers valid
        cs.push_back(0);
        loc.push_back(0);
    }
    cs_.upload(cs, s);
    loc_.upload(loc, s);
    first_.upload(first, s);
}

namespace {

// the stencils of all cases, indices shifted by each case's segment offset
void cat_stencils(const std::vector<const Stencils *> &src, const std::vector<long> &off, long limit, Stencils &dst,
                  hipStream_t s) {
    std::vector<int> node, idx;
    std::vector<double> coef, nrm, extra;
    int n = 0;
    long per_nrm = -1, per_extra = -1;
    for (size_t c = 0; c < src.size(); c++) {
        const Stencils &S = *src[c];
        if (S.n == 0) continue;
        std::vector<int> nd = S.node.to_host(s), ix = S.idx.to_host(s);
        std::vector<double> cf = S.coef.to_host(s);
        nd.resize(S.n);
        ix.resize(4 * (size_t)S.n);
        cf.resize(4 * (size_t)S.n);
        for (int &v : nd) v += (int)off[c];
        for (int &v : ix)
            if (v >= 0) v += (int)off[c];
        node.insert(node.end(), nd.begin(), nd.end());
        idx.insert(idx.end(), ix.begin(), ix.end());
        coef.insert(coef.end(), cf.begin(), cf.end());
        const long pn = S.nrm_per, pe = S.extra_per;
        if ((per_nrm >= 0 && pn != per_nrm) || (per_extra >= 0 && pe != per_extra))
            throw Error("lockstep batch: interpolation stencils of different layouts", ERR_ARGUMENT);
        per_nrm = pn;
        per_extra = pe;
        if (pn) {
            std::vector<double> v = S.nrm.to_host(s);
            nrm.insert(nrm.end(), v.begin(), v.begin() + pn * S.n);
        }
        if (pe) {
            std::vector<double> v = S.extra.to_host(s);
            extra.insert(extra.end(), v.begin(), v.begin() + pe * S.n);
        }
        n += S.n;
    }
    dst.n = n;
    if (!n) return;
    for (size_t t = 0; t < node.size(); t++)
        if (node[t] < 0 || node[t] >= limit) throw Error("lockstep batch: a stencil node outside the fields", ERR_GENERIC);
    for (int v : idx)
        if (v >= limit) throw Error("lockstep batch: a stencil corner outside the fields", ERR_GENERIC);
    dst.node.upload(node, s);
    dst.idx.upload(idx, s);
    dst.coef.upload(coef, s);
    if (!nrm.empty()) dst.nrm.upload(nrm, s);
    if (!extra.empty()) dst.extra.upload(extra, s);
    dst.tmp.alloc(n);
}

// device-to-device copy of one case segment
void seg_copy(double *dst, const double *src, long n, hipStream_t s) {
    if (n) CUIBM_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, s));
}

}  // namespace

struct LockstepBatch::Impl {
    hipStream_t s = nullptr;
    int nc = 0;
    std::vector<int> order;  // batch position -> input index
    std::vector<NavierStokesSolver *> solver;  // batch order
    std::vector<long> offUV, offP;
    std::vector<int> offXP;
    long nUV = 0, nP = 0, nXP = 0;
    DArray<double> u, uhat, uold, p, N, Nold, rhs1, rhs2, ustar, pstar, xp, dinv1, defl;
    DArray<double> sc;  // 2 scalar blocks per case (velocity, Poisson)
    std::vector<double> host_sc;
    DArray<BCase> tab;
    std::vector<BCase> tab_host;
    BMap mUV, mP, mXP;
    Stencils gnV, hnV, hnP, gnP;
    DArray<int> solid;
    int nsolid = 0;
    std::unique_ptr<BatchAmg> amg;
    // Krylov workspace per operator
    struct Work {
        DArray<double> r, rs, p, Mp, AMp, sv, Ms, AMs, part;
    } wV, wP;
    DArray<double> fpart, fout;
    double tol1 = 0, tol2 = 0;
    int maxit1 = 0, maxit2 = 0;
    // eight BiCGSTAB iterations of each operator as one captured graph
    static constexpr int GRAPH_ITS = 8;
    bool graphs = true;
    hipGraphExec_t block[2] = {nullptr, nullptr};
    // reliable updates of the Poisson solve (the cases' restart_reduction()): a case whose
    // residual has fallen by this factor since its last (re)start restarts from b - A x
    double reliable2 = 0.0;  // the factor squared (compared with r.r)
    DArray<int> mask;
    std::vector<int> host_mask;

    ~Impl() {
        for (hipGraphExec_t &g : block)
            if (g) (void)hipGraphExecDestroy(g);
    }

    void fetch() {
        CUIBM_HIP(hipMemcpyAsync(host_sc.data(), sc.p(), host_sc.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        CUIBM_HIP(hipStreamSynchronize(s));
    }
    const double *scv(int c, int op) const { return host_sc.data() + (2 * c + op) * SC_COUNT; }

    // BiCGSTAB of every case at once (cusp::monitor semantics per case); returns the
    // largest iteration count
    template <class M>
    int solve(int op, double *x, const double *b, M precond, double rtol, int maxit, const char *what) {
        const BMap &m = op == BOP_VEL ? mUV : mP;
        Work &w = op == BOP_VEL ? wV : wP;
        const BMapDev md = m.dev();
        launch_b_init(op, tab.p(), md, x, b, w.r.p(), w.p.p(), w.rs.p(), w.part.p(), s);
        launch_b_finish(op, tab.p(), nc, m.first(), md.nblk, w.part.p(), FIN_INIT, rtol, maxit, s);
        // one iteration: every launch is the same in every iteration (the stop codes and the
        // half-step state live in the per-case scalar blocks), so a block of them can be
        // captured once and replayed
        auto iteration = [&]() {
            precond(w.p.p(), w.Mp.p());
            launch_b_spmv_dot(op, tab.p(), md, w.Mp.p(), w.AMp.p(), w.rs.p(), false, w.part.p(), s);
            launch_b_finish(op, tab.p(), nc, m.first(), md.nblk, w.part.p(), FIN_ALPHA, rtol, maxit, s);
            launch_b_bicg_s(op, tab.p(), md, w.r.p(), w.AMp.p(), w.sv.p(), w.part.p(), s);
            launch_b_finish(op, tab.p(), nc, m.first(), md.nblk, w.part.p(), FIN_SS, rtol, maxit, s);
            precond(w.sv.p(), w.Ms.p());
            launch_b_spmv_dot(op, tab.p(), md, w.Ms.p(), w.AMs.p(), w.sv.p(), true, w.part.p(), s);
            launch_b_finish(op, tab.p(), nc, m.first(), md.nblk, w.part.p(), FIN_OMEGA, rtol, maxit, s);
            launch_b_bicg_xr(op, tab.p(), md, x, w.Mp.p(), w.Ms.p(), w.r.p(), w.sv.p(), w.AMs.p(), w.rs.p(),
                             w.part.p(), s);
            launch_b_finish(op, tab.p(), nc, m.first(), md.nblk, w.part.p(), FIN_BETA, rtol, maxit, s);
            launch_b_bicg_p(op, tab.p(), md, w.p.p(), w.r.p(), w.AMp.p(), s);
        };
        int it = 0;
        const double rel2 = op == BOP_POIS ? reliable2 : 0.0;
        std::vector<double> rr_ref;
        for (;;) {
            fetch();
            if (rr_ref.empty())
                for (int c = 0; c < nc; c++) rr_ref.push_back(scv(c, op)[SC_RR]);
            bool running = false;
            for (int c = 0; c < nc; c++)
                if (scv(c, op)[SC_STOP] == 0.0) running = true;
            if (!running || it >= maxit) break;
            if (rel2 > 0) {  // reliable updates of the running cases whose residual fell far enough
                bool any = false;
                for (int c = 0; c < nc; c++) {
                    const double *v = scv(c, op);
                    host_mask[c] = v[SC_STOP] == 0.0 && v[SC_RR] < rel2 * rr_ref[c];
                    any = any || host_mask[c];
                }
                if (any) {
                    mask.upload(host_mask, s);
                    launch_b_init(op, tab.p(), md, x, b, w.r.p(), w.p.p(), w.rs.p(), w.part.p(), s, mask.p());
                    launch_b_finish(op, tab.p(), nc, m.first(), md.nblk, w.part.p(), FIN_RESTART, rtol, maxit, s,
                                    mask.p());
                    fetch();
                    for (int c = 0; c < nc; c++)
                        if (host_mask[c]) rr_ref[c] = scv(c, op)[SC_RR];
                    running = false;
                    for (int c = 0; c < nc; c++)
                        if (scv(c, op)[SC_STOP] == 0.0) running = true;
                    if (!running) break;
                }
            }
            const int batch = std::min(GRAPH_ITS, maxit - it);
            if (graphs && batch == GRAPH_ITS) {
                // the block reads and writes the same arrays every time (x and the
                // operator's workspace): capture it at its first use
                hipGraphExec_t &g = block[op == BOP_VEL ? 0 : 1];
                if (!g) {
                    hipGraph_t graph = nullptr;
                    CUIBM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    try {
                        for (int j = 0; j < GRAPH_ITS; j++) iteration();
                    } catch (...) {  // leave the stream out of capture mode
                        hipGraph_t partial = nullptr;
                        (void)hipStreamEndCapture(s, &partial);
                        if (partial) (void)hipGraphDestroy(partial);
                        throw;
                    }
                    CUIBM_HIP(hipStreamEndCapture(s, &graph));
                    const hipError_t e = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(graph);
                    CUIBM_HIP(e);
                }
                CUIBM_HIP(hipGraphLaunch(g, s));
                it += GRAPH_ITS;
            } else {
                for (int j = 0; j < batch; j++, it++) iteration();
            }
        }
        int most = 0;
        for (int c = 0; c < nc; c++) {
            const double *v = scv(c, op);
            const int code = (int)v[SC_STOP];
            if (code != STOP_CONVERGED && code != STOP_HALF)
                throw Error(std::string("lockstep batch: ") + what + " solve of case " + std::to_string(order[c]) +
                                (code == STOP_MAXIT || code == 0 ? " reached its iteration limit"
                                                                 : " broke down"),
                            ERR_NOT_CONVERGED);
            most = std::max(most, (int)v[SC_ITS]);
        }
        return most;
    }
};

LockstepBatch::LockstepBatch(const std::vector<NavierStokesSolver *> &cases, hipStream_t s, bool graphs)
    : im_(new Impl) {
    Impl &B = *im_;
    B.s = s;
    B.graphs = graphs;
    B.nc = (int)cases.size();
    if (B.nc == 0) throw Error("lockstep batch without cases", ERR_ARGUMENT);
    std::vector<luoIBM *> luo(B.nc);
    std::vector<AmgPrecond *> amg(B.nc);
    for (int c = 0; c < B.nc; c++) {
        luo[c] = dynamic_cast<luoIBM *>(cases[c]);
        if (!luo[c]) throw Error("lockstep batch: LUO cases only", ERR_UNSUPPORTED);
        luoIBM &L = *luo[c];
        if (!L.tagged || L.moving || L.perPointUB || L.db->VIV != 0)
            throw Error("lockstep batch: initialised static bodies only", ERR_UNSUPPORTED);
        if ((!L.bUB.empty() && L.bUB[0] != 0.0) || (!L.bVB.empty() && L.bVB[0] != 0.0))
            throw Error("lockstep batch: bodies at rest only", ERR_UNSUPPORTED);
        if (!dynamic_cast<VelocityStencil *>(L.op1.get()) || !dynamic_cast<PoissonStencil *>(L.op2.get()))
            throw Error("lockstep batch: stencil operators only", ERR_UNSUPPORTED);
        if (!dynamic_cast<JacobiPrecond *>(L.PC1.get()))
            throw Error("lockstep batch: DIAGONAL velocity preconditioner only", ERR_UNSUPPORTED);
        amg[c] = dynamic_cast<AmgPrecond *>(L.PC2.get());
        if (!amg[c] || !amg[c]->has_stencil_level())
            throw Error("lockstep batch: SMOOTHED_AGGREGATION Poisson preconditioner on a stencil finest level only",
                        ERR_UNSUPPORTED);
    }
    const ParameterDB &d0 = *luo[0]->db;
    B.tol1 = d0.velocitySolve.tolerance;
    B.tol2 = d0.PoissonSolve.tolerance;
    B.maxit1 = d0.velocitySolve.maxIterations;
    B.maxit2 = d0.PoissonSolve.maxIterations;
    for (int c = 1; c < B.nc; c++) {
        const ParameterDB &d = *luo[c]->db;
        if (d.velocitySolve.tolerance != B.tol1 || d.PoissonSolve.tolerance != B.tol2 ||
            d.velocitySolve.maxIterations != B.maxit1 || d.PoissonSolve.maxIterations != B.maxit2)
            throw Error("lockstep batch: the cases' solver tolerances and limits must agree", ERR_ARGUMENT);
    }
    // deepest hierarchy first (the batched V-cycle's level prefixes), then largest first
    B.order.resize(B.nc);
    std::iota(B.order.begin(), B.order.end(), 0);
    std::stable_sort(B.order.begin(), B.order.end(), [&](int a, int b) {
        const int la = amg[a]->info().levels, lb = amg[b]->info().levels;
        if (la != lb) return la > lb;
        return luo[a]->numP > luo[b]->numP;
    });
    std::vector<luoIBM *> L(B.nc);
    std::vector<AmgPrecond *> A(B.nc);
    for (int k = 0; k < B.nc; k++) {
        L[k] = luo[B.order[k]];
        A[k] = amg[B.order[k]];
        B.solver.push_back(L[k]);
    }
    // segments
    for (int c = 0; c < B.nc; c++) {
        B.offUV.push_back(B.nUV);
        B.offP.push_back(B.nP);
        B.offXP.push_back((int)B.nXP);
        B.nUV += batch_pad(L[c]->numUV);
        B.nP += batch_pad(L[c]->numP);
        B.nXP += batch_pad(2 * L[c]->ny - 1);
        cells_ += (long long)L[c]->numP;
    }
    for (auto *a : {&B.u, &B.uhat, &B.uold, &B.N, &B.Nold, &B.rhs1, &B.ustar, &B.dinv1}) {
        a->alloc(B.nUV);
        launch_fill(a->p(), 0.0, B.nUV, s);
    }
    for (auto *a : {&B.p, &B.rhs2, &B.pstar}) {
        a->alloc(B.nP);
        launch_fill(a->p(), 0.0, B.nP, s);
    }
    B.xp.alloc(B.nXP);
    launch_fill(B.xp.p(), 0.0, B.nXP, s);
    for (int c = 0; c < B.nc; c++) {
        luoIBM &S = *L[c];
        CUIBM_HIP(hipStreamSynchronize(S.stream_));
        seg_copy(B.u.p() + B.offUV[c], S.u.p(), S.numUV, s);
        seg_copy(B.uhat.p() + B.offUV[c], S.uhat.p(), S.numUV, s);
        seg_copy(B.uold.p() + B.offUV[c], S.uold.p(), S.numUV, s);
        seg_copy(B.N.p() + B.offUV[c], S.N.p(), S.numUV, s);
        seg_copy(B.Nold.p() + B.offUV[c], S.Nold.p(), S.numUV, s);
        seg_copy(B.ustar.p() + B.offUV[c], S.ustar.p(), S.numUV, s);
        seg_copy(B.p.p() + B.offP[c], S.pressure.p(), S.numP, s);
        seg_copy(B.pstar.p() + B.offP[c], S.pstar.p(), S.numP, s);
        seg_copy(B.xp.p() + B.offXP[c], S.bc[XPLUS].p(), 2 * S.ny - 1, s);
        launch_diag_inv(S.LHS1, B.dinv1.p() + B.offUV[c], s);
    }
    // scalar blocks, deflation constants, the case table
    B.sc.alloc((size_t)2 * B.nc * SC_COUNT);
    launch_fill(B.sc.p(), 0.0, 2 * B.nc * SC_COUNT, s);
    B.host_sc.assign((size_t)2 * B.nc * SC_COUNT, 0.0);
    const bool deflate = A[0]->fine_dev().defl != nullptr;
    if (deflate) {
        B.defl.alloc(B.nc);
        launch_fill(B.defl.p(), 0.0, B.nc, s);
    }
    B.tab_host.resize(B.nc);
    for (int c = 0; c < B.nc; c++) {
        luoIBM &S = *L[c];
        BCase &C = B.tab_host[c];
        C.g = S.grid();
        C.bc = S.bcs();
        C.bc.xp = B.xp.p() + B.offXP[c];
        C.nu = S.nu;
        C.dt = S.dt;
        C.offUV = B.offUV[c];
        C.offP = B.offP[c];
        C.offXP = B.offXP[c];
        C.numUV = S.numUV;
        C.numP = S.numP;
        C.v5 = static_cast<VelocityStencil *>(S.op1.get())->dev();
        C.v5.stop = nullptr;
        C.scV = B.sc.p() + (size_t)(2 * c) * SC_COUNT;
        C.scP = B.sc.p() + (size_t)(2 * c + 1) * SC_COUNT;
        C.p5 = A[c]->fine_dev();
        C.p5.stop = C.scP + SC_STOP;
        C.p5.defl = deflate ? B.defl.p() + c : nullptr;
        C.p5.defl_part = nullptr;  // the batch forms each case's c in k_b_wdot_finish
        if (deflate != (A[c]->fine_dev().defl != nullptr))
            throw Error("lockstep batch: the cases' AMG precisions must agree", ERR_ARGUMENT);
        C.fI = S.startI;
        C.fJ = S.startJ;
        C.fncx = S.numCellsX;
        C.fncy = S.numCellsY;
        C.ghostUV = S.ghostTagsUV.p();
    }
    B.tab.upload(B.tab_host, s);
    // block maps
    std::vector<long> cUV(B.nc), cP(B.nc), cXP(B.nc);
    for (int c = 0; c < B.nc; c++) {
        cUV[c] = L[c]->numUV;
        cP[c] = L[c]->numP;
        cXP[c] = 2 * L[c]->ny - 1;
    }
    B.mUV.build(cUV, BS, s);
    B.mP.build(cP, BS, s);
    B.mXP.build(cXP, BS, s);
    // interpolation stencils and solid nodes, concatenated
    std::vector<const Stencils *> sg, sh, shp, sgp;
    std::vector<int> solid;
    for (int c = 0; c < B.nc; c++) {
        sg.push_back(&L[c]->gnV);
        sh.push_back(&L[c]->hnV);
        shp.push_back(&L[c]->hnP);
        sgp.push_back(&L[c]->gnP);
        if (L[c]->nsolid) {
            std::vector<int> v = L[c]->solid.to_host(s);
            v.resize(L[c]->nsolid);
            for (int &i : v) {
                if (i < 0 || i >= L[c]->numUV) throw Error("lockstep batch: a solid node outside its case", ERR_GENERIC);
                i += (int)B.offUV[c];
            }
            solid.insert(solid.end(), v.begin(), v.end());
        }
    }
    cat_stencils(sg, B.offUV, B.nUV, B.gnV, s);
    cat_stencils(sh, B.offUV, B.nUV, B.hnV, s);
    cat_stencils(shp, B.offP, B.nP, B.hnP, s);
    cat_stencils(sgp, B.offP, B.nP, B.gnP, s);
    B.nsolid = (int)solid.size();
    if (B.nsolid) B.solid.upload(solid, s);
    // Krylov workspaces and the force
    for (auto *w : {&B.wV, &B.wP}) {
        const long n = w == &B.wV ? B.nUV : B.nP;
        const int nb = w == &B.wV ? B.mUV.nblk() : B.mP.nblk();
        for (auto *a : {&w->r, &w->rs, &w->p, &w->Mp, &w->AMp, &w->sv, &w->Ms, &w->AMs}) {
            a->alloc(n);
            launch_fill(a->p(), 0.0, n, s);
        }
        w->part.alloc(2 * (size_t)std::max(nb, 1));
    }
    B.fpart.alloc((size_t)6 * B.nc * BFORCE_BLOCKS);
    B.fout.alloc((size_t)6 * B.nc);
    launch_fill(B.fout.p(), 0.0, 6 * B.nc, s);
    B.amg.reset(new BatchAmg(A, B.offP, B.tab.p(), B.tab_host, B.mP.dev(), B.mP.first(), s));
    const double f = A[0]->restart_reduction();
    for (auto *a : A)
        if (a->restart_reduction() != f) throw Error("lockstep batch: cases of different -poissonRestart", ERR_ARGUMENT);
    B.reliable2 = f * f;
    B.mask.alloc(B.nc);
    B.host_mask.assign(B.nc, 0);
    CUIBM_HIP(hipStreamSynchronize(s));
}

LockstepBatch::~LockstepBatch() = default;

void LockstepBatch::step() {
    Impl &B = *im_;
    hipStream_t s = B.s;
    std::swap(B.N, B.Nold);
    // intermediate velocity (luoIBM.cu: generateRHS1, solve, weightUhat)
    launch_b_robin(B.tab.p(), B.mXP.dev(), B.u.p(), B.xp.p(), s);
    if (B.gnV.n) launch_gn_vel_apply(B.u.p(), B.gnV, 0, 0.0, 0.0, s);
    if (B.nsolid) launch_set_inside(B.u.p(), B.solid.p(), B.nsolid, 0, 0.0, 0.0, s);
    if (B.hnV.n) launch_hn_vel_apply(B.u.p(), B.ustar.p(), B.hnV, 0, 0.0, 0.0, s);
    launch_b_rhs1(B.tab.p(), B.mUV.dev(), B.u.p(), B.Nold.p(), B.N.p(), B.rhs1.p(), s);
    it1_ = B.solve(BOP_VEL, B.uhat.p(), B.rhs1.p(),
                   [&](const double *r, double *z) { launch_b_jacobi(B.nUV, B.dinv1.p(), r, z, s); }, B.tol1, B.maxit1,
                   "velocity");
    if (B.hnV.n) launch_weight(B.uhat.p(), B.ustar.p(), B.hnV, s);
    // pressure (generateRHS2, solvePoisson, weightPressure)
    if (B.gnV.n) launch_gn_vel_apply(B.uhat.p(), B.gnV, 0, 0.0, 0.0, s);
    launch_b_rhs2(B.tab.p(), B.mP.dev(), B.uhat.p(), B.rhs2.p(), s);
    it2_ = B.solve(BOP_POIS, B.p.p(), B.rhs2.p(), [&](const double *r, double *z) { B.amg->apply(r, z, s); }, B.tol2,
                   B.maxit2, "Poisson");
    if (B.hnP.n) {
        launch_hn_p_apply(B.p.p(), B.pstar.p(), B.hnP, 0.0, 0.0, s);
        launch_weight(B.p.p(), B.pstar.p(), B.hnP, s);
    }
    if (B.gnP.n) launch_gn_p_apply(B.p.p(), B.gnP, 0.0, 0.0, s);
    // projection and the body values (_project_velocity), the force (calculateForce)
    launch_b_project(B.tab.p(), B.mUV.dev(), B.u.p(), B.uhat.p(), B.uold.p(), B.p.p(), s);
    if (B.gnV.n) launch_gn_vel_apply(B.u.p(), B.gnV, 0, 0.0, 0.0, s);
    if (B.nsolid) launch_set_inside(B.u.p(), B.solid.p(), B.nsolid, 0, 0.0, 0.0, s);
    launch_b_force(B.tab.p(), B.nc, B.u.p(), B.uold.p(), B.p.p(), B.fpart.p(), B.fout.p(), s);
    for (auto *S : B.solver) static_cast<luoIBM *>(S)->timeStep++;
}

void LockstepBatch::coefficients(std::vector<double> &cd, std::vector<double> &cl) {
    Impl &B = *im_;
    std::vector<double> f(6 * (size_t)B.nc);
    B.fout.download(f.data(), f.size(), B.s);
    CUIBM_HIP(hipStreamSynchronize(B.s));
    cd.assign(B.nc, 0.0);
    cl.assign(B.nc, 0.0);
    for (int c = 0; c < B.nc; c++) {
        const double *o = f.data() + 6 * c;
        cd[B.order[c]] = 2.0 * (o[0] + o[1] + o[2]);
        cl[B.order[c]] = 2.0 * (o[3] + o[4] + o[5]);
    }
}

void LockstepBatch::copy_back() {
    Impl &B = *im_;
    hipStream_t s = B.s;
    for (int c = 0; c < B.nc; c++) {
        luoIBM &S = *static_cast<luoIBM *>(B.solver[c]);
        seg_copy(S.u.p(), B.u.p() + B.offUV[c], S.numUV, s);
        seg_copy(S.uhat.p(), B.uhat.p() + B.offUV[c], S.numUV, s);
        seg_copy(S.uold.p(), B.uold.p() + B.offUV[c], S.numUV, s);
        seg_copy(S.N.p(), B.N.p() + B.offUV[c], S.numUV, s);
        seg_copy(S.Nold.p(), B.Nold.p() + B.offUV[c], S.numUV, s);
        seg_copy(S.pressure.p(), B.p.p() + B.offP[c], S.numP, s);
        seg_copy(S.bc[XPLUS].p(), B.xp.p() + B.offXP[c], 2 * S.ny - 1, s);
    }
    std::vector<double> f(6 * (size_t)B.nc);
    B.fout.download(f.data(), f.size(), s);
    CUIBM_HIP(hipStreamSynchronize(s));
    for (int c = 0; c < B.nc; c++) {
        luoIBM &S = *static_cast<luoIBM *>(B.solver[c]);
        seg_copy(S.forceOut.p(), B.fout.p() + 6 * c, 6, s);
        CUIBM_HIP(hipStreamSynchronize(s));
        const double *o = f.data() + 6 * c;
        S.fxx = o[0];
        S.fxy = o[1];
        S.fxu = o[2];
        S.forceX = o[0] + o[1] + o[2];
        S.forceY = o[3] + o[4] + o[5];
    }
}

}  // namespace cuibm
