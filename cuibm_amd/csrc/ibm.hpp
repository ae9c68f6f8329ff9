// Immersed-boundary device data and kernel launchers (Luo external and modified
// Fadlun methods): ray-cast tagging, interpolation stencils, blending, force.
#pragma once
#include "kernels.hpp"

namespace cuibm {

// one velocity component (u or v) of the staggered grid as seen by the tagging and
// interpolation kernels
struct CompDev {
    const double *xs, *ys;  // node coordinates: (xu, yu) for u, (xv, yv) for v
    int ni, nj;             // nodes per row / rows
    int stride, off;        // row stride, offset inside the packed (u, v) vector
};

struct BoxDev {
    int i_start, j_start, w, h;  // tagging box in cell indices
    int nthr;                     // threads of the reference launch (w*h rounded up to 256)
};

// per-thread tagging records (struct of arrays, nthr entries each)
struct TagRecords {
    DArray<int> own, nb0, nb1, nb2, nb3, flags;
    DArray<double> dv0, dv1;  // distance_from_{u,v}_to_body writes (own, next)
    void alloc(int n) {
        for (auto *a : {&own, &nb0, &nb1, &nb2, &nb3, &flags}) a->alloc(n);
        dv0.alloc(n);
        dv1.alloc(n);
    }
};

struct TagOut {
    int *ghostUV, *hybridUV, *hybridUV2, *ghostP, *hybridP;
    double *bi_x, *bi_y, *ip_x, *ip_y;          // numUV
    double *bip_x, *bip_y, *ipp_x, *ipp_y;      // numP
    double *dist_int, *dist_between, *uv;       // numUV
    double *du, *dv;                            // numP
};

struct BodyDev {
    const double *x, *y, *uB, *vB;
    int n;
    double midX, midY;
};

// tagging: records, then resolution of the neighbour writes, then the interior fill
void launch_tag_velocity(const BodyDev &b, const CompDev &c, const BoxDev &box, bool vdir, TagRecords &rec,
                         const TagOut &o, int nx, hipStream_t s);
void launch_tag_velocity_resolve(const CompDev &c, const BoxDev &box, const BodyDev &b, bool vdir,
                                 const TagRecords &rec, const TagOut &o, int nx, int ny, hipStream_t s);
void launch_tag_pressure(const BodyDev &b, const GridDev &g, const BoxDev &box, TagRecords &rec, hipStream_t s);
void launch_tag_pressure_resolve(const GridDev &g, const BoxDev &box, const BodyDev &b, const TagRecords &rec,
                                 const TagOut &o, hipStream_t s);
void launch_zero_rows(int *tags, const BoxDev &box, int rowlen, int kmax, int off, int nrows, long len,
                      hipStream_t s);

// interpolation stencils: value at a node = sum_l coef_l q_l with q_l = field[idx_l],
// or, where idx_l < 0, a boundary value (body velocity / Neumann datum)
struct Stencils {
    int n = 0;
    int nrm_per = 0, extra_per = 0;  // values of nrm and extra per node (arrays may be longer)
    DArray<int> node, idx;        // node[n], idx[4n]
    DArray<int> idx0;             // idx before the boundary-corner substitution (4n)
    DArray<double> coef;          // coef[4n]
    DArray<double> nrm;           // Neumann corner normals (nx, ny) per corner: 8n
    DArray<double> extra;         // per node: alpha (hybrid nodes) or dist/close-normal (pressure ghost: 3n)
    DArray<double> tmp;           // staging for the ghost-node velocity update
    DArray<double> ub;            // per-node body velocity at the node's intercept (empty: uB[0] for all)
};

// The ordered node lists of one tagging, built on the device (two passes over fixed chunks
// of candidates: counts, then in-order writes): [0] ghost and [1] hybrid velocity nodes
// (tag > 0), [2] ghost and [3] hybrid pressure nodes, [4] solid velocity nodes (ghost tag
// 0) in setInsideVelocity's order (u node i, then the v node J nx + I + numU of the same
// I, J = i / (nx - 1)). out[k] must hold every candidate of list k; counts[5] receives
// the lengths; part holds node_list_parts(...) ints.
struct NodeListIn {
    const int *ghostUV, *hybridUV, *ghostP, *hybridP;
    int numU, numUV, numP, nx;
};
int node_list_parts(int numUV, int numP, int numU);
void launch_node_lists(const NodeListIn &in, int *const out[5], int *part, int *counts, hipStream_t s);

struct IbmGeom {
    const int *ghostUV, *hybridUV, *ghostP, *hybridP;
    const double *bi_x, *bi_y, *ip_x, *ip_y, *bip_x, *bip_y, *ipp_x, *ipp_y;
};

void launch_gn_vel_coef(const GridDev &g, const IbmGeom &m, Stencils &st, hipStream_t s);
void launch_hn_vel_coef(const GridDev &g, const IbmGeom &m, Stencils &st, hipStream_t s);
void launch_hn_p_coef(const GridDev &g, const IbmGeom &m, Stencils &st, hipStream_t s);
void launch_gn_p_coef(const GridDev &g, const IbmGeom &m, Stencils &st, hipStream_t s);

void launch_gn_vel_apply(double *f, Stencils &st, int numU, double uB, double vB, hipStream_t s);
// solid nodes take uB / vB, or ubn[t] for the t-th solid node when ubn is given
void launch_set_inside(double *f, const int *solid, int nsolid, int numU, double uB, double vB, hipStream_t s,
                       const double *ubn = nullptr);
void launch_hn_vel_apply(const double *f, double *ustar, const Stencils &st, int numU, double uB, double vB,
                         hipStream_t s);
void launch_weight(double *f, const double *star, const Stencils &st, hipStream_t s);
void launch_hn_p_apply(const double *p, double *pstar, const Stencils &st, double dudt, double dvdt, hipStream_t s);
void launch_gn_p_apply(double *p, const Stencils &st, double dudt, double dvdt, hipStream_t s);

// control-volume force (Lai & Peskin): out[0..5] = FxX, FxY, FxU, FyX, FyY, FyU
void launch_force(const double *u, const double *uold, const double *p, const int *ghostUV, const GridDev &g,
                  double nu, double dt, int I, int J, int ncx, int ncy, double *part, double *out, hipStream_t s);

}  // namespace cuibm
