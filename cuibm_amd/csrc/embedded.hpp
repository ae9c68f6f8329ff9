// The Luo embedded method (SolverType OSC_CYLINDER, luo_iter.h): the ghost- and
// hybrid-node interpolation is folded into the velocity and Poisson matrices.
//
// Assembly (luo_iter/kernels/LHS1.cu:12-408, LHS2.cu:12-224) runs once per tagging:
// every matrix entry depends on the geometry only, so for a static body the
// reference's per-step reassembly and preconditioner rebuild reproduce the same
// matrices. One thread per row writes at most six sorted entries into a row buffer;
// the host turns the row counts into row pointers and a second kernel compacts the
// buffers into CSR. The right-hand-side update rhs = rhs * ns_rhs + interp_rhs
// (luo_iter/kernels/intermediateVelocity.cu:48-73, pressure.cu:36-46) runs every step,
// with interp_rhs evaluated from the live fields and body velocity.
#pragma once
#include "ibm.hpp"

namespace cuibm {

// one interpolation stencil set as the embedded kernels read it (slot-indexed)
struct EmbStencil {
    const int *idx0;      // corner indices before substitution (4 per slot)
    const int *idx;       // < 0 where the corner was replaced by the body intercept
    const double *coef;   // interpolation weights (4 per slot)
    const double *nrm;    // Neumann normals of replaced pressure corners (8 per slot)
    const double *extra;  // alpha (hybrid) or (distance, close normal x, y) (pressure ghost)
};

// one system's assembly inputs and outputs
struct EmbAsm {
    int n;                          // rows
    const int *map;                 // row -> -1 plain, t >= 0 hybrid slot t, -(t + 2) ghost slot t
    EmbStencil hyb, gho;
    const int *ghost;               // ghost tags of the system (ghostTagsUV or ghostTagsP)
    const int *prp, *pci;           // plain CSR, whose boundary rows are kept
    const double *pv;
    int *cnt;                       // entries per row
    int *bc;                        // row buffers: 6 columns per row
    double *bv;                     //              6 values per row
    double *ns;                     // right-hand-side row scale
    int *stats;                     // slot anomalies: unwritten 5-point slots, tail slots
                                    // never written, written twice, dropped corners
};

constexpr int EMB_ROW = 6;

void launch_emb_lhs1(const GridDev &g, double nu, double dt, const EmbAsm &a, hipStream_t s);
void launch_emb_lhs2(const GridDev &g, double dt, const EmbAsm &a, hipStream_t s);
void launch_emb_compact(const EmbAsm &a, const int *rp, int *ci, double *v, hipStream_t s);
// row pointers from per-row counts on the device (replaces the reference's single-thread
// size_LHS kernels, luo_iter/kernels/LHS1.cu, LHS2.cu): rp[0] = 0, rp[i+1] = rp[i] + cnt[i];
// `part` holds at least scan_parts(n) ints
int scan_parts(int n);
void launch_row_scan(const int *cnt, int *rp, int n, int *part, hipStream_t s);
// row map of the embedded assembly: -1, then -(t + 2) at ghost node t, then t at hybrid
// node t (a node in both lists is a hybrid row)
void launch_emb_map(int *map, int n, const int *ghost, int ng, const int *hybrid, int nh, hipStream_t s);
// rhs1 = rhs1 * ns + interp for interior velocity rows (v rows only below the thread
// count of the u launch, as in the reference)
void launch_emb_rhs1(double *rhs1, const double *u, const EmbAsm &a, const GridDev &g, double uB, double vB,
                     hipStream_t s);
// rhs2 = rhs2 * ns + interp for interior pressure rows below (nx - 1) ny
void launch_emb_rhs2(double *rhs2, const EmbAsm &a, const GridDev &g, double dudt, double dvdt, hipStream_t s);

}  // namespace cuibm
