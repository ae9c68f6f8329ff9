// Device-side views and kernel launchers shared by the solver classes.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "dev.hpp"

namespace cuibm {

// read-only grid geometry on the device (domain.h: x, dx, y, dy, xu, yv; xv = x, yu = y)
struct GridDev {
    const double *x, *dx, *y, *dy, *xu, *yv;
    int nx, ny;
};

// boundary velocity arrays (NavierStokes/cast.cu:48-51): xm/xp hold u per J then v per J,
// ym/yp hold u per I then v per I
struct BcDev {
    double *xm, *xp, *ym, *yp;
};

// modified-Fadlun rhs1 override data; hybrid == nullptr disables it
struct FadlunRhsDev {
    const int *hybrid, *ghost;
    const double *dist_int, *dist_between, *uv;
};

// modified-Fadlun rhs2 data; ghostP == nullptr selects the plain divergence
struct FadlunP2Dev {
    const int *ghostP, *hybridP;
    const double *du, *dv;
};

// modified-Fadlun matrix data; ghostUV == nullptr selects the plain operators
struct FadlunLhsDev {
    const int *ghostUV, *hybridUV, *hybridUV2, *ghostP, *hybridP;
    const double *dist_int, *dist_between, *du, *dv;
};

// ---- k_ns.hip
void launch_rhs1(const double *u, const double *Nold, double *N, double *rhs, double *L, double *bc1, const GridDev &g,
                 const BcDev &b, double nu, double dt, const FadlunRhsDev &fad, hipStream_t s);
void launch_rhs2(const double *uhat, double *rhs2, const GridDev &g, const BcDev &b, const FadlunP2Dev &fad,
                 hipStream_t s);
void launch_project(double *u, const double *uhat, double *uold, const double *p, const GridDev &g, double dt,
                    const int *ghostP, hipStream_t s);
void launch_robin(const double *u, double *xp, const GridDev &g, double dt, hipStream_t s);
// maximum CFL number over the cells and the first cell index attaining it: out[0], out[1];
// part holds CFL_BLOCKS doubles then CFL_BLOCKS ints
constexpr int CFL_BLOCKS = 256;
void launch_cfl(const double *u, const GridDev &g, double dt, double *part, double *out, hipStream_t s);

// ---- k_lhs.hip
void lhs1_rowptr(int nx, int ny, std::vector<int> &rp);
void lhs2_rowptr(int nx, int ny, std::vector<int> &rp);
void launch_lhs1(const int *rp, int *ci, double *v, const GridDev &g, double nu, double dt, const FadlunLhsDev &fad,
                 hipStream_t s);
void launch_lhs2(const int *rp, int *ci, double *v, const GridDev &g, double dt, bool centre_pin,
                 const FadlunLhsDev &fad, hipStream_t s);

}  // namespace cuibm
