// ORACLE — test infrastructure only (see orc.hpp).
// Immersed-boundary pieces of one time step, restating (paths under src/solvers/NavierStokes/):
//   src/bodies.cu:257-326                               bounding box (boundingBox kernel)
//   luo_base/tagPoints.cu:17-73, luo_base/kernels/tagPoints.cu:14-1298
//                                                        ray-cast tagging + interior fill
//   luo_base/kernels/biLinearInterpolation.cu:14-1066    ghost/hybrid bilinear interpolation
//   luo_base/kernels/intermediateVelocity.cu:18-31       setInsideVelocity
//   luoIBM/kernels/weight.cu:11-134                      alpha blending
//   FadlunModified/kernels/intermediateVelocity.cu:19-63 convective outflow, Fadlun rhs1
//   FadlunModified/kernels/intermediatePressure.cu:17-138 Fadlun rhs2
//   FadlunModified/kernels/projectVelocity.cu:14-44       masked projection
//   FadlunModified/kernels/calculateForce.cu:32-283       control-volume force
//
// Racy writes. The tagging kernels write their neighbours' ghost tags and intercept
// data, and the ghost-node velocity interpolation reads corner values that other
// threads of the same launch overwrite. The oracle fixes both:
//   tagging: every initialisation write first, then the box threads one after
//            another in increasing thread index, each in program order (so among
//            several writers of one entry the highest node index wins);
//   ghost-node velocity interpolation: all reads see the field as it was before
//            the launch (results are computed first, then stored).
// The diagnostic-only arrays (x1..y4, q1..q4 of the tagging/interpolation kernels)
// are not reproduced.
#include <algorithm>
#include <cmath>
#include "orc.hpp"

namespace orc {

static const double EPS = 1.e-10;

// squares are products (pow(x, 2) in the reference): the same rounding on every platform
static inline double sq(double x) { return x * x; }

// foot of the perpendicular from (px,py) to the line through body points k,l
// (tagPoints.cu:122-133, same construction at every call site)
static inline void foot(const double *bx, const double *by, int k, int l, double px, double py,
                        double &fx, double &fy) {
    // a / tan(theta) of the reference, with theta = acos(c), a = b sin(theta) and
    // c = (p^2 + b^2 - o^2) / (2 p b), is b cos(theta) = (p^2 + b^2 - o^2) / (2 p)
    const double p = std::sqrt(sq((bx[l] - bx[k])) + sq((by[l] - by[k])));
    const double o = std::sqrt(sq((bx[l] - px)) + sq((by[l] - py)));
    const double b = std::sqrt(sq((px - bx[k])) + sq((py - by[k])));
    const double d = (p * p + b * b - o * o) / (2 * p);
    fy = by[k] + (by[l] - by[k]) / p * d;
    fx = bx[k] + (bx[l] - bx[k]) / p * d;
}

// src/bodies.cu:257-326 (the box is computed from the body points; x,y are cell centres)
void Solver::boundingBox() {
    const double scale = P.scaleCV;
    double xmax = *std::max_element(B.x.begin(), B.x.end());
    double xmin = *std::min_element(B.x.begin(), B.x.end());
    double ymax = *std::max_element(B.y.begin(), B.y.end());
    double ymin = *std::min_element(B.y.begin(), B.y.end());
    const double dxb = xmax - xmin, dyb = ymax - ymin;
    xmax += 0.5 * dxb * (scale - 1.0);
    xmin -= 0.5 * dxb * (scale - 1.0);
    ymax += 0.5 * dyb * (scale - 1.0);
    ymin -= 0.5 * dyb * (scale - 1.0);
    int i = 0, j = 0;
    while (G.x[i + 1] < xmin) i++;
    while (G.y[j + 1] < ymin) j++;
    B.startI = i;
    B.startJ = j;
    while (G.x[i] < xmax) i++;
    while (G.y[j] < ymax) j++;
    B.numCellsX = i - B.startI;
    B.numCellsY = j - B.startJ;
}

// One velocity-node thread of tag_u_luo (tagPoints.cu:14-427, vdir = false) or
// tag_v_luo (tagPoints.cu:430-826, vdir = true).
struct TagCtx {
    Solver *S;
    const double *bx, *by, *ub;  // body points and the body velocity component
    const double *xs, *ys;       // node coordinates (xu,yu) or (xv,yv)
    int stride, off;             // row stride and offset of this component in the UV vector
    double midX, midY;
    int n;
};

static void tag_velocity_thread(TagCtx &c, int I, int J, bool vdir) {
    Solver &S = *c.S;
    const int nx = S.nx;
    const int st = c.stride;
    const int iu = J * st + I + c.off, ip = J * nx + I;
    const double *xs = c.xs, *ys = c.ys, *bx = c.bx, *by = c.by;
    int k = c.n - 1, l = 0;
    bool outsideX = true, outsideY = true, flag = false;
    int bdryFlagX = -1, bdryFlagY = -1, bdryFlag2X = -1, bdryFlag2Y = -1;
    double uvX = 0, uvY = 0, Xa = 1, Ya = 1, Xb = 1, Yb = 1, x = 0, y = 0;
    auto ghost_from = [&](int g, double gx, double gy, int kk, int ll) {
        S.ghostTagsUV[g] = g;
        double fx, fy;
        foot(bx, by, kk, ll, gx, gy, fx, fy);
        S.bi_y[g] = fy;
        S.bi_x[g] = fx;
        S.ip_y[g] = fy + (fy - gy);
        S.ip_x[g] = fx + (fx - gx);
    };
    auto hybrid_self = [&](int kk, int ll) {
        double fx, fy;
        foot(bx, by, kk, ll, xs[I], ys[J], fx, fy);
        S.bi_y[iu] = fy;
        S.bi_x[iu] = fx;
        S.ip_y[iu] = 2 * ys[J] - fy;
        S.ip_x[iu] = 2 * xs[I] - fx;
    };
    auto coincident = [&](int g, double ipx, double ipy) {
        S.ghostTagsUV[g] = g;
        S.bi_x[g] = xs[I];
        S.bi_y[g] = ys[J];
        S.ip_x[g] = ipx;
        S.ip_y[g] = ipy;
        S.bi_x[iu] = xs[I];
        S.bi_y[iu] = ys[J];
        S.ip_x[iu] = ipx;
        S.ip_y[iu] = ipy;
    };
    while (l < c.n && (vdir || !flag)) {
        int bottom, top, left, right;
        if (by[k] > by[l]) { bottom = l; top = k; } else { bottom = k; top = l; }
        if (bx[k] > bx[l]) { left = l; right = k; } else { left = k; right = l; }
        // rays along x
        if (by[bottom] - EPS < ys[J] && by[top] - EPS > ys[J] && (!vdir || !flag)) {
            if (std::fabs(by[l] - by[k]) > EPS) {
                x = bx[k] + (bx[l] - bx[k]) * (ys[J] - by[k]) / (by[l] - by[k]);
                uvX = c.ub[k] + (c.ub[l] - c.ub[k]) * (ys[J] - by[k]) / (by[l] - by[k]);
                if (x > xs[I] + EPS) outsideX = !outsideX;
                if (x > xs[I - 1] + EPS && x < xs[I] - EPS) {
                    bdryFlagX = iu;
                    bdryFlag2X = iu + 1;
                    if (x > c.midX + EPS) {
                        ghost_from(iu - 1, xs[I - 1], ys[J], k, l);
                        hybrid_self(k, l);
                    }
                    Xa = xs[I] - x;
                    Xb = xs[I + 1] - xs[I];
                    if (!vdir && x > c.midX) S.dist_u_body[ip] = Xa;
                } else if (x > xs[I] + EPS && x < xs[I + 1] - EPS) {
                    bdryFlagX = iu;
                    bdryFlag2X = iu - 1;
                    if (x < c.midX - EPS) {
                        ghost_from(iu + 1, xs[I + 1], ys[J], k, l);
                        hybrid_self(k, l);
                    }
                    Xa = x - xs[I];
                    Xb = xs[I] - xs[I - 1];
                    if (!vdir && x < c.midX) S.dist_u_body[ip + 1] = Xa;
                } else if (std::fabs(x - xs[I]) < EPS && std::fabs(c.midX - xs[I]) > std::fabs(c.midY - ys[J])) {
                    outsideX = true;
                    bdryFlagX = iu;
                    if (x < c.midX) {
                        bdryFlag2X = iu - 1;
                        coincident(iu + 1, xs[I - 1], ys[J]);
                        Xa = 0;
                        Xb = xs[I] - xs[I - 1];
                    }
                    if (x > c.midX) {
                        bdryFlag2X = iu + 1;
                        coincident(iu - 1, xs[I + 1], ys[J]);
                        Xa = 0;
                        Xb = xs[I + 1] - xs[I];
                    }
                    flag = true;
                }
            }
        }
        // rays along y
        if (bx[left] - EPS < xs[I] && bx[right] - EPS > xs[I] && (!vdir || !flag)) {
            if (std::fabs(bx[l] - bx[k]) > EPS) {
                y = by[k] + (by[l] - by[k]) * (xs[I] - bx[k]) / (bx[l] - bx[k]);
                uvY = c.ub[k] + (c.ub[l] - c.ub[k]) * (xs[I] - bx[k]) / (bx[l] - bx[k]);
                if (y > ys[J] + EPS) outsideY = !outsideY;
                if (y > ys[J - 1] + EPS && y < ys[J] - EPS) {
                    bdryFlagY = iu;
                    bdryFlag2Y = iu + st;
                    if (y > c.midY + EPS) {
                        ghost_from(iu - st, xs[I], ys[J - 1], k, l);
                        hybrid_self(k, l);
                    }
                    Ya = ys[J] - y;
                    Yb = ys[J + 1] - ys[J];
                    if (vdir && outsideY) S.dist_v_body[ip] = Ya;
                } else if (y > ys[J] + EPS && y < ys[J + 1] - EPS) {
                    bdryFlagY = iu;
                    bdryFlag2Y = iu - st;
                    if (y < c.midY - EPS) {
                        ghost_from(iu + st, xs[I], ys[J + 1], k, l);
                        hybrid_self(k, l);
                    }
                    Ya = y - ys[J];
                    Yb = ys[J] - ys[J - 1];
                    if (vdir && outsideY) S.dist_v_body[ip + nx] = Ya;
                } else if (vdir ? (std::fabs(y - ys[J]) < EPS && std::fabs(c.midY - ys[J]) > std::fabs(c.midX - xs[I]))
                                : (std::fabs(y - ys[J]) < EPS && std::fabs(y - c.midY) > std::fabs(x - c.midX))) {
                    outsideY = true;
                    bdryFlagY = iu;
                    if (y < c.midY) {
                        bdryFlag2X = iu - st;  // the reference sets the X flag here
                        coincident(iu + st, xs[I], ys[J - 1]);
                        Ya = 0;
                        Yb = ys[J] - ys[J - 1];
                    }
                    if (y > c.midY) {
                        bdryFlag2X = iu + st;
                        coincident(iu - st, xs[I], ys[J + 1]);
                        Ya = 0;
                        Yb = ys[J + 1] - ys[J];
                    }
                    flag = true;
                }
            }
        }
        k = l;
        l = l + 1;
    }
    auto finalX = [&]() {
        S.ghostTagsUV[iu] = -1;
        S.hybridTagsUV[iu] = bdryFlagX;
        S.hybridTagsUV2[iu] = bdryFlag2X;
        S.dist_int[iu] = Xa;
        S.dist_between[iu] = Xb;
        S.uv[iu] = uvX;
    };
    auto finalY = [&]() {
        S.ghostTagsUV[iu] = -1;
        S.hybridTagsUV[iu] = bdryFlagY;
        S.hybridTagsUV2[iu] = bdryFlag2Y;
        S.dist_int[iu] = Ya;
        S.dist_between[iu] = Yb;
        S.uv[iu] = uvY;
    };
    if (!vdir) {
        if (outsideX && bdryFlagX >= 0) finalX();
        else if (outsideY && bdryFlagY >= 0) finalY();
    } else {
        if (outsideY && bdryFlagY >= 0) finalY();
        else if (outsideX && bdryFlagX >= 0) finalX();
    }
}

// one pressure-node thread of tag_p_luo (tagPoints.cu:829-1154)
static void tag_pressure_thread(Solver &S, const double *bx, const double *by, int n, double midX,
                                double midY, int I, int J) {
    const int nx = S.nx, ip = J * nx + I;
    const double *xv = S.G.x.data(), *yu = S.G.y.data();
    int k = n - 1, l = 0;
    bool flag = false;
    double x, y;
    auto ghost_self = [&](int kk, int ll) {
        S.ghostTagsP[ip] = ip;
        double fx, fy;
        foot(bx, by, kk, ll, xv[I], yu[J], fx, fy);
        S.bip_y[ip] = fy;
        S.bip_x[ip] = fx;
        S.ipp_y[ip] = fy + (fy - yu[J]);
        S.ipp_x[ip] = fx + (fx - xv[I]);
    };
    auto hybrid_self = [&](int kk, int ll) {
        S.hybridTagsP[ip] = ip;
        double fx, fy;
        foot(bx, by, kk, ll, xv[I], yu[J], fx, fy);
        S.bip_y[ip] = fy;
        S.bip_x[ip] = fx;
        S.ipp_y[ip] = 2 * yu[J] - fy;
        S.ipp_x[ip] = 2 * xv[I] - fx;
    };
    auto coincident = [&](int g, double ipx, double ipy) {
        S.ghostTagsP[g] = g;
        S.bip_x[g] = xv[I];
        S.bip_y[g] = yu[J];
        S.ipp_x[g] = ipx;
        S.ipp_y[g] = ipy;
        S.bip_x[ip] = xv[I];
        S.bip_y[ip] = yu[J];
        S.ipp_x[ip] = ipx;
        S.ipp_y[ip] = ipy;
    };
    while (l < n) {
        int bottom, top, left, right;
        if (by[k] > by[l]) { bottom = l; top = k; } else { bottom = k; top = l; }
        if (bx[k] > bx[l]) { left = l; right = k; } else { left = k; right = l; }
        if (by[bottom] - EPS < yu[J] && by[top] - EPS > yu[J] && !flag) {
            if (std::fabs(by[l] - by[k]) > EPS) {
                x = bx[k] + (bx[l] - bx[k]) * (yu[J] - by[k]) / (by[l] - by[k]);
                if (x > midX + EPS && x > xv[I] + EPS + EPS && x < xv[I + 1] - EPS) ghost_self(k, l);
                else if (x < midX + EPS && x < xv[I] - EPS && x > xv[I - 1] + EPS) ghost_self(k, l);
                if (x > midX + EPS && x < xv[I] - EPS && x > xv[I - 1] + EPS) hybrid_self(k, l);
                else if (x < midX + EPS && x > xv[I] + EPS && x < xv[I + 1] - EPS) hybrid_self(k, l);
                if (std::fabs(x - xv[I]) < EPS && std::fabs(midX - xv[I]) > std::fabs(midY - yu[J])) {
                    S.hybridTagsP[ip] = ip;
                    if (x < midX) coincident(ip + 1, xv[I - 1], yu[J]);
                    else if (x > midX) coincident(ip - 1, xv[I + 1], yu[J]);
                    flag = true;
                }
            }
        }
        if (bx[left] - EPS < xv[I] && bx[right] - EPS > xv[I] && !flag) {
            if (std::fabs(bx[l] - bx[k]) > EPS) {
                y = by[k] + (by[l] - by[k]) * (xv[I] - bx[k]) / (bx[l] - bx[k]);
                if (y > midY + EPS && y > yu[J] + EPS && y < yu[J + 1] - EPS) ghost_self(k, l);
                else if (y < midY + EPS && y < yu[J] - EPS && y > yu[J - 1] + EPS) ghost_self(k, l);
                if (y > midY + EPS && y < yu[J] - EPS && y > yu[J - 1] + EPS) hybrid_self(k, l);
                else if (y < midY + EPS && y > yu[J] + EPS && y < yu[J + 1] - EPS) hybrid_self(k, l);
                if (std::fabs(y - yu[J]) < EPS && std::fabs(midY - yu[J]) > std::fabs(midX - xv[I])) {
                    S.hybridTagsP[ip] = ip;
                    if (y < midY) coincident(ip + nx, xv[I], yu[J - 1]);
                    if (y > midY) coincident(ip - nx, xv[I], yu[J + 1]);
                    flag = true;
                }
            }
        }
        k = l;
        l = l + 1;
    }
}

// zero_pressure_luo / zero_x_luo / zero_y_luo (tagPoints.cu:1157-1298): serial row
// scans that mark the body interior solid (tag 0). `rowlen` is the row stride,
// `last` the last index k of the look-ahead loop, `off` the component offset.
static void zero_rows(std::vector<int> &t, int i_start, int j_start, int i_end, int j_end, int rowlen,
                      int kmax, int off, int nthreads, int nrows) {
    for (int jt = 0; jt < nthreads; jt++) {
        const int J = j_start + jt;
        if (J > j_end || J >= nrows) continue;
        bool rowIsntDone = true, flag = false;
        for (int i = i_start; i < i_end; i++) {
            const int I = J * rowlen + i + off;
            if (t[I - 1] >= 0) {
                if (t[I] != -1 && t[I - 1] == 0) rowIsntDone = false;
                if (t[I] == -1 && rowIsntDone) t[I] = 0;
                if (t[I] == 0 && t[I - 1] != 0) {
                    int k = i;
                    while (k < kmax) {
                        k++;
                        const long q = (long)J * rowlen + k + off;
                        if (q < (long)t.size() && t[q] != -1) flag = true;
                    }
                    if (!flag) {
                        t[I] = -1;
                        rowIsntDone = false;
                    }
                }
            }
        }
    }
}

// luo_base/tagPoints.cu:17-73
void Solver::tagPoints() {
    std::fill(ghostTagsUV.begin(), ghostTagsUV.end(), -1);
    std::fill(ghostTagsP.begin(), ghostTagsP.end(), -1);
    std::fill(hybridTagsUV.begin(), hybridTagsUV.end(), -1);
    std::fill(hybridTagsP.begin(), hybridTagsP.end(), -1);
    std::fill(hybridTagsUV2.begin(), hybridTagsUV2.end(), -1);
    std::fill(dist_int.begin(), dist_int.end(), 1.0);
    std::fill(dist_between.begin(), dist_between.end(), 1.0);
    std::fill(bi_x.begin(), bi_x.end(), 0.0);
    std::fill(bi_y.begin(), bi_y.end(), 0.0);
    std::fill(ip_x.begin(), ip_x.end(), 0.0);
    std::fill(ip_y.begin(), ip_y.end(), 0.0);
    const int i_start = B.startI, j_start = B.startJ, w = B.numCellsX, h = B.numCellsY;
    const int i_end = i_start + w, j_end = j_start + h;
    const int nthr = ((int)((w * (double)h - 0.5) / 256) + 1) * 256;
    // tag_u_luo: its per-thread initialisation first, then the threads in order
    for (int idx = 0; idx < nthr; idx++) {
        const int I = i_start + idx % w, J = j_start + idx / w;
        if (J * (nx - 1) + I >= numU) continue;
        dist_u_body[J * nx + I] = 0;
    }
    TagCtx cu{this, B.x.data(), B.y.data(), B.uB.data(), G.xu.data(), G.y.data(), nx - 1, 0, B.midX, B.midY, B.n};
    // nodes on the outer grid line are skipped: their stencils would read outside the
    // coordinate arrays (the reference reads past them there)
    for (int idx = 0; idx < nthr; idx++) {
        const int I = i_start + idx % w, J = j_start + idx / w;
        if (J * (nx - 1) + I >= numU) continue;
        if (I < 1 || J < 1 || I >= nx - 2 || J >= ny - 1) continue;
        tag_velocity_thread(cu, I, J, false);
    }
    TagCtx cv{this, B.x.data(), B.y.data(), B.vB.data(), G.x.data(), G.yv.data(), nx, numU, B.midX, B.midY, B.n};
    for (int idx = 0; idx < nthr; idx++) {
        const int I = i_start + idx % w, J = j_start + idx / w;
        if (J * nx + I + numU >= numUV) continue;
        if (I < 1 || J < 1 || I >= nx - 1 || J >= ny - 2) continue;
        tag_velocity_thread(cv, I, J, true);
    }
    for (int idx = 0; idx < nthr; idx++) {
        const int I = i_start + idx % w, J = j_start + idx / w;
        if (J * nx + I >= numP) continue;
        if (I < 1 || J < 1 || I >= nx - 1 || J >= ny - 1) continue;
        tag_pressure_thread(*this, B.x.data(), B.y.data(), B.n, B.midX, B.midY, I, J);
    }
    const int nthr0 = ((int)((i_end - i_start - 0.5) / 256) + 1) * 256;
    zero_rows(ghostTagsP, i_start, j_start, i_end, j_end, nx, nx, 0, nthr0, ny);
    zero_rows(ghostTagsUV, i_start, j_start, i_end, j_end, nx - 1, nx - 1, 0, nthr0, ny);
    zero_rows(ghostTagsUV, i_start, j_start, i_end, j_end, nx, nx - 1, numU, nthr0, ny - 1);
}

// 4x4 cofactor inverse of the bilinear system (biLinearInterpolation.cu:129-168)
struct Bil {
    double det, b[4][4];  // b[r][c] = b_{r+1,c+1}
};
static Bil bil_inverse(const double a[16]) {
    const double a11 = a[0], a12 = a[1], a13 = a[2], a14 = a[3];
    const double a21 = a[4], a22 = a[5], a23 = a[6], a24 = a[7];
    const double a31 = a[8], a32 = a[9], a33 = a[10], a34 = a[11];
    const double a41 = a[12], a42 = a[13], a43 = a[14], a44 = a[15];
    Bil r;
    r.det = a11 * a22 * a33 * a44 + a11 * a23 * a34 * a42 + a11 * a24 * a32 * a43
          + a12 * a21 * a34 * a43 + a12 * a23 * a31 * a44 + a12 * a24 * a33 * a41
          + a13 * a21 * a32 * a44 + a13 * a22 * a34 * a41 + a13 * a24 * a31 * a42
          + a14 * a21 * a33 * a42 + a14 * a22 * a31 * a43 + a14 * a23 * a32 * a41
          - a11 * a22 * a34 * a43 - a11 * a23 * a32 * a44 - a11 * a24 * a33 * a42
          - a12 * a21 * a33 * a44 - a12 * a23 * a34 * a41 - a12 * a24 * a31 * a43
          - a13 * a21 * a34 * a42 - a13 * a22 * a31 * a44 - a13 * a24 * a32 * a41
          - a14 * a21 * a32 * a43 - a14 * a22 * a33 * a41 - a14 * a23 * a31 * a42;
    r.b[0][0] = a22 * a33 * a44 + a23 * a34 * a42 + a24 * a32 * a43 - a22 * a34 * a43 - a23 * a32 * a44 - a24 * a33 * a42;
    r.b[0][1] = a12 * a34 * a43 + a13 * a32 * a44 + a14 * a33 * a42 - a12 * a33 * a44 - a13 * a34 * a42 - a14 * a32 * a43;
    r.b[0][2] = a12 * a23 * a44 + a13 * a24 * a42 + a14 * a22 * a43 - a12 * a24 * a43 - a13 * a22 * a44 - a14 * a23 * a42;
    r.b[0][3] = a12 * a24 * a33 + a13 * a22 * a34 + a14 * a23 * a32 - a12 * a23 * a34 - a13 * a24 * a32 - a14 * a22 * a33;
    r.b[1][0] = a21 * a34 * a43 + a23 * a31 * a44 + a24 * a33 * a41 - a21 * a33 * a44 - a23 * a34 * a41 - a24 * a31 * a43;
    r.b[1][1] = a11 * a33 * a44 + a13 * a34 * a41 + a14 * a31 * a43 - a11 * a34 * a43 - a13 * a31 * a44 - a14 * a33 * a41;
    r.b[1][2] = a11 * a24 * a43 + a13 * a21 * a44 + a14 * a23 * a41 - a11 * a23 * a44 - a13 * a24 * a41 - a14 * a21 * a43;
    r.b[1][3] = a11 * a23 * a34 + a13 * a24 * a31 + a14 * a21 * a33 - a11 * a24 * a33 - a13 * a21 * a34 - a14 * a23 * a31;
    r.b[2][0] = a21 * a32 * a44 + a22 * a34 * a41 + a24 * a31 * a42 - a21 * a34 * a42 - a22 * a31 * a44 - a24 * a32 * a41;
    r.b[2][1] = a11 * a34 * a42 + a12 * a31 * a44 + a14 * a32 * a41 - a11 * a32 * a44 - a12 * a34 * a41 - a14 * a31 * a42;
    r.b[2][2] = a11 * a22 * a44 + a12 * a24 * a41 + a14 * a21 * a42 - a11 * a24 * a42 - a12 * a21 * a44 - a14 * a22 * a41;
    r.b[2][3] = a11 * a24 * a32 + a12 * a21 * a34 + a14 * a22 * a31 - a11 * a22 * a34 - a12 * a24 * a31 - a14 * a21 * a32;
    r.b[3][0] = a21 * a33 * a42 + a22 * a31 * a43 + a23 * a32 * a41 - a21 * a32 * a43 - a22 * a33 * a41 - a23 * a31 * a42;
    r.b[3][1] = a11 * a32 * a43 + a12 * a33 * a41 + a13 * a31 * a42 - a11 * a33 * a42 - a12 * a31 * a43 - a13 * a32 * a41;
    r.b[3][2] = a11 * a23 * a42 + a12 * a21 * a43 + a13 * a22 * a41 - a11 * a22 * a43 - a12 * a23 * a41 - a13 * a21 * a42;
    r.b[3][3] = a11 * a22 * a33 + a12 * a23 * a31 + a13 * a21 * a32 - a11 * a23 * a32 - a12 * a21 * a33 - a13 * a22 * a31;
    return r;
}

static inline void bil_rows(const double x[4], const double y[4], double a[16]) {
    for (int l = 0; l < 4; l++) {
        a[l * 4] = 1;
        a[l * 4 + 1] = x[l];
        a[l * 4 + 2] = y[l];
        a[l * 4 + 3] = x[l] * y[l];
    }
}

// per-component geometry of the staggered grid
struct Comp {
    const double *xs, *ys;
    int stride, off, ni, nj;
};

// interpolateVelocityToGhostNodeX/Y (biLinearInterpolation.cu:14-353) over every ghost
// node; reads see `f` as before the call.
void Solver::gn_velocity(std::vector<double> &f, bool set) {
    const Comp comps[2] = {{G.xu.data(), G.y.data(), nx - 1, 0, nx - 1, ny},
                           {G.x.data(), G.yv.data(), nx, numU, nx, ny - 1}};
    const double ubody[2] = {B.uB[0], B.vB[0]};
    for (int c = 0; c < 2; c++) {
        const Comp &C = comps[c];
        std::vector<std::pair<int, double>> out;
        for (int J = 0; J < C.nj; J++)
            for (int I = 0; I < C.ni; I++) {
                const int g = J * C.stride + I + C.off;
                if (ghostTagsUV[g] <= 0) continue;
                int ii = I - 5 < 1 ? 1 : I - 5, jj = J - 5 < 1 ? 1 : J - 5;
                while (ii < C.ni - 1 && C.xs[ii] < ip_x[g]) ii++;
                while (jj < C.nj - 1 && C.ys[jj] < ip_y[g]) jj++;
                double x[4] = {C.xs[ii - 1], C.xs[ii], C.xs[ii - 1], C.xs[ii]};
                double y[4] = {C.ys[jj - 1], C.ys[jj - 1], C.ys[jj], C.ys[jj]};
                int index[4] = {(jj - 1) * C.stride + ii - 1 + C.off, (jj - 1) * C.stride + ii + C.off,
                                jj * C.stride + ii - 1 + C.off, jj * C.stride + ii + C.off};
                double q[4] = {f[index[0]], f[index[1]], f[index[2]], f[index[3]]};
                const double ubg = node_ub.empty() ? ubody[c] : node_ub[g];
                double mn = 1.0;
                int close_index = -1;
                bool inflag = false;
                for (int l = 0; l < 4; l++) {
                    const double s = std::sqrt(sq(x[l] - bi_x[g]) + sq(y[l] - bi_y[g]));
                    if (s < mn) { mn = s; close_index = index[l]; }
                    if (ghostTagsUV[index[l]] > 0) inflag = true;
                }
                for (int l = 0; l < 4; l++) {
                    if (ghostTagsUV[index[l]] == g) {
                        x[l] = bi_x[index[l]];
                        y[l] = bi_y[index[l]];
                        q[l] = ubg;
                    } else if (index[l] == close_index && !inflag) {
                        x[l] = bi_x[g];
                        y[l] = bi_y[g];
                        q[l] = ubg;
                    }
                }
                double a[16];
                bil_rows(x, y, a);
                const Bil B4 = bil_inverse(a);
                const double d = B4.det;
                double av[4];
                for (int r = 0; r < 4; r++)
                    av[r] = B4.b[r][0] / d * q[0] + B4.b[r][1] / d * q[1] + B4.b[r][2] / d * q[2] + B4.b[r][3] / d * q[3];
                const double X = ip_x[g], Y = ip_y[g];
                for (int l = 0; l < 4; l++) {
                    qcoef[l][g] = (B4.b[0][l] + B4.b[1][l] * X + B4.b[2][l] * Y + B4.b[3][l] * X * Y) / d;
                    qidx[l][g] = index[l];
                    qval[l][g] = q[l];
                }
                image_point_u[g] = av[0] + av[1] * X + av[2] * Y + av[3] * X * Y;
                if (set) out.push_back({g, 2 * ubg - image_point_u[g]});
            }
        for (auto &e : out) f[e.first] = e.second;
    }
}

// setInsideVelocity (luo_base/kernels/intermediateVelocity.cu:18-31); the kernel's
// index map covers v only for I < nx-1 and stops at the end of the array
void Solver::set_inside(std::vector<double> &f) {
    for (int i = 0; i < numU; i++) {
        const int I = i % (nx - 1), J = i / (nx - 1);
        const int iu = J * (nx - 1) + I, iv = J * nx + I + (nx - 1) * ny;
        const double bu = node_ub.empty() ? B.uB[0] : node_ub[iu];
        f[iu] = (ghostTagsUV[iu] != 0) * f[iu] + (ghostTagsUV[iu] == 0) * bu;
        if (iv < numUV) {
            const double bv = node_ub.empty() ? B.vB[0] : node_ub[iv];
            f[iv] = (ghostTagsUV[iv] != 0) * f[iv] + (ghostTagsUV[iv] == 0) * bv;
        }
    }
}

// Per-point body velocities, a build extension with no reference counterpart (the
// reference gives every point uB[0], biLinearInterpolation.cu:87,181): a ghost or hybrid
// node takes the value interpolated linearly along the body segment (k, k+1) whose
// nearest point to the node's body intercept is closest (the first such segment on a
// tie), and a node inside the body (tag 0) the rigid rotation omega about the mean of the
// body points, u = -omega (y - yc), v = omega (x - xc), at the node's position. The
// solid nodes of u are updated at once.
void Solver::set_body_velocity(const double *ub, const double *vb, double omega) {
    const int np = B.n;
    node_ub.assign(numUV, 0.0);
    double cx = 0, cy = 0;
    for (int k = 0; k < np; k++) {
        cx += B.x[k];
        cy += B.y[k];
    }
    cx /= np;
    cy /= np;
    for (int i = 0; i < numUV; i++) {
        const bool isU = i < numU;
        if (ghostTagsUV[i] > 0 || hybridTagsUV[i] > 0) {
            const double px = bi_x[i], py = bi_y[i];
            double best = 1e300, wbest = 0;
            int kbest = 0;
            for (int k = 0; k < np; k++) {
                const int l = (k + 1) % np;
                const double ex = B.x[l] - B.x[k], ey = B.y[l] - B.y[k];
                const double len2 = ex * ex + ey * ey;
                double t = len2 > 0 ? ((px - B.x[k]) * ex + (py - B.y[k]) * ey) / len2 : 0.0;
                t = std::min(1.0, std::max(0.0, t));
                const double d = sq(B.x[k] + t * ex - px) + sq(B.y[k] + t * ey - py);
                if (d < best) {
                    best = d;
                    kbest = k;
                    wbest = t;
                }
            }
            const double *c = isU ? ub : vb;
            node_ub[i] = (1 - wbest) * c[kbest] + wbest * c[(kbest + 1) % np];
        } else if (ghostTagsUV[i] == 0) {
            if (isU) node_ub[i] = -omega * (G.y[i / (nx - 1)] - cy);
            else node_ub[i] = omega * (G.x[(i - numU) % nx] - cx);
        }
    }
    set_inside(u);
}

// interpolateVelocityToHybridNodeX/Y (biLinearInterpolation.cu:356-668)
void Solver::hn_velocity(const std::vector<double> &f) {
    const Comp comps[2] = {{G.xu.data(), G.y.data(), nx - 1, 0, nx - 1, ny},
                           {G.x.data(), G.yv.data(), nx, numU, nx, ny - 1}};
    const double ubody[2] = {B.uB[0], B.vB[0]};
    for (int c = 0; c < 2; c++) {
        const Comp &C = comps[c];
        for (int J = 0; J < C.nj; J++)
            for (int I = 0; I < C.ni; I++) {
                const int h = J * C.stride + I + C.off;
                if (hybridTagsUV[h] <= 0) continue;
                int ii = I - 5 < 1 ? 1 : I - 5, jj = J - 5 < 1 ? 1 : J - 5;
                while (ii < C.ni - 1 && C.xs[ii] < ip_x[h]) ii++;
                while (jj < C.nj - 1 && C.ys[jj] < ip_y[h]) jj++;
                double x[4] = {C.xs[ii - 1], C.xs[ii], C.xs[ii - 1], C.xs[ii]};
                double y[4] = {C.ys[jj - 1], C.ys[jj - 1], C.ys[jj], C.ys[jj]};
                int index[4] = {(jj - 1) * C.stride + ii - 1 + C.off, (jj - 1) * C.stride + ii + C.off,
                                jj * C.stride + ii - 1 + C.off, jj * C.stride + ii + C.off};
                double q[4];
                for (int l = 0; l < 4; l++) q[l] = f[index[l]];
                for (int l = 0; l < 4; l++)
                    if (hybridTagsUV[index[l]] == h) {
                        x[l] = bi_x[h];
                        y[l] = bi_y[h];
                        q[l] = node_ub.empty() ? ubody[c] : node_ub[h];
                    }
                double a[16];
                bil_rows(x, y, a);
                const Bil B4 = bil_inverse(a);
                const double d = B4.det;
                double av[4];
                for (int r = 0; r < 4; r++)
                    av[r] = (B4.b[r][0] * q[0] + B4.b[r][1] * q[1] + B4.b[r][2] * q[2] + B4.b[r][3] * q[3]) / d;
                const double X = C.xs[I], Y = C.ys[J];
                for (int l = 0; l < 4; l++) {
                    qcoef[l][h] = (B4.b[0][l] + B4.b[1][l] * X + B4.b[2][l] * Y + B4.b[3][l] * X * Y) / d;
                    qidx[l][h] = index[l];
                    qval[l][h] = q[l];
                }
                ustar[h] = av[0] + av[1] * X + av[2] * Y + av[3] * Y * X;
                image_point_u[h] = av[0] + av[1] * ip_x[h] + av[2] * ip_y[h] + av[3] * ip_x[h] * ip_y[h];
            }
    }
}

// weightX / weightY (luoIBM/kernels/weight.cu:11-92)
void Solver::weight_uhat() {
    const Comp comps[2] = {{G.xu.data(), G.y.data(), nx - 1, 0, nx - 1, ny},
                           {G.x.data(), G.yv.data(), nx, numU, nx, ny - 1}};
    for (int c = 0; c < 2; c++) {
        const Comp &C = comps[c];
        const double *xs = C.xs, *ys = C.ys;
        const int st = C.stride;
        for (int J = 0; J < C.nj; J++)
            for (int I = 0; I < C.ni; I++) {
                const int h = J * st + I + C.off;
                if (hybridTagsUV[h] <= 0) continue;
                double d1 = 0, d2 = 0;
                const double dxl = xs[I] - xs[I - 1], dyl = ys[J] - ys[J - 1];
                if (ghostTagsUV[h - 1] > 0)
                    d1 = std::sqrt(sq(bi_x[h - 1] - xs[I - 1]) + sq((bi_y[h - 1] - ys[J])));
                else if (ghostTagsUV[h + 1] > 0)
                    d1 = std::sqrt(sq((bi_x[h + 1] - xs[I + 1])) + sq((bi_y[h + 1] - ys[J])));
                if (ghostTagsUV[h - st] > 0)
                    d2 = std::sqrt(sq(bi_x[h - st] - xs[I]) + sq(bi_y[h - st] - ys[J - 1]));
                if (ghostTagsUV[h + st] > 0)
                    d2 = std::sqrt(sq(bi_x[h + st] - xs[I]) + sq(bi_y[h + st] - ys[J + 1]));
                const double alpha = std::sqrt(sq(d1 / dxl) + sq(d2 / dyl));
                uhat[h] = (1 - alpha) * uhat[h] + alpha * ustar[h];
            }
    }
}

// interpolatePressureToHybridNode (biLinearInterpolation.cu:671-856)
void Solver::hn_pressure() {
    const double *xv = G.x.data(), *yu = G.y.data();
    const double dt = P.dt;
    for (int J = 0; J < ny; J++)
        for (int I = 0; I < nx; I++) {
            const int ip = J * nx + I;
            if (hybridTagsP[ip] <= 0) continue;
            int ii = I - 5 < 1 ? 1 : I - 5, jj = J - 5 < 1 ? 1 : J - 5;
            while (ii < nx - 1 && xv[ii] < ipp_x[ip]) ii++;
            while (jj < ny - 1 && yu[jj] < ipp_y[ip]) jj++;
            double x[4] = {xv[ii - 1], xv[ii], xv[ii - 1], xv[ii]};
            double y[4] = {yu[jj - 1], yu[jj - 1], yu[jj], yu[jj]};
            int index[4] = {(jj - 1) * nx + ii - 1, (jj - 1) * nx + ii, jj * nx + ii - 1, jj * nx + ii};
            double q[4];
            for (int l = 0; l < 4; l++) q[l] = p[index[l]];
            double a[16];
            bil_rows(x, y, a);
            for (int l = 0; l < 4; l++)
                if (hybridTagsP[index[l]] == ip) {
                    const double dudt = (B.uB[0] - B.uBk[0]) / dt, dvdt = (B.vB[0] - B.vBk[0]) / dt;
                    x[l] = bip_x[ip];
                    y[l] = bip_y[ip];
                    const double n_x = ipp_x[ip] - x[l], n_y = ipp_y[ip] - y[l];
                    const double nl = std::sqrt(n_x * n_x + n_y * n_y);
                    q[l] = -(n_x / nl * dudt + n_y / nl * dvdt);
                    a[l * 4] = 0;
                    a[l * 4 + 1] = n_x / nl;
                    a[l * 4 + 2] = n_y / nl;
                    a[l * 4 + 3] = n_y / nl * x[l] + n_x / nl * y[l];
                }
            const Bil B4 = bil_inverse(a);
            const double d = B4.det;
            double av[4];
            for (int r = 0; r < 4; r++)
                av[r] = B4.b[r][0] / d * q[0] + B4.b[r][1] / d * q[1] + B4.b[r][2] / d * q[2] + B4.b[r][3] / d * q[3];
            pstar[ip] = av[0] + av[1] * xv[I] + av[2] * yu[J] + av[3] * xv[I] * yu[J];
            // biLinearInterpolation.cu:775-785, 850-853
            for (int l = 0; l < 4; l++) {
                pqcoef[l][ip] = (B4.b[0][l] + B4.b[1][l] * xv[I] + B4.b[2][l] * yu[J] + B4.b[3][l] * xv[I] * yu[J]) / d;
                pqidx[l][ip] = index[l];
                pqval[l][ip] = q[l];
            }
        }
}

// weightP (luoIBM/kernels/weight.cu:95-134)
void Solver::weight_p() {
    const double *xv = G.x.data(), *yu = G.y.data();
    for (int J = 0; J < ny; J++)
        for (int I = 0; I < nx; I++) {
            const int ip = J * nx + I;
            if (hybridTagsP[ip] <= 0) continue;
            double d1 = 0, d2 = 0;
            const double dxl = xv[I] - xv[I - 1], dyl = yu[J] - yu[J - 1];
            if (ghostTagsP[ip - 1] > 0)
                d1 = std::sqrt(sq(bip_x[ip - 1] - xv[I - 1]) + sq(bip_y[ip - 1] - yu[J]));
            else if (ghostTagsP[ip + 1] > 0)
                d1 = std::sqrt(sq(bip_x[ip + 1] - xv[I + 1]) + sq(bip_y[ip + 1] - yu[J]));
            if (ghostTagsP[ip - nx] > 0)
                d2 = std::sqrt(sq(bip_x[ip - nx] - xv[I]) + sq(bip_y[ip - nx] - yu[J - 1]));
            if (ghostTagsP[ip + nx] > 0)
                d2 = std::sqrt(sq(bip_x[ip + nx] - xv[I]) + sq(bip_y[ip + nx] - yu[J + 1]));
            const double alpha = std::sqrt(sq(d1 / dxl) + sq(d2 / dyl));
            p[ip] = (1 - alpha) * p[ip] + alpha * pstar[ip];
        }
}

// interpolatePressureToGhostNode (biLinearInterpolation.cu:859-1066)
void Solver::gn_pressure(bool set) {
    const double *xv = G.x.data(), *yu = G.y.data();
    const double dt = P.dt;
    for (int J = 0; J < ny; J++)
        for (int I = 0; I < nx; I++) {
            const int ip = J * nx + I;
            if (ghostTagsP[ip] <= 0) continue;
            int ii = I - 5 < 1 ? 1 : I - 5, jj = J - 5 < 1 ? 1 : J - 5;
            while (ii < nx - 1 && xv[ii] < ipp_x[ip]) ii++;
            while (jj < ny - 1 && yu[jj] < ipp_y[ip]) jj++;
            double x[4] = {xv[ii - 1], xv[ii], xv[ii - 1], xv[ii]};
            double y[4] = {yu[jj - 1], yu[jj - 1], yu[jj], yu[jj]};
            int index[4] = {(jj - 1) * nx + ii - 1, (jj - 1) * nx + ii, jj * nx + ii - 1, jj * nx + ii};
            double q[4] = {p[index[0]], p[index[1]], p[index[2]], p[index[3]]};
            double a[16];
            bil_rows(x, y, a);
            double mn = 1.0, matDClose = 0;
            int close_index = -1;
            for (int l = 0; l < 4; l++) {
                const double s = std::sqrt(sq(x[l] - bip_x[ip]) + sq(y[l] - bip_y[ip]));
                if (s < mn) { mn = s; close_index = index[l]; }
            }
            for (int l = 0; l < 4; l++) {
                const double dudt = (B.uB[0] - B.uBk[0]) / dt, dvdt = (B.vB[0] - B.vBk[0]) / dt;
                if (ghostTagsP[index[l]] > 0) {
                    x[l] = bip_x[index[l]];
                    y[l] = bip_y[index[l]];
                    const double n_x = ipp_x[index[l]] - x[l], n_y = ipp_y[index[l]] - y[l];
                    const double nl = std::sqrt(n_x * n_x + n_y * n_y);
                    q[l] = -(n_x / nl * dudt + n_y / nl * dvdt);
                    a[l * 4] = 0;
                    a[l * 4 + 1] = n_x / nl;
                    a[l * 4 + 2] = n_y / nl;
                    a[l * 4 + 3] = n_y / nl * x[l] + n_x / nl * y[l];
                }
                if (index[l] == close_index) {
                    x[l] = bip_x[ip];
                    y[l] = bip_y[ip];
                    const double n_x = ipp_x[ip] - x[l], n_y = ipp_y[ip] - y[l];
                    const double nl = std::sqrt(n_x * n_x + n_y * n_y);
                    matDClose = (n_x / nl * dudt + n_y / nl * dvdt);
                }
            }
            const Bil B4 = bil_inverse(a);
            const double d = B4.det;
            double av[4];
            for (int r = 0; r < 4; r++)
                av[r] = B4.b[r][0] / d * q[0] + B4.b[r][1] / d * q[1] + B4.b[r][2] / d * q[2] + B4.b[r][3] / d * q[3];
            const double X = ipp_x[ip], Y = ipp_y[ip];
            const double ipp = av[0] + av[1] * X + av[2] * Y + av[3] * Y * X;
            dpdn[ip] = std::sqrt(sq(X - xv[I]) + sq(Y - yu[J])) * matDClose;
            // biLinearInterpolation.cu:972-982, 1053-1056
            for (int l = 0; l < 4; l++) {
                pqcoef[l][ip] = (B4.b[0][l] + B4.b[1][l] * X + B4.b[2][l] * Y + B4.b[3][l] * X * Y) / d;
                pqidx[l][ip] = index[l];
                pqval[l][ip] = q[l];
            }
            if (set) p[ip] = ipp + dpdn[ip];
        }
}

// updateBoundaryX/Y (FadlunModified/kernels/intermediateVelocity.cu:19-41), with the
// hard-coded Uinf = Vinf = 1 of luo_base/intermediateVelocity.cu:13-14
void Solver::robin() {
    double *xp = bc[XPLUS].data();
    const double dt = P.dt, Uinf = 1, Vinf = 1;
    for (int J = 0; J < ny; J++) {
        const int I = nx - 1, i = J * (nx - 1) + I;
        const double beta = Uinf * dt / G.dx[I];
        xp[J] = xp[J] * (1 - beta) + beta * u[i - 1];
    }
    for (int J = 0; J < ny - 1; J++) {
        const int I = nx, i = J * nx + I;
        const double beta = Vinf * dt / G.dx[I - 1];
        xp[J + ny] = xp[J + ny] * (1 - beta) + beta * u[i + numU - 1];
    }
}

// updateRHS1forIBX/Y (FadlunModified/kernels/intermediateVelocity.cu:44-63)
void Solver::fadlun_rhs1() {
    for (int i = 0; i < numUV; i++)
        rhs1[i] = (hybridTagsUV[i] == -1) * (ghostTagsUV[i] <= 0) * (rhs1[i]) +
                  (hybridTagsUV[i] != -1) * dist_between[i] / (dist_int[i] + dist_between[i]) * uv[i];
}

// FadlunModified/kernels/intermediatePressure.cu:17-138
void Solver::fadlun_rhs2() {
    const double *dx = G.dx.data(), *dy = G.dy.data();
    const double *xm = bc[XMINUS].data(), *xp = bc[XPLUS].data();
    const double *ym = bc[YMINUS].data(), *yp = bc[YPLUS].data();
    for (int ip = 0; ip < numP; ip++) {
        const int I = ip % nx, J = ip / nx, iu = (nx - 1) * J + I, iv = (nx - 1) * ny + nx * J + I;
        double t = 0;
        if (hybridTagsP[ip] != -1) {
            const double du = dist_u_body[ip], dv = dist_v_body[ip];
            if (ghostTagsP[ip + 1] == -1) {
                if (du > dx[I] / 2 && du < dx[I]) t -= uhat[iu] / du;
                else t -= uhat[iu] / dx[I];
            }
            if (ghostTagsP[ip - 1] == -1) {
                if (du > dx[I] / 2 && du < dx[I]) t += uhat[iu - 1] / du;
                else t += uhat[iu - 1] / dx[I];
            }
            if (ghostTagsP[ip + nx] == -1) {
                if (dv > dy[J] / 2 && dv < dy[J]) t -= uhat[iv] / dv;
                else t -= uhat[iv] / dy[J];
            }
            if (ghostTagsP[ip - nx] == -1) {
                if (dv > dy[J] / 2 && dv < dy[J]) t += uhat[iv - nx] / dv;
                else t += uhat[iv - nx] / dy[J];
            }
        } else if (ghostTagsP[ip] > 0) {
            if (ghostTagsP[ip + 1] == 0) t -= uhat[iu] / dx[I];
            if (ghostTagsP[ip - 1] == 0) t += uhat[iu - 1] / dx[I];
            if (ghostTagsP[ip + nx] == 0) t -= uhat[iv] / dy[J];
            if (ghostTagsP[ip - nx] == 0) t += uhat[iv - nx] / dy[J];
        } else {
            if (I != nx - 1) t -= uhat[iu] / dx[I];
            else t -= xp[J] / dx[I];
            if (I != 0) t += uhat[iu - 1] / dx[I];
            else t += xm[J] / dx[I];
            if (J != ny - 1) t -= uhat[iv] / dy[J];
            else t -= yp[(nx - 1) + I] / dy[J];
            if (J != 0) t += uhat[iv - nx] / dy[J];
            else t += ym[(nx - 1) + I] / dy[J];
        }
        rhs2[ip] = t;
    }
}

// FadlunModified/kernels/projectVelocity.cu:14-44 (no uold update)
void Solver::fadlun_project() {
    const double *dx = G.dx.data(), *dy = G.dy.data();
    const double dt = P.dt;
    for (int i = 0; i < numU; i++) {
        const int I = i % (nx - 1), J = i / (nx - 1), ip = J * nx + I;
        u[i] = uhat[i] - (ghostTagsP[ip + 1] == -1 && ghostTagsP[ip] == -1) * dt * (p[ip + 1] - p[ip]) /
                             (0.5 * dx[I + 1] + 0.5 * dx[I]);
    }
    for (int k = 0; k < numV; k++) {
        const int I = k % nx, J = k / nx, ip = J * nx + I, i = k + numU;
        u[i] = uhat[i] - (ghostTagsP[ip + nx] == -1 && ghostTagsP[ip] == -1) * dt * (p[ip + nx] - p[ip]) /
                             (0.5 * dy[J + 1] + 0.5 * dy[J]);
    }
}

// luo_base/calculateForce.cu:21-74 with FadlunModified/kernels/calculateForce.cu:32-283
void Solver::force() {
    const double *dx = G.dx.data(), *dy = G.dy.data();
    const double nu = P.nu, dt = P.dt;
    const int I = B.startI, J = B.startJ, ncx = B.numCellsX, ncy = B.numCellsY;
    const double *uu = u.data(), *pp = p.data();
    const int *tagsIn = ghostTagsUV.data();
    double sx = 0, sy = 0, su = 0;
    for (int idx = 0; idx < ncy; idx++) {  // dragLeftRight
        const int Ip = (J + idx) * nx + I, Iu = (J + idx) * (nx - 1) + I;
        sx += -((pp[Ip + ncx] - pp[Ip]) * dy[J + idx]
                + ((uu[Iu + ncx] + uu[Iu + ncx - 1]) * (uu[Iu + ncx] + uu[Iu + ncx - 1]) / 4
                   - (uu[Iu - 1] + uu[Iu]) * (uu[Iu - 1] + uu[Iu]) / 4) * dy[J + idx]
                - nu * ((uu[Iu + ncx] - uu[Iu + ncx - 1]) / dx[I + ncx] - (uu[Iu] - uu[Iu - 1]) / dx[I]) * dy[J + idx]);
    }
    for (int idx = 0; idx <= ncx; idx++) {  // dragBottomTop
        const int Iu = J * (nx - 1) + (I - 1 + idx), Iv = (nx - 1) * ny + (J - 1) * nx + I + idx;
        sy += -((0.25 * (uu[Iu + ncy * (nx - 1)] + uu[Iu + (ncy - 1) * (nx - 1)]) * (uu[Iv + ncy * nx] + uu[Iv + ncy * nx - 1]) -
                          0.25 * (uu[Iu] + uu[Iu - (nx - 1)]) * (uu[Iv] + uu[Iv - 1]))
                - nu * (((uu[Iu + ncy * (nx - 1)] - uu[Iu + (ncy - 1) * (nx - 1)]) / 2.0 / (dy[J + ncy] + dy[J + ncy - 1]) +
                         (uu[Iv + ncy * nx] - uu[Iv + ncy * nx - 1]) / 2.0 / (dx[I + idx] + dx[I + idx - 1])) -
                        ((uu[Iu] - uu[Iu - (nx - 1)]) / 2.0 / (dy[J] + dy[J - 1]) +
                         (uu[Iv] - uu[Iv - 1]) / 2.0 / (dx[I + idx] + dx[I + idx - 1])))) *
              0.5 * (dx[I + idx] + dx[I + idx - 1]);
    }
    for (int idx = 0; idx < (ncx + 1) * ncy; idx++) {  // dragUnsteady
        const int i = idx % (ncx + 1), j = idx / (ncx + 1);
        const int Iu = (J + j) * (nx - 1) + (I - 1 + i);
        su += -(tagsIn[Iu] == -1) * ((uu[Iu] * dy[J + j] - uold[Iu] * dy[J + j]) / dt * 0.5 * (dx[I + i] + dx[I - 1 + i]));
    }
    fxx = sx;
    fxy = sy;
    fxu = su;
    B.forceX = fxx + fxy + fxu;
    double lx = 0, ly = 0, lu = 0;
    for (int idx = 0; idx <= ncy; idx++) {  // liftLeftRight
        const int Iu = (J + idx) * (nx - 1) + (I - 1), Iv = (nx - 1) * ny + (J - 1 + idx) * nx + I;
        lx += -((0.25 * (uu[Iu + ncx] + uu[Iu + ncx - (nx - 1)]) * (uu[Iv + ncx] + uu[Iv + ncx - 1]) -
                          0.25 * (uu[Iu] + uu[Iu - (nx - 1)]) * (uu[Iv] + uu[Iv - 1]))
                - nu * (((uu[Iu + ncx] - uu[Iu + ncx - (nx - 1)]) / 2.0 / (dy[J + idx] + dy[J - 1 + idx]) +
                         (uu[Iv + ncx] - uu[Iv + ncx - 1]) / 2.0 / (dx[I + ncx] + dx[I + ncx - 1])) -
                        ((uu[Iu] - uu[Iu - (nx - 1)]) / 2.0 / (dy[J + idx] + dy[J - 1 + idx]) +
                         (uu[Iv] - uu[Iv - 1]) / 2.0 / (dx[I] + dx[I - 1])))) *
              0.5 * (dy[J + idx] + dy[J - 1 + idx]);
    }
    for (int idx = 0; idx < ncx; idx++) {  // liftBottomTop
        const int Ip = J * nx + I + idx, Iv = (nx - 1) * ny + (J - 1) * nx + I + idx;
        ly += -((pp[Ip + ncy * nx] - pp[Ip - nx]) * dx[I + idx]
                + (0.25 * (uu[Iv + (ncy + 1) * nx] + uu[Iv + ncy * nx]) * (uu[Iv + (ncy + 1) * nx] + uu[Iv + ncy * nx])
                   - 0.25 * (uu[Iv] + uu[Iv - nx]) * (uu[Iv] + uu[Iv - nx])) * dx[I + idx]
                - nu * ((uu[Iv + (ncy + 1) * nx] - uu[Iv + ncy * nx]) * dx[I + idx] / dy[J + ncy]
                        - (uu[Iv] - uu[Iv - nx]) * dx[I] / dy[J - 1]));
    }
    for (int idx = 0; idx < ncx * (ncy + 1); idx++) {  // liftUnsteady
        const int i = idx % ncx, j = idx / ncx;
        const int Iv = (J - 1 + j) * nx + (I + i) + (nx - 1) * ny;
        lu += -(tagsIn[Iv] == -1) * ((uu[Iv] * dx[I + i] - uold[Iv] * dx[I + i]) / dt * 0.5 * (dy[J + j] + dy[J - 1 + j]));
    }
    B.forceY = lx + ly + lu;
}

}  // namespace orc
