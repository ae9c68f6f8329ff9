// Immersed-boundary kernels (Luo external / modified Fadlun), FP64.
//
// Tagging restates tag_u_luo / tag_v_luo / tag_p_luo and the zero_* interior fills
// (src/solvers/NavierStokes/luo_base/kernels/tagPoints.cu:14-1298). The reference
// threads write their neighbours' ghost tags and intercept data, so several threads
// can write one entry in an unspecified order. Here every box thread runs the
// segment loop once and records, per target (itself and its 4 neighbours), only the
// last write it would make; a second, gather-style kernel then resolves each entry
// as if the threads had run one after another in increasing index order (all
// per-thread initialisation writes first). Tags are therefore deterministic.
//
// Interpolation (luo_base/kernels/biLinearInterpolation.cu:14-1066) is split into a
// set-up kernel that turns each ghost/hybrid node's bilinear (or Neumann-augmented)
// 4x4 system into four corner weights, and per-step apply kernels that are pure
// gathers over compact node lists. Ghost-node velocity updates are staged so that
// every read sees the field as it was before the update.
#include <algorithm>

#include "batch.hpp"
#include "ibm.hpp"
#include "krylov.hpp"

namespace cuibm {

namespace {

constexpr double EPS = 1.e-10;
enum { DIR_W = 0, DIR_E = 1, DIR_S = 2, DIR_N = 3 };
enum { T_GHOST = 0, T_C1 = 1, T_C2 = 2, T_C3 = 3, T_C4 = 4, T_HYBP = 5, T_HYB = 6 };
enum { F_FINAL = 1, F_D0 = 2, F_D1 = 4, F_GHOSTP = 8, F_HYBP = 16 };

int grid_for(long n) {
    const long b = (n + BS - 1) / BS;
    return (int)(b < 65536 ? b : 65536);
}

// squares are products: pow(x, 2) in the reference
__device__ __forceinline__ double sq(double x) { return x * x; }

// foot of the perpendicular from (px,py) to the body line through points k,l
// (tagPoints.cu:122-133). The reference forms the distance along the line as
// a / tan(theta) with theta = acos(c), a = b sin(theta), c = (p^2 + b^2 - o^2) / (2 p b);
// that is b cos(theta) = (p^2 + b^2 - o^2) / (2 p), evaluated here in that form so that
// only correctly rounded operations are involved and a CPU evaluation agrees bit for bit.
__device__ void foot(const double *bx, const double *by, int k, int l, double px, double py, double &fx, double &fy) {
    const double p = sqrt(sq((bx[l] - bx[k])) + sq((by[l] - by[k])));
    const double o = sqrt(sq((bx[l] - px)) + sq((by[l] - py)));
    const double b = sqrt(sq((px - bx[k])) + sq((py - by[k])));
    const double d = (p * p + b * b - o * o) / (2 * p);
    fy = by[k] + (by[l] - by[k]) / p * d;
    fx = bx[k] + (bx[l] - bx[k]) / p * d;
}

__device__ __forceinline__ int seg_k(int l, int n) { return l == 0 ? n - 1 : l - 1; }

// BI / IP written by a record `code` for target node (tx,ty); writer node (wx.., wy..)
__device__ void eval_code(int code, const double *bx, const double *by, int nb, double tx, double ty, double wx,
                          double wy, double wxm, double wxp, double wym, double wyp, double &bix, double &biy,
                          double &ipx, double &ipy) {
    const int type = code & 7, l = code >> 3;
    double fx, fy;
    switch (type) {
        case T_GHOST:
            foot(bx, by, seg_k(l, nb), l, tx, ty, fx, fy);
            bix = fx; biy = fy;
            ipy = fy + (fy - ty);
            ipx = fx + (fx - tx);
            break;
        case T_HYB:
        case T_HYBP:
            foot(bx, by, seg_k(l, nb), l, tx, ty, fx, fy);
            bix = fx; biy = fy;
            ipy = 2 * ty - fy;
            ipx = 2 * tx - fx;
            break;
        case T_C1: bix = wx; biy = wy; ipx = wxm; ipy = wy; break;
        case T_C2: bix = wx; biy = wy; ipx = wxp; ipy = wy; break;
        case T_C3: bix = wx; biy = wy; ipx = wx; ipy = wym; break;
        case T_C4: bix = wx; biy = wy; ipx = wx; ipy = wyp; break;
    }
}

// ---- velocity tagging: one thread of tag_u_luo (VDIR=false) / tag_v_luo (VDIR=true)
template <bool VDIR>
__global__ void __launch_bounds__(BS) k_tag_vel(BodyDev B, CompDev c, BoxDev box, int nx, int *own, int *nb0, int *nb1,
                                                int *nb2, int *nb3, int *flags, double *dv0, double *dv1, TagOut o) {
    const int T = blockIdx.x * blockDim.x + threadIdx.x;
    if (T >= box.nthr) return;
    own[T] = -1; nb0[T] = -1; nb1[T] = -1; nb2[T] = -1; nb3[T] = -1; flags[T] = 0;
    const int I = box.i_start + T % box.w, J = box.j_start + T / box.w;
    if (J * c.stride + I >= c.ni * c.nj) return;
    // nodes on the outer grid line would read outside the coordinate arrays
    if (I < 1 || J < 1 || I >= c.ni - 1 || J >= c.nj - 1) return;
    const int st = c.stride, iu = J * st + I + c.off;
    const double *xs = c.xs, *ys = c.ys, *bx = B.x, *by = B.y;
    const double *ub = VDIR ? B.vB : B.uB;
    const int n = B.n;
    int k = n - 1, l = 0;
    bool outsideX = true, outsideY = true, flag = false;
    int bdryFlagX = -1, bdryFlagY = -1, bdryFlag2X = -1, bdryFlag2Y = -1;
    double uvX = 0, uvY = 0, Xa = 1, Ya = 1, Xb = 1, Yb = 1, x = 0, y = 0;
    int rown = -1, rnb[4] = {-1, -1, -1, -1}, fl = 0;
    double d0 = 0, d1 = 0;
    while (l < n && (VDIR || !flag)) {
        int bottom, top, left, right;
        if (by[k] > by[l]) { bottom = l; top = k; } else { bottom = k; top = l; }
        if (bx[k] > bx[l]) { left = l; right = k; } else { left = k; right = l; }
        if (by[bottom] - EPS < ys[J] && by[top] - EPS > ys[J] && (!VDIR || !flag)) {
            if (fabs(by[l] - by[k]) > EPS) {
                x = bx[k] + (bx[l] - bx[k]) * (ys[J] - by[k]) / (by[l] - by[k]);
                uvX = ub[k] + (ub[l] - ub[k]) * (ys[J] - by[k]) / (by[l] - by[k]);
                if (x > xs[I] + EPS) outsideX = !outsideX;
                if (x > xs[I - 1] + EPS && x < xs[I] - EPS) {
                    bdryFlagX = iu;
                    bdryFlag2X = iu + 1;
                    if (x > B.midX + EPS) { rnb[DIR_W] = l * 8 + T_GHOST; rown = l * 8 + T_HYB; }
                    Xa = xs[I] - x;
                    Xb = xs[I + 1] - xs[I];
                    if (!VDIR && x > B.midX) { d0 = Xa; fl |= F_D0; }
                } else if (x > xs[I] + EPS && x < xs[I + 1] - EPS) {
                    bdryFlagX = iu;
                    bdryFlag2X = iu - 1;
                    if (x < B.midX - EPS) { rnb[DIR_E] = l * 8 + T_GHOST; rown = l * 8 + T_HYB; }
                    Xa = x - xs[I];
                    Xb = xs[I] - xs[I - 1];
                    if (!VDIR && x < B.midX) { d1 = Xa; fl |= F_D1; }
                } else if (fabs(x - xs[I]) < EPS && fabs(B.midX - xs[I]) > fabs(B.midY - ys[J])) {
                    outsideX = true;
                    bdryFlagX = iu;
                    if (x < B.midX) {
                        bdryFlag2X = iu - 1;
                        rnb[DIR_E] = T_C1; rown = T_C1;
                        Xa = 0;
                        Xb = xs[I] - xs[I - 1];
                    }
                    if (x > B.midX) {
                        bdryFlag2X = iu + 1;
                        rnb[DIR_W] = T_C2; rown = T_C2;
                        Xa = 0;
                        Xb = xs[I + 1] - xs[I];
                    }
                    flag = true;
                }
            }
        }
        if (bx[left] - EPS < xs[I] && bx[right] - EPS > xs[I] && (!VDIR || !flag)) {
            if (fabs(bx[l] - bx[k]) > EPS) {
                y = by[k] + (by[l] - by[k]) * (xs[I] - bx[k]) / (bx[l] - bx[k]);
                uvY = ub[k] + (ub[l] - ub[k]) * (xs[I] - bx[k]) / (bx[l] - bx[k]);
                if (y > ys[J] + EPS) outsideY = !outsideY;
                if (y > ys[J - 1] + EPS && y < ys[J] - EPS) {
                    bdryFlagY = iu;
                    bdryFlag2Y = iu + st;
                    if (y > B.midY + EPS) { rnb[DIR_S] = l * 8 + T_GHOST; rown = l * 8 + T_HYB; }
                    Ya = ys[J] - y;
                    Yb = ys[J + 1] - ys[J];
                    if (VDIR && outsideY) { d0 = Ya; fl |= F_D0; }
                } else if (y > ys[J] + EPS && y < ys[J + 1] - EPS) {
                    bdryFlagY = iu;
                    bdryFlag2Y = iu - st;
                    if (y < B.midY - EPS) { rnb[DIR_N] = l * 8 + T_GHOST; rown = l * 8 + T_HYB; }
                    Ya = y - ys[J];
                    Yb = ys[J] - ys[J - 1];
                    if (VDIR && outsideY) { d1 = Ya; fl |= F_D1; }
                } else if (VDIR ? (fabs(y - ys[J]) < EPS && fabs(B.midY - ys[J]) > fabs(B.midX - xs[I]))
                                : (fabs(y - ys[J]) < EPS && fabs(y - B.midY) > fabs(x - B.midX))) {
                    outsideY = true;
                    bdryFlagY = iu;
                    if (y < B.midY) {
                        bdryFlag2X = iu - st;
                        rnb[DIR_N] = T_C3; rown = T_C3;
                        Ya = 0;
                        Yb = ys[J] - ys[J - 1];
                    }
                    if (y > B.midY) {
                        bdryFlag2X = iu + st;
                        rnb[DIR_S] = T_C4; rown = T_C4;
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
    int which = 0;
    if (!VDIR) {
        if (outsideX && bdryFlagX >= 0) which = 1;
        else if (outsideY && bdryFlagY >= 0) which = 2;
    } else {
        if (outsideY && bdryFlagY >= 0) which = 2;
        else if (outsideX && bdryFlagX >= 0) which = 1;
    }
    if (which) {
        fl |= F_FINAL;
        o.hybridUV[iu] = which == 1 ? bdryFlagX : bdryFlagY;
        o.hybridUV2[iu] = which == 1 ? bdryFlag2X : bdryFlag2Y;
        o.dist_int[iu] = which == 1 ? Xa : Ya;
        o.dist_between[iu] = which == 1 ? Xb : Yb;
        o.uv[iu] = which == 1 ? uvX : uvY;
    }
    own[T] = rown; nb0[T] = rnb[0]; nb1[T] = rnb[1]; nb2[T] = rnb[2]; nb3[T] = rnb[3]; flags[T] = fl;
    dv0[T] = d0; dv1[T] = d1;
}

// thread index of the box thread that owns node (I,J), or -1
__device__ __forceinline__ int box_thread(const BoxDev &box, int I, int J) {
    const int i = I - box.i_start, j = J - box.j_start;
    if (i < 0 || i >= box.w || j < 0) return -1;
    const int T = j * box.w + i;
    return T < box.nthr ? T : -1;
}

// resolve ghost tags and intercept data of one velocity component, and the
// distance_from_{u,v}_to_body entries, in writer-index order
template <bool VDIR>
__global__ void __launch_bounds__(BS) k_tag_vel_resolve(CompDev c, BoxDev box, BodyDev B, const int *own,
                                                        const int *nb0, const int *nb1, const int *nb2, const int *nb3,
                                                        const int *flags, const double *dv0, const double *dv1,
                                                        TagOut o, int nx, int ny) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncomp = c.ni * c.nj;
    if (t < ncomp) {
        const int I = t % c.ni, J = t / c.ni, m = t + c.off;
        // candidates in increasing writer index: (I,J-1)->N, (I-1,J)->E, self, (I+1,J)->W, (I,J+1)->S
        const int wI[5] = {I, I - 1, I, I + 1, I};
        const int wJ[5] = {J - 1, J, J, J, J + 1};
        int gt = o.ghostUV[m];
        int code = -1, cw = -1;
        for (int q = 0; q < 5; q++) {
            if (wI[q] < 0 || wI[q] >= c.ni || wJ[q] < 0 || wJ[q] >= c.nj) continue;
            const int T = box_thread(box, wI[q], wJ[q]);
            if (T < 0) continue;
            int cc;
            if (q == 2) {
                cc = own[T];
                if (flags[T] & F_FINAL) gt = -1;
            } else {
                cc = (q == 0) ? nb3[T] : (q == 1) ? nb1[T] : (q == 3) ? nb0[T] : nb2[T];
                if (cc != -1) gt = m;
            }
            if (cc != -1) { code = cc; cw = q; }
        }
        o.ghostUV[m] = gt;
        if (code != -1) {
            const int Iw = wI[cw], Jw = wJ[cw];
            double bix = 0, biy = 0, ipx = 0, ipy = 0;
            eval_code(code, B.x, B.y, B.n, c.xs[I], c.ys[J], c.xs[Iw], c.ys[Jw], Iw > 0 ? c.xs[Iw - 1] : 0.0,
                      Iw + 1 < c.ni ? c.xs[Iw + 1] : 0.0, Jw > 0 ? c.ys[Jw - 1] : 0.0, Jw + 1 < c.nj ? c.ys[Jw + 1] : 0.0,
                      bix, biy, ipx, ipy);
            o.bi_x[m] = bix; o.bi_y[m] = biy; o.ip_x[m] = ipx; o.ip_y[m] = ipy;
        }
    }
    // distance_from_{u,v}_to_body[ip] for pressure node (I,J)
    if (t < nx * ny) {
        const int I = t % nx, J = t / nx;
        double *d = VDIR ? o.dv : o.du;
        double val = d[t];
        bool any = false;
        if (!VDIR) {
            // writers: u-thread (I-1,J) "next", u-thread (I,J) "own"; (I,J) first zeroes it
            const int Ts = (I < c.ni && J < c.nj) ? box_thread(box, I, J) : -1;
            if (Ts >= 0) { val = 0; any = true; }
            const int Tw = (I - 1 >= 0 && I - 1 < c.ni && J < c.nj) ? box_thread(box, I - 1, J) : -1;
            if (Tw >= 0 && (flags[Tw] & F_D1)) { val = dv1[Tw]; any = true; }
            if (Ts >= 0 && (flags[Ts] & F_D0)) { val = dv0[Ts]; any = true; }
        } else {
            // writers: v-thread (I,J-1) "next", v-thread (I,J) "own"
            const int Tb = (J - 1 >= 0 && J - 1 < c.nj && I < c.ni) ? box_thread(box, I, J - 1) : -1;
            const int Ts = (J < c.nj && I < c.ni) ? box_thread(box, I, J) : -1;
            if (Tb >= 0 && (flags[Tb] & F_D1)) { val = dv1[Tb]; any = true; }
            if (Ts >= 0 && (flags[Ts] & F_D0)) { val = dv0[Ts]; any = true; }
        }
        if (any) d[t] = val;
    }
}

// ---- pressure tagging: one thread of tag_p_luo
__global__ void __launch_bounds__(BS) k_tag_p(BodyDev B, GridDev g, BoxDev box, int *own, int *nb0, int *nb1, int *nb2,
                                              int *nb3, int *flags) {
    const int T = blockIdx.x * blockDim.x + threadIdx.x;
    if (T >= box.nthr) return;
    own[T] = -1; nb0[T] = -1; nb1[T] = -1; nb2[T] = -1; nb3[T] = -1; flags[T] = 0;
    const int nx = g.nx, I = box.i_start + T % box.w, J = box.j_start + T / box.w;
    if (J * nx + I >= nx * g.ny) return;
    if (I < 1 || J < 1 || I >= nx - 1 || J >= g.ny - 1) return;
    const double *xv = g.x, *yu = g.y, *bx = B.x, *by = B.y;
    const double midX = B.midX, midY = B.midY;
    const int n = B.n;
    int k = n - 1, l = 0;
    bool flag = false;
    int rown = -1, rnb[4] = {-1, -1, -1, -1}, fl = 0;
    while (l < n) {
        int bottom, top, left, right;
        if (by[k] > by[l]) { bottom = l; top = k; } else { bottom = k; top = l; }
        if (bx[k] > bx[l]) { left = l; right = k; } else { left = k; right = l; }
        if (by[bottom] - EPS < yu[J] && by[top] - EPS > yu[J] && !flag) {
            if (fabs(by[l] - by[k]) > EPS) {
                const double x = bx[k] + (bx[l] - bx[k]) * (yu[J] - by[k]) / (by[l] - by[k]);
                int ev = -1;
                if (x > midX + EPS && x > xv[I] + EPS + EPS && x < xv[I + 1] - EPS) ev = T_GHOST;
                else if (x < midX + EPS && x < xv[I] - EPS && x > xv[I - 1] + EPS) ev = T_GHOST;
                if (x > midX + EPS && x < xv[I] - EPS && x > xv[I - 1] + EPS) ev = T_HYBP;
                else if (x < midX + EPS && x > xv[I] + EPS && x < xv[I + 1] - EPS) ev = T_HYBP;
                if (ev == T_GHOST) fl |= F_GHOSTP;
                if (ev == T_HYBP) fl |= F_HYBP;
                if (ev >= 0) rown = l * 8 + ev;
                if (fabs(x - xv[I]) < EPS && fabs(midX - xv[I]) > fabs(midY - yu[J])) {
                    fl |= F_HYBP;
                    if (x < midX) { rnb[DIR_E] = T_C1; rown = T_C1; }
                    else if (x > midX) { rnb[DIR_W] = T_C2; rown = T_C2; }
                    flag = true;
                }
            }
        }
        if (bx[left] - EPS < xv[I] && bx[right] - EPS > xv[I] && !flag) {
            if (fabs(bx[l] - bx[k]) > EPS) {
                const double y = by[k] + (by[l] - by[k]) * (xv[I] - bx[k]) / (bx[l] - bx[k]);
                int ev = -1;
                if (y > midY + EPS && y > yu[J] + EPS && y < yu[J + 1] - EPS) ev = T_GHOST;
                else if (y < midY + EPS && y < yu[J] - EPS && y > yu[J - 1] + EPS) ev = T_GHOST;
                if (y > midY + EPS && y < yu[J] - EPS && y > yu[J - 1] + EPS) ev = T_HYBP;
                else if (y < midY + EPS && y > yu[J] + EPS && y < yu[J + 1] - EPS) ev = T_HYBP;
                if (ev == T_GHOST) fl |= F_GHOSTP;
                if (ev == T_HYBP) fl |= F_HYBP;
                if (ev >= 0) rown = l * 8 + ev;
                if (fabs(y - yu[J]) < EPS && fabs(midY - yu[J]) > fabs(midX - xv[I])) {
                    fl |= F_HYBP;
                    if (y < midY) { rnb[DIR_N] = T_C3; rown = T_C3; }
                    if (y > midY) { rnb[DIR_S] = T_C4; rown = T_C4; }
                    flag = true;
                }
            }
        }
        k = l;
        l = l + 1;
    }
    own[T] = rown; nb0[T] = rnb[0]; nb1[T] = rnb[1]; nb2[T] = rnb[2]; nb3[T] = rnb[3]; flags[T] = fl;
}

__global__ void __launch_bounds__(BS) k_tag_p_resolve(GridDev g, BoxDev box, BodyDev B, const int *own,
                                                      const int *nb0, const int *nb1, const int *nb2, const int *nb3,
                                                      const int *flags, TagOut o) {
    const int nx = g.nx, ny = g.ny;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nx * ny) return;
    const int I = t % nx, J = t / nx;
    const int wI[5] = {I, I - 1, I, I + 1, I};
    const int wJ[5] = {J - 1, J, J, J, J + 1};
    int gp = o.ghostP[t], hp = o.hybridP[t];
    int code = -1, cw = -1;
    for (int q = 0; q < 5; q++) {
        if (wI[q] < 0 || wI[q] >= nx || wJ[q] < 0 || wJ[q] >= ny) continue;
        const int T = box_thread(box, wI[q], wJ[q]);
        if (T < 0) continue;
        int cc;
        if (q == 2) {
            cc = own[T];
            if (flags[T] & F_GHOSTP) gp = t;
            if (flags[T] & F_HYBP) hp = t;
        } else {
            cc = (q == 0) ? nb3[T] : (q == 1) ? nb1[T] : (q == 3) ? nb0[T] : nb2[T];
            if (cc != -1) gp = t;
        }
        if (cc != -1) { code = cc; cw = q; }
    }
    o.ghostP[t] = gp;
    o.hybridP[t] = hp;
    if (code != -1) {
        const int Iw = wI[cw], Jw = wJ[cw];
        double bix = 0, biy = 0, ipx = 0, ipy = 0;
        eval_code(code, B.x, B.y, B.n, g.x[I], g.y[J], g.x[Iw], g.y[Jw], Iw > 0 ? g.x[Iw - 1] : 0.0,
                  Iw + 1 < nx ? g.x[Iw + 1] : 0.0, Jw > 0 ? g.y[Jw - 1] : 0.0, Jw + 1 < ny ? g.y[Jw + 1] : 0.0, bix,
                  biy, ipx, ipy);
        o.bip_x[t] = bix; o.bip_y[t] = biy; o.ipp_x[t] = ipx; o.ipp_y[t] = ipy;
    }
}

// zero_pressure_luo / zero_x_luo / zero_y_luo (tagPoints.cu:1157-1298): each row of the
// box is scanned left to right, marking the body interior solid (tag 0). One block per
// row stages the row segment [i_start-1, max(kmax, i_end-1)] in LDS. The reference's
// look-ahead loop at a hit i asks whether any position k in (i, kmax] holds a tag other
// than -1; the scan writes only at positions <= i, so those positions still hold their
// entry values and the answer is L > i, with L the last such position of the row at entry
// (a block max). Lane 0 then runs the serial scan on LDS and stores each change; a
// position past the end of the tag array reads as -1, as in the look-ahead's bound.
__global__ void __launch_bounds__(BS) k_zero_rows(int *t, BoxDev box, int rowlen, int kmax, int off, long len) {
    extern __shared__ int seg[];
    __shared__ int red[BS / 64];
    const int J = box.j_start + blockIdx.x;
    const int i0 = box.i_start - 1, i_end = box.i_start + box.w;
    const int nseg = max(kmax, i_end - 1) - i0 + 1;
    const long base = (long)J * rowlen + off;
    int last = i0;
    for (int k = threadIdx.x; k < nseg; k += blockDim.x) {
        const long q = base + i0 + k;
        const int v = (q >= 0 && q < len) ? t[q] : -1;
        seg[k] = v;
        const int pos = i0 + k;
        if (pos > box.i_start && pos <= kmax && v != -1) last = max(last, pos);
    }
    for (int o = 32; o > 0; o >>= 1) last = max(last, __shfl_xor(last, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = last;
    __syncthreads();
    if (threadIdx.x != 0) return;
    int L = red[0];
    for (int w = 1; w < BS / 64; w++) L = max(L, red[w]);
    bool rowIsntDone = true, flag = false;
    for (int i = box.i_start; i < i_end; i++) {
        const int prev = seg[i - 1 - i0];
        int cur = seg[i - i0];
        if (prev < 0) continue;
        if (cur != -1 && prev == 0) rowIsntDone = false;
        int nv = cur;
        if (cur == -1 && rowIsntDone) nv = 0;
        if (nv == 0 && prev != 0) {
            flag = flag || L > i;
            if (!flag) {
                nv = -1;
                rowIsntDone = false;
            }
        }
        if (nv != cur) {
            seg[i - i0] = nv;
            if (base + i >= 0 && base + i < len) t[base + i] = nv;
        }
    }
}

// ---- bilinear interpolation weights ------------------------------------------------

// cofactors of the 4x4 system (biLinearInterpolation.cu:129-168); weights at (X,Y):
// w_l = (b_1l + b_2l X + b_3l Y + b_4l X Y) / det
__device__ void bil_weights(const double a[16], double X, double Y, double w[4]) {
    const double a11 = a[0], a12 = a[1], a13 = a[2], a14 = a[3];
    const double a21 = a[4], a22 = a[5], a23 = a[6], a24 = a[7];
    const double a31 = a[8], a32 = a[9], a33 = a[10], a34 = a[11];
    const double a41 = a[12], a42 = a[13], a43 = a[14], a44 = a[15];
    const double det = a11 * a22 * a33 * a44 + a11 * a23 * a34 * a42 + a11 * a24 * a32 * a43
                     + a12 * a21 * a34 * a43 + a12 * a23 * a31 * a44 + a12 * a24 * a33 * a41
                     + a13 * a21 * a32 * a44 + a13 * a22 * a34 * a41 + a13 * a24 * a31 * a42
                     + a14 * a21 * a33 * a42 + a14 * a22 * a31 * a43 + a14 * a23 * a32 * a41
                     - a11 * a22 * a34 * a43 - a11 * a23 * a32 * a44 - a11 * a24 * a33 * a42
                     - a12 * a21 * a33 * a44 - a12 * a23 * a34 * a41 - a12 * a24 * a31 * a43
                     - a13 * a21 * a34 * a42 - a13 * a22 * a31 * a44 - a13 * a24 * a32 * a41
                     - a14 * a21 * a32 * a43 - a14 * a22 * a33 * a41 - a14 * a23 * a31 * a42;
    double b[4][4];
    b[0][0] = a22 * a33 * a44 + a23 * a34 * a42 + a24 * a32 * a43 - a22 * a34 * a43 - a23 * a32 * a44 - a24 * a33 * a42;
    b[0][1] = a12 * a34 * a43 + a13 * a32 * a44 + a14 * a33 * a42 - a12 * a33 * a44 - a13 * a34 * a42 - a14 * a32 * a43;
    b[0][2] = a12 * a23 * a44 + a13 * a24 * a42 + a14 * a22 * a43 - a12 * a24 * a43 - a13 * a22 * a44 - a14 * a23 * a42;
    b[0][3] = a12 * a24 * a33 + a13 * a22 * a34 + a14 * a23 * a32 - a12 * a23 * a34 - a13 * a24 * a32 - a14 * a22 * a33;
    b[1][0] = a21 * a34 * a43 + a23 * a31 * a44 + a24 * a33 * a41 - a21 * a33 * a44 - a23 * a34 * a41 - a24 * a31 * a43;
    b[1][1] = a11 * a33 * a44 + a13 * a34 * a41 + a14 * a31 * a43 - a11 * a34 * a43 - a13 * a31 * a44 - a14 * a33 * a41;
    b[1][2] = a11 * a24 * a43 + a13 * a21 * a44 + a14 * a23 * a41 - a11 * a23 * a44 - a13 * a24 * a41 - a14 * a21 * a43;
    b[1][3] = a11 * a23 * a34 + a13 * a24 * a31 + a14 * a21 * a33 - a11 * a24 * a33 - a13 * a21 * a34 - a14 * a23 * a31;
    b[2][0] = a21 * a32 * a44 + a22 * a34 * a41 + a24 * a31 * a42 - a21 * a34 * a42 - a22 * a31 * a44 - a24 * a32 * a41;
    b[2][1] = a11 * a34 * a42 + a12 * a31 * a44 + a14 * a32 * a41 - a11 * a32 * a44 - a12 * a34 * a41 - a14 * a31 * a42;
    b[2][2] = a11 * a22 * a44 + a12 * a24 * a41 + a14 * a21 * a42 - a11 * a24 * a42 - a12 * a21 * a44 - a14 * a22 * a41;
    b[2][3] = a11 * a24 * a32 + a12 * a21 * a34 + a14 * a22 * a31 - a11 * a22 * a34 - a12 * a24 * a31 - a14 * a21 * a32;
    b[3][0] = a21 * a33 * a42 + a22 * a31 * a43 + a23 * a32 * a41 - a21 * a32 * a43 - a22 * a33 * a41 - a23 * a31 * a42;
    b[3][1] = a11 * a32 * a43 + a12 * a33 * a41 + a13 * a31 * a42 - a11 * a33 * a42 - a12 * a31 * a43 - a13 * a32 * a41;
    b[3][2] = a11 * a23 * a42 + a12 * a21 * a43 + a13 * a22 * a41 - a11 * a22 * a43 - a12 * a23 * a41 - a13 * a21 * a42;
    b[3][3] = a11 * a22 * a33 + a12 * a23 * a31 + a13 * a21 * a32 - a11 * a23 * a32 - a12 * a21 * a33 - a13 * a22 * a31;
    for (int l = 0; l < 4; l++) w[l] = (b[0][l] + b[1][l] * X + b[2][l] * Y + b[3][l] * X * Y) / det;
}

__device__ __forceinline__ void bil_rows(const double x[4], const double y[4], double a[16]) {
    for (int l = 0; l < 4; l++) {
        a[l * 4] = 1;
        a[l * 4 + 1] = x[l];
        a[l * 4 + 2] = y[l];
        a[l * 4 + 3] = x[l] * y[l];
    }
}

struct CompGeom {
    const double *xs, *ys;
    int ni, nj, stride, off;
};

__device__ __forceinline__ CompGeom comp_of(const GridDev &g, int node) {
    const int numU = (g.nx - 1) * g.ny;
    if (node < numU) return {g.xu, g.y, g.nx - 1, g.ny, g.nx - 1, 0};
    return {g.x, g.yv, g.nx, g.ny - 1, g.nx, numU};
}

// interpolateVelocityToGhostNodeX/Y weights (biLinearInterpolation.cu:14-353)
__global__ void k_gn_vel_coef(GridDev g, IbmGeom m, int n, const int *node, int *idx, int *idx0, double *coef) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int gi = node[t];
    const CompGeom c = comp_of(g, gi);
    const int I = (gi - c.off) % c.stride, J = (gi - c.off) / c.stride;
    int ii = I - 5 < 1 ? 1 : I - 5, jj = J - 5 < 1 ? 1 : J - 5;
    while (ii < c.ni - 1 && c.xs[ii] < m.ip_x[gi]) ii++;
    while (jj < c.nj - 1 && c.ys[jj] < m.ip_y[gi]) jj++;
    double x[4] = {c.xs[ii - 1], c.xs[ii], c.xs[ii - 1], c.xs[ii]};
    double y[4] = {c.ys[jj - 1], c.ys[jj - 1], c.ys[jj], c.ys[jj]};
    int index[4] = {(jj - 1) * c.stride + ii - 1 + c.off, (jj - 1) * c.stride + ii + c.off,
                    jj * c.stride + ii - 1 + c.off, jj * c.stride + ii + c.off};
    double mn = 1.0;
    int close_index = -1;
    bool inflag = false;
    for (int l = 0; l < 4; l++) {
        const double s = sqrt(sq(x[l] - m.bi_x[gi]) + sq(y[l] - m.bi_y[gi]));
        if (s < mn) { mn = s; close_index = index[l]; }
        if (m.ghostUV[index[l]] > 0) inflag = true;
    }
    for (int l = 0; l < 4; l++) {
        idx0[4 * t + l] = index[l];
        if (m.ghostUV[index[l]] == gi) {
            x[l] = m.bi_x[index[l]];
            y[l] = m.bi_y[index[l]];
            index[l] = -1;
        } else if (index[l] == close_index && !inflag) {
            x[l] = m.bi_x[gi];
            y[l] = m.bi_y[gi];
            index[l] = -1;
        }
    }
    double a[16], w[4];
    bil_rows(x, y, a);
    bil_weights(a, m.ip_x[gi], m.ip_y[gi], w);
    for (int l = 0; l < 4; l++) {
        idx[4 * t + l] = index[l];
        coef[4 * t + l] = w[l];
    }
}

// interpolateVelocityToHybridNodeX/Y weights and the luoIBM alpha (weight.cu:11-92)
__global__ void k_hn_vel_coef(GridDev g, IbmGeom m, int n, const int *node, int *idx, int *idx0, double *coef,
                              double *alpha) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int h = node[t];
    const CompGeom c = comp_of(g, h);
    const int I = (h - c.off) % c.stride, J = (h - c.off) / c.stride;
    int ii = I - 5 < 1 ? 1 : I - 5, jj = J - 5 < 1 ? 1 : J - 5;
    while (ii < c.ni - 1 && c.xs[ii] < m.ip_x[h]) ii++;
    while (jj < c.nj - 1 && c.ys[jj] < m.ip_y[h]) jj++;
    double x[4] = {c.xs[ii - 1], c.xs[ii], c.xs[ii - 1], c.xs[ii]};
    double y[4] = {c.ys[jj - 1], c.ys[jj - 1], c.ys[jj], c.ys[jj]};
    int index[4] = {(jj - 1) * c.stride + ii - 1 + c.off, (jj - 1) * c.stride + ii + c.off,
                    jj * c.stride + ii - 1 + c.off, jj * c.stride + ii + c.off};
    for (int l = 0; l < 4; l++) {
        idx0[4 * t + l] = index[l];
        if (m.hybridUV[index[l]] == h) {
            x[l] = m.bi_x[h];
            y[l] = m.bi_y[h];
            index[l] = -1;
        }
    }
    double a[16], w[4];
    bil_rows(x, y, a);
    bil_weights(a, c.xs[I], c.ys[J], w);
    for (int l = 0; l < 4; l++) {
        idx[4 * t + l] = index[l];
        coef[4 * t + l] = w[l];
    }
    const int st = c.stride;
    double d1 = 0, d2 = 0;
    const double dxl = c.xs[I] - c.xs[I - 1], dyl = c.ys[J] - c.ys[J - 1];
    if (m.ghostUV[h - 1] > 0) d1 = sqrt(sq(m.bi_x[h - 1] - c.xs[I - 1]) + sq((m.bi_y[h - 1] - c.ys[J])));
    else if (m.ghostUV[h + 1] > 0) d1 = sqrt(sq((m.bi_x[h + 1] - c.xs[I + 1])) + sq((m.bi_y[h + 1] - c.ys[J])));
    if (m.ghostUV[h - st] > 0) d2 = sqrt(sq(m.bi_x[h - st] - c.xs[I]) + sq(m.bi_y[h - st] - c.ys[J - 1]));
    if (m.ghostUV[h + st] > 0) d2 = sqrt(sq(m.bi_x[h + st] - c.xs[I]) + sq(m.bi_y[h + st] - c.ys[J + 1]));
    alpha[t] = sqrt(sq(d1 / dxl) + sq(d2 / dyl));
}

// interpolatePressureToHybridNode weights with the Neumann row (biLinearInterpolation.cu:671-856)
// and weightP's alpha (luoIBM/kernels/weight.cu:95-134)
__global__ void k_hn_p_coef(GridDev g, IbmGeom m, int n, const int *node, int *idx, int *idx0, double *coef,
                            double *nrm, double *alpha) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int ip = node[t], nx = g.nx;
    const int I = ip % nx, J = ip / nx;
    const double *xv = g.x, *yu = g.y;
    int ii = I - 5 < 1 ? 1 : I - 5, jj = J - 5 < 1 ? 1 : J - 5;
    while (ii < nx - 1 && xv[ii] < m.ipp_x[ip]) ii++;
    while (jj < g.ny - 1 && yu[jj] < m.ipp_y[ip]) jj++;
    double x[4] = {xv[ii - 1], xv[ii], xv[ii - 1], xv[ii]};
    double y[4] = {yu[jj - 1], yu[jj - 1], yu[jj], yu[jj]};
    int index[4] = {(jj - 1) * nx + ii - 1, (jj - 1) * nx + ii, jj * nx + ii - 1, jj * nx + ii};
    double a[16];
    bil_rows(x, y, a);
    for (int l = 0; l < 4; l++) {
        idx0[4 * t + l] = index[l];
        nrm[8 * t + 2 * l] = 0;
        nrm[8 * t + 2 * l + 1] = 0;
        if (m.hybridP[index[l]] == ip) {
            x[l] = m.bip_x[ip];
            y[l] = m.bip_y[ip];
            const double n_x = m.ipp_x[ip] - x[l], n_y = m.ipp_y[ip] - y[l];
            const double nl = sqrt(n_x * n_x + n_y * n_y);
            a[l * 4] = 0;
            a[l * 4 + 1] = n_x / nl;
            a[l * 4 + 2] = n_y / nl;
            a[l * 4 + 3] = n_y / nl * x[l] + n_x / nl * y[l];
            nrm[8 * t + 2 * l] = n_x / nl;
            nrm[8 * t + 2 * l + 1] = n_y / nl;
            index[l] = -1;
        }
    }
    double w[4];
    bil_weights(a, xv[I], yu[J], w);
    for (int l = 0; l < 4; l++) {
        idx[4 * t + l] = index[l];
        coef[4 * t + l] = w[l];
    }
    double d1 = 0, d2 = 0;
    const double dxl = xv[I] - xv[I - 1], dyl = yu[J] - yu[J - 1];
    if (m.ghostP[ip - 1] > 0) d1 = sqrt(sq(m.bip_x[ip - 1] - xv[I - 1]) + sq(m.bip_y[ip - 1] - yu[J]));
    else if (m.ghostP[ip + 1] > 0) d1 = sqrt(sq(m.bip_x[ip + 1] - xv[I + 1]) + sq(m.bip_y[ip + 1] - yu[J]));
    if (m.ghostP[ip - nx] > 0) d2 = sqrt(sq(m.bip_x[ip - nx] - xv[I]) + sq(m.bip_y[ip - nx] - yu[J - 1]));
    if (m.ghostP[ip + nx] > 0) d2 = sqrt(sq(m.bip_x[ip + nx] - xv[I]) + sq(m.bip_y[ip + nx] - yu[J + 1]));
    alpha[t] = sqrt(sq(d1 / dxl) + sq(d2 / dyl));
}

// interpolatePressureToGhostNode weights (biLinearInterpolation.cu:859-1066); extra =
// (distance image point - node, close-corner normal x, y)
__global__ void k_gn_p_coef(GridDev g, IbmGeom m, int n, const int *node, int *idx, int *idx0, double *coef,
                            double *nrm, double *extra) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int ip = node[t], nx = g.nx;
    const int I = ip % nx, J = ip / nx;
    const double *xv = g.x, *yu = g.y;
    int ii = I - 5 < 1 ? 1 : I - 5, jj = J - 5 < 1 ? 1 : J - 5;
    while (ii < nx - 1 && xv[ii] < m.ipp_x[ip]) ii++;
    while (jj < g.ny - 1 && yu[jj] < m.ipp_y[ip]) jj++;
    double x[4] = {xv[ii - 1], xv[ii], xv[ii - 1], xv[ii]};
    double y[4] = {yu[jj - 1], yu[jj - 1], yu[jj], yu[jj]};
    int index[4] = {(jj - 1) * nx + ii - 1, (jj - 1) * nx + ii, jj * nx + ii - 1, jj * nx + ii};
    double a[16];
    bil_rows(x, y, a);
    double mn = 1.0, cnx = 0, cny = 0;
    int close_index = -1;
    for (int l = 0; l < 4; l++) {
        const double s = sqrt(sq(x[l] - m.bip_x[ip]) + sq(y[l] - m.bip_y[ip]));
        if (s < mn) { mn = s; close_index = index[l]; }
    }
    int out_idx[4];
    for (int l = 0; l < 4; l++) {
        idx0[4 * t + l] = index[l];
        out_idx[l] = index[l];
        nrm[8 * t + 2 * l] = 0;
        nrm[8 * t + 2 * l + 1] = 0;
        if (m.ghostP[index[l]] > 0) {
            x[l] = m.bip_x[index[l]];
            y[l] = m.bip_y[index[l]];
            const double n_x = m.ipp_x[index[l]] - x[l], n_y = m.ipp_y[index[l]] - y[l];
            const double nl = sqrt(n_x * n_x + n_y * n_y);
            a[l * 4] = 0;
            a[l * 4 + 1] = n_x / nl;
            a[l * 4 + 2] = n_y / nl;
            a[l * 4 + 3] = n_y / nl * x[l] + n_x / nl * y[l];
            nrm[8 * t + 2 * l] = n_x / nl;
            nrm[8 * t + 2 * l + 1] = n_y / nl;
            out_idx[l] = -1;
        }
        if (index[l] == close_index) {
            x[l] = m.bip_x[ip];
            y[l] = m.bip_y[ip];
            const double n_x = m.ipp_x[ip] - x[l], n_y = m.ipp_y[ip] - y[l];
            const double nl = sqrt(n_x * n_x + n_y * n_y);
            cnx = n_x / nl;
            cny = n_y / nl;
        }
    }
    double w[4];
    bil_weights(a, m.ipp_x[ip], m.ipp_y[ip], w);
    for (int l = 0; l < 4; l++) {
        idx[4 * t + l] = out_idx[l];
        coef[4 * t + l] = w[l];
    }
    extra[3 * t] = sqrt(sq(m.ipp_x[ip] - xv[I]) + sq(m.ipp_y[ip] - yu[J]));
    extra[3 * t + 1] = cnx;
    extra[3 * t + 2] = cny;
}

// ---- per-step apply kernels ----------------------------------------------------------

__global__ void k_gn_vel_eval(const double *__restrict__ f, int n, const int *__restrict__ node,
                              const int *__restrict__ idx, const double *__restrict__ coef, double *__restrict__ tmp,
                              int numU, double uB, double vB, const double *__restrict__ ubn) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double ub = ubn ? ubn[t] : (node[t] < numU ? uB : vB);
    double s = 0;
    for (int l = 0; l < 4; l++) {
        const int j = idx[4 * t + l];
        s += coef[4 * t + l] * (j >= 0 ? f[j] : ub);
    }
    tmp[t] = 2 * ub - s;
}

__global__ void k_scatter(double *__restrict__ f, int n, const int *__restrict__ node, const double *__restrict__ tmp) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) f[node[t]] = tmp[t];
}

__global__ void k_set_inside(double *__restrict__ f, const int *__restrict__ solid, int n, int numU, double uB,
                             double vB, const double *__restrict__ ubn) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        const int s = solid[t];
        f[s] = ubn ? ubn[t] : (s < numU ? uB : vB);
    }
}

__global__ void k_hn_vel_eval(const double *__restrict__ f, double *__restrict__ ustar, int n,
                              const int *__restrict__ node, const int *__restrict__ idx,
                              const double *__restrict__ coef, int numU, double uB, double vB,
                              const double *__restrict__ ubn) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double ub = ubn ? ubn[t] : (node[t] < numU ? uB : vB);
    double s = 0;
    for (int l = 0; l < 4; l++) {
        const int j = idx[4 * t + l];
        s += coef[4 * t + l] * (j >= 0 ? f[j] : ub);
    }
    ustar[node[t]] = s;
}

__global__ void k_weight(double *__restrict__ f, const double *__restrict__ star, int n, const int *__restrict__ node,
                         const double *__restrict__ alpha) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int h = node[t];
    const double a = alpha[t];
    f[h] = (1 - a) * f[h] + a * star[h];
}

__global__ void k_hn_p_eval(const double *__restrict__ p, double *__restrict__ pstar, int n,
                            const int *__restrict__ node, const int *__restrict__ idx, const double *__restrict__ coef,
                            const double *__restrict__ nrm, double dudt, double dvdt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double s = 0;
    for (int l = 0; l < 4; l++) {
        const int j = idx[4 * t + l];
        const double q = j >= 0 ? p[j] : -(nrm[8 * t + 2 * l] * dudt + nrm[8 * t + 2 * l + 1] * dvdt);
        s += coef[4 * t + l] * q;
    }
    pstar[node[t]] = s;
}

__global__ void k_gn_p_eval(double *__restrict__ p, int n, const int *__restrict__ node, const int *__restrict__ idx,
                            const double *__restrict__ coef, const double *__restrict__ nrm,
                            const double *__restrict__ extra, double dudt, double dvdt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double s = 0;
    for (int l = 0; l < 4; l++) {
        const int j = idx[4 * t + l];
        const double q = j >= 0 ? p[j] : -(nrm[8 * t + 2 * l] * dudt + nrm[8 * t + 2 * l + 1] * dvdt);
        s += coef[4 * t + l] * q;
    }
    const double matDClose = extra[3 * t + 1] * dudt + extra[3 * t + 2] * dvdt;
    p[node[t]] = s + extra[3 * t] * matDClose;
}

// ---- control-volume force (FadlunModified/kernels/calculateForce.cu:32-283) --------------

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// the force terms of work item t of the box (FadlunModified/kernels/calculateForce.cu:32-283),
// added into acc[0..5]
__device__ __forceinline__ void force_term(long t, const double *__restrict__ u, const double *__restrict__ uold,
                                           const double *__restrict__ p, const int *__restrict__ tagsIn,
                                           const GridDev &g, double nu, double dt, int I, int J, int ncx, int ncy,
                                           double (&acc)[6]) {
    const int nx = g.nx, ny = g.ny;
    const double *dx = g.dx, *dy = g.dy;
    const long n0 = ncy, n1 = ncx + 1, n2 = (long)(ncx + 1) * ncy, n3 = ncy + 1, n4 = ncx;
    long idx = t;
    if (idx < n0) {  // dragLeftRight
        const int Ip = (J + idx) * nx + I, Iu = (J + idx) * (nx - 1) + I;
        acc[0] += -((p[Ip + ncx] - p[Ip]) * dy[J + idx] +
                    ((u[Iu + ncx] + u[Iu + ncx - 1]) * (u[Iu + ncx] + u[Iu + ncx - 1]) / 4 -
                     (u[Iu - 1] + u[Iu]) * (u[Iu - 1] + u[Iu]) / 4) * dy[J + idx] -
                    nu * ((u[Iu + ncx] - u[Iu + ncx - 1]) / dx[I + ncx] - (u[Iu] - u[Iu - 1]) / dx[I]) * dy[J + idx]);
        return;
    }
    idx -= n0;
    if (idx < n1) {  // dragBottomTop
        const int Iu = J * (nx - 1) + (I - 1 + idx), Iv = (nx - 1) * ny + (J - 1) * nx + I + idx;
        acc[1] += -((0.25 * (u[Iu + ncy * (nx - 1)] + u[Iu + (ncy - 1) * (nx - 1)]) * (u[Iv + ncy * nx] + u[Iv + ncy * nx - 1]) -
                      0.25 * (u[Iu] + u[Iu - (nx - 1)]) * (u[Iv] + u[Iv - 1])) -
                    nu * (((u[Iu + ncy * (nx - 1)] - u[Iu + (ncy - 1) * (nx - 1)]) / 2.0 / (dy[J + ncy] + dy[J + ncy - 1]) +
                           (u[Iv + ncy * nx] - u[Iv + ncy * nx - 1]) / 2.0 / (dx[I + idx] + dx[I + idx - 1])) -
                          ((u[Iu] - u[Iu - (nx - 1)]) / 2.0 / (dy[J] + dy[J - 1]) +
                           (u[Iv] - u[Iv - 1]) / 2.0 / (dx[I + idx] + dx[I + idx - 1])))) *
                  0.5 * (dx[I + idx] + dx[I + idx - 1]);
        return;
    }
    idx -= n1;
    if (idx < n2) {  // dragUnsteady
        const int i = idx % (ncx + 1), j = idx / (ncx + 1);
        const int Iu = (J + j) * (nx - 1) + (I - 1 + i);
        acc[2] += -(tagsIn[Iu] == -1) * ((u[Iu] * dy[J + j] - uold[Iu] * dy[J + j]) / dt * 0.5 * (dx[I + i] + dx[I - 1 + i]));
        return;
    }
    idx -= n2;
    if (idx < n3) {  // liftLeftRight
        const int Iu = (J + idx) * (nx - 1) + (I - 1), Iv = (nx - 1) * ny + (J - 1 + idx) * nx + I;
        acc[3] += -((0.25 * (u[Iu + ncx] + u[Iu + ncx - (nx - 1)]) * (u[Iv + ncx] + u[Iv + ncx - 1]) -
                      0.25 * (u[Iu] + u[Iu - (nx - 1)]) * (u[Iv] + u[Iv - 1])) -
                    nu * (((u[Iu + ncx] - u[Iu + ncx - (nx - 1)]) / 2.0 / (dy[J + idx] + dy[J - 1 + idx]) +
                           (u[Iv + ncx] - u[Iv + ncx - 1]) / 2.0 / (dx[I + ncx] + dx[I + ncx - 1])) -
                          ((u[Iu] - u[Iu - (nx - 1)]) / 2.0 / (dy[J + idx] + dy[J - 1 + idx]) +
                           (u[Iv] - u[Iv - 1]) / 2.0 / (dx[I] + dx[I - 1])))) *
                  0.5 * (dy[J + idx] + dy[J - 1 + idx]);
        return;
    }
    idx -= n3;
    if (idx < n4) {  // liftBottomTop
        const int Ip = J * nx + I + idx, Iv = (nx - 1) * ny + (J - 1) * nx + I + idx;
        acc[4] += -((p[Ip + ncy * nx] - p[Ip - nx]) * dx[I + idx] +
                    (0.25 * (u[Iv + (ncy + 1) * nx] + u[Iv + ncy * nx]) * (u[Iv + (ncy + 1) * nx] + u[Iv + ncy * nx]) -
                     0.25 * (u[Iv] + u[Iv - nx]) * (u[Iv] + u[Iv - nx])) * dx[I + idx] -
                    nu * ((u[Iv + (ncy + 1) * nx] - u[Iv + ncy * nx]) * dx[I + idx] / dy[J + ncy] -
                          (u[Iv] - u[Iv - nx]) * dx[I] / dy[J - 1]));
        return;
    }
    idx -= n4;
    {  // liftUnsteady
        const int i = idx % ncx, j = idx / ncx;
        const int Iv = (J - 1 + j) * nx + (I + i) + (nx - 1) * ny;
        acc[5] += -(tagsIn[Iv] == -1) * ((u[Iv] * dx[I + i] - uold[Iv] * dx[I + i]) / dt * 0.5 * (dy[J + j] + dy[J - 1 + j]));
    }
}

__device__ __forceinline__ long force_items(int ncx, int ncy) {
    return (long)ncy + (ncx + 1) + (long)(ncx + 1) * ncy + (ncy + 1) + ncx + (long)ncx * (ncy + 1);
}

__global__ void __launch_bounds__(BS) k_force(const double *__restrict__ u, const double *__restrict__ uold,
                                              const double *__restrict__ p, const int *__restrict__ tagsIn, GridDev g,
                                              double nu, double dt, int I, int J, int ncx, int ncy,
                                              double *__restrict__ part) {
    const long total = force_items(ncx, ncy);
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x)
        force_term(t, u, uold, p, tagsIn, g, nu, dt, I, J, ncx, ncy, acc);
    __shared__ double sh[6][BS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int k = 0; k < 6; k++) {
        const double v = wsum(acc[k]);
        if (lane == 0) sh[k][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        part[k * gridDim.x + blockIdx.x] = (sh[k][0] + sh[k][1]) + (sh[k][2] + sh[k][3]);
    }
}

__global__ void k_force_finish(const double *__restrict__ part, int nblocks, double *__restrict__ out) {
    const int k = threadIdx.x;
    if (k < 6) {
        double s = 0;
        for (int b = 0; b < nblocks; b++) s += part[k * nblocks + b];
        out[k] = s;
    }
}

}  // namespace

void launch_tag_velocity(const BodyDev &b, const CompDev &c, const BoxDev &box, bool vdir, TagRecords &rec,
                         const TagOut &o, int nx, hipStream_t s) {
    if (vdir)
        hipLaunchKernelGGL(k_tag_vel<true>, dim3(nblk(box.nthr)), dim3(BS), 0, s, b, c, box, nx, rec.own.p(), rec.nb0.p(),
                           rec.nb1.p(), rec.nb2.p(), rec.nb3.p(), rec.flags.p(), rec.dv0.p(), rec.dv1.p(), o);
    else
        hipLaunchKernelGGL(k_tag_vel<false>, dim3(nblk(box.nthr)), dim3(BS), 0, s, b, c, box, nx, rec.own.p(),
                           rec.nb0.p(), rec.nb1.p(), rec.nb2.p(), rec.nb3.p(), rec.flags.p(), rec.dv0.p(), rec.dv1.p(), o);
    CUIBM_LAUNCH_CHECK();
}

void launch_tag_velocity_resolve(const CompDev &c, const BoxDev &box, const BodyDev &b, bool vdir,
                                 const TagRecords &rec, const TagOut &o, int nx, int ny, hipStream_t s) {
    const long n = std::max((long)c.ni * c.nj, (long)nx * ny);
    if (vdir)
        hipLaunchKernelGGL(k_tag_vel_resolve<true>, dim3(nblk(n)), dim3(BS), 0, s, c, box, b, rec.own.p(), rec.nb0.p(),
                           rec.nb1.p(), rec.nb2.p(), rec.nb3.p(), rec.flags.p(), rec.dv0.p(), rec.dv1.p(), o, nx, ny);
    else
        hipLaunchKernelGGL(k_tag_vel_resolve<false>, dim3(nblk(n)), dim3(BS), 0, s, c, box, b, rec.own.p(),
                           rec.nb0.p(), rec.nb1.p(), rec.nb2.p(), rec.nb3.p(), rec.flags.p(), rec.dv0.p(), rec.dv1.p(),
                           o, nx, ny);
    CUIBM_LAUNCH_CHECK();
}

void launch_tag_pressure(const BodyDev &b, const GridDev &g, const BoxDev &box, TagRecords &rec, hipStream_t s) {
    hipLaunchKernelGGL(k_tag_p, dim3(nblk(box.nthr)), dim3(BS), 0, s, b, g, box, rec.own.p(), rec.nb0.p(), rec.nb1.p(),
                       rec.nb2.p(), rec.nb3.p(), rec.flags.p());
    CUIBM_LAUNCH_CHECK();
}

void launch_tag_pressure_resolve(const GridDev &g, const BoxDev &box, const BodyDev &b, const TagRecords &rec,
                                 const TagOut &o, hipStream_t s) {
    hipLaunchKernelGGL(k_tag_p_resolve, dim3(nblk((long)g.nx * g.ny)), dim3(BS), 0, s, g, box, b, rec.own.p(),
                       rec.nb0.p(), rec.nb1.p(), rec.nb2.p(), rec.nb3.p(), rec.flags.p(), o);
    CUIBM_LAUNCH_CHECK();
}

namespace {

constexpr int NL_ITEMS = 8, NL_CHUNK = BS * NL_ITEMS;

// candidate c of list k: its node index, or -1 when it is not on the list
__device__ inline int nl_pick(int k, int c, const NodeListIn &in) {
    switch (k) {
        case 0: return c < in.numUV && in.ghostUV[c] > 0 ? c : -1;
        case 1: return c < in.numUV && in.hybridUV[c] > 0 ? c : -1;
        case 2: return c < in.numP && in.ghostP[c] > 0 ? c : -1;
        case 3: return c < in.numP && in.hybridP[c] > 0 ? c : -1;
        default: {
            if (c >= 2 * in.numU) return -1;
            const int i = c >> 1;
            if ((c & 1) == 0) return in.ghostUV[i] == 0 ? i : -1;
            const int iv = (i / (in.nx - 1)) * in.nx + i % (in.nx - 1) + in.numU;
            return iv < in.numUV && in.ghostUV[iv] == 0 ? iv : -1;
        }
    }
}

// pass 1: members of list blockIdx.y in chunk blockIdx.x
__global__ void __launch_bounds__(BS) k_nl_count(NodeListIn in, int *__restrict__ part) {
    __shared__ int sh[BS];
    const int k = blockIdx.y, c0 = blockIdx.x * NL_CHUNK + threadIdx.x * NL_ITEMS;
    int m = 0;
    for (int e = 0; e < NL_ITEMS; e++) m += nl_pick(k, c0 + e, in) >= 0;
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int o = BS / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[k * gridDim.x + blockIdx.x] = sh[0];
}

// pass 2: the chunk's members written in candidate order after those of earlier chunks
__global__ void __launch_bounds__(BS) k_nl_write(NodeListIn in, const int *__restrict__ part, int *o0, int *o1, int *o2,
                                                 int *o3, int *o4, int *__restrict__ counts) {
    __shared__ int sh[BS];
    __shared__ int base_s;
    const int k = blockIdx.y, nb = gridDim.x, c0 = blockIdx.x * NL_CHUNK + threadIdx.x * NL_ITEMS;
    int *out = k == 0 ? o0 : k == 1 ? o1 : k == 2 ? o2 : k == 3 ? o3 : o4;
    // members of the earlier chunks of this list
    int b = 0;
    for (int q = threadIdx.x; q < (int)blockIdx.x; q += BS) b += part[k * nb + q];
    sh[threadIdx.x] = b;
    __syncthreads();
    for (int o = BS / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) base_s = sh[0];
    __syncthreads();
    const int base = base_s;
    int m = 0;
    for (int e = 0; e < NL_ITEMS; e++) m += nl_pick(k, c0 + e, in) >= 0;
    // exclusive scan of the per-thread counts
    __syncthreads();
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int o = 1; o < BS; o <<= 1) {
        const int v = (int)threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += v;
        __syncthreads();
    }
    int at = base + sh[threadIdx.x] - m;
    for (int e = 0; e < NL_ITEMS; e++) {
        const int v = nl_pick(k, c0 + e, in);
        if (v >= 0) out[at++] = v;
    }
    if (blockIdx.x == nb - 1 && threadIdx.x == BS - 1) counts[k] = base + sh[BS - 1];
}

int nl_chunks(int numUV, int numP, int numU) {
    const long m = std::max<long>(std::max<long>(numUV, numP), 2L * numU);
    return (int)std::max<long>(1, (m + NL_CHUNK - 1) / NL_CHUNK);
}

}  // namespace

int node_list_parts(int numUV, int numP, int numU) { return 5 * nl_chunks(numUV, numP, numU); }

void launch_node_lists(const NodeListIn &in, int *const out[5], int *part, int *counts, hipStream_t s) {
    if (in.nx < 2) throw Error("node lists: grid narrower than two cells", ERR_ARGUMENT);
    const int nb = nl_chunks(in.numUV, in.numP, in.numU);
    hipLaunchKernelGGL(k_nl_count, dim3(nb, 5), dim3(BS), 0, s, in, part);
    CUIBM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_nl_write, dim3(nb, 5), dim3(BS), 0, s, in, part, out[0], out[1], out[2], out[3], out[4],
                       counts);
    CUIBM_LAUNCH_CHECK();
}

void launch_zero_rows(int *tags, const BoxDev &box, int rowlen, int kmax, int off, int nrows, long len,
                      hipStream_t s) {
    // rows j_start .. j_end (inclusive) below nrows, and no more than the reference's thread
    // count over the box width
    const int nthr0 = ((int)((box.w - 0.5) / 256) + 1) * 256;
    const int jlast = std::min({box.j_start + box.h, nrows - 1, box.j_start + nthr0 - 1});
    const int rows = jlast - box.j_start + 1;
    if (rows <= 0 || box.w <= 0) return;
    const int nseg = std::max(kmax, box.i_start + box.w - 1) - (box.i_start - 1) + 1;
    const size_t lds = (size_t)nseg * sizeof(int);
    if (lds > 65536) throw Error("tagging: grid row of " + std::to_string(nseg) + " nodes exceeds the 16384-node row limit");
    hipLaunchKernelGGL(k_zero_rows, dim3(rows), dim3(BS), lds, s, tags, box, rowlen, kmax, off, len);
    CUIBM_LAUNCH_CHECK();
}

void launch_gn_vel_coef(const GridDev &g, const IbmGeom &m, Stencils &st, hipStream_t s) {
    if (!st.n) return;
    hipLaunchKernelGGL(k_gn_vel_coef, dim3(nblk(st.n)), dim3(BS), 0, s, g, m, st.n, st.node.p(), st.idx.p(), st.idx0.p(),
                       st.coef.p());
    CUIBM_LAUNCH_CHECK();
}
void launch_hn_vel_coef(const GridDev &g, const IbmGeom &m, Stencils &st, hipStream_t s) {
    if (!st.n) return;
    hipLaunchKernelGGL(k_hn_vel_coef, dim3(nblk(st.n)), dim3(BS), 0, s, g, m, st.n, st.node.p(), st.idx.p(), st.idx0.p(),
                       st.coef.p(),
                       st.extra.p());
    CUIBM_LAUNCH_CHECK();
}
void launch_hn_p_coef(const GridDev &g, const IbmGeom &m, Stencils &st, hipStream_t s) {
    if (!st.n) return;
    hipLaunchKernelGGL(k_hn_p_coef, dim3(nblk(st.n)), dim3(BS), 0, s, g, m, st.n, st.node.p(), st.idx.p(), st.idx0.p(),
                       st.coef.p(),
                       st.nrm.p(), st.extra.p());
    CUIBM_LAUNCH_CHECK();
}
void launch_gn_p_coef(const GridDev &g, const IbmGeom &m, Stencils &st, hipStream_t s) {
    if (!st.n) return;
    hipLaunchKernelGGL(k_gn_p_coef, dim3(nblk(st.n)), dim3(BS), 0, s, g, m, st.n, st.node.p(), st.idx.p(), st.idx0.p(),
                       st.coef.p(),
                       st.nrm.p(), st.extra.p());
    CUIBM_LAUNCH_CHECK();
}

void launch_gn_vel_apply(double *f, Stencils &st, int numU, double uB, double vB, hipStream_t s) {
    if (!st.n) return;
    hipLaunchKernelGGL(k_gn_vel_eval, dim3(nblk(st.n)), dim3(BS), 0, s, f, st.n, st.node.p(), st.idx.p(), st.coef.p(),
                       st.tmp.p(), numU, uB, vB, st.ub.empty() ? nullptr : (const double *)st.ub.p());
    CUIBM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scatter, dim3(nblk(st.n)), dim3(BS), 0, s, f, st.n, st.node.p(), st.tmp.p());
    CUIBM_LAUNCH_CHECK();
}
void launch_set_inside(double *f, const int *solid, int nsolid, int numU, double uB, double vB, hipStream_t s,
                       const double *ubn) {
    if (!nsolid) return;
    hipLaunchKernelGGL(k_set_inside, dim3(nblk(nsolid)), dim3(BS), 0, s, f, solid, nsolid, numU, uB, vB, ubn);
    CUIBM_LAUNCH_CHECK();
}
void launch_hn_vel_apply(const double *f, double *ustar, const Stencils &st, int numU, double uB, double vB,
                         hipStream_t s) {
    if (!st.n) return;
    hipLaunchKernelGGL(k_hn_vel_eval, dim3(nblk(st.n)), dim3(BS), 0, s, f, ustar, st.n, st.node.p(), st.idx.p(),
                       st.coef.p(), numU, uB, vB, st.ub.empty() ? nullptr : (const double *)st.ub.p());
    CUIBM_LAUNCH_CHECK();
}
void launch_weight(double *f, const double *star, const Stencils &st, hipStream_t s) {
    if (!st.n) return;
    hipLaunchKernelGGL(k_weight, dim3(nblk(st.n)), dim3(BS), 0, s, f, star, st.n, st.node.p(), st.extra.p());
    CUIBM_LAUNCH_CHECK();
}
void launch_hn_p_apply(const double *p, double *pstar, const Stencils &st, double dudt, double dvdt, hipStream_t s) {
    if (!st.n) return;
    hipLaunchKernelGGL(k_hn_p_eval, dim3(nblk(st.n)), dim3(BS), 0, s, p, pstar, st.n, st.node.p(), st.idx.p(),
                       st.coef.p(), st.nrm.p(), dudt, dvdt);
    CUIBM_LAUNCH_CHECK();
}
void launch_gn_p_apply(double *p, const Stencils &st, double dudt, double dvdt, hipStream_t s) {
    if (!st.n) return;
    hipLaunchKernelGGL(k_gn_p_eval, dim3(nblk(st.n)), dim3(BS), 0, s, p, st.n, st.node.p(), st.idx.p(), st.coef.p(),
                       st.nrm.p(), st.extra.p(), dudt, dvdt);
    CUIBM_LAUNCH_CHECK();
}

constexpr int FORCE_BLOCKS = 64;
void launch_force(const double *u, const double *uold, const double *p, const int *ghostUV, const GridDev &g,
                  double nu, double dt, int I, int J, int ncx, int ncy, double *part, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_force, dim3(FORCE_BLOCKS), dim3(BS), 0, s, u, uold, p, ghostUV, g, nu, dt, I, J, ncx, ncy, part);
    CUIBM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_force_finish, dim3(1), dim3(64), 0, s, part, FORCE_BLOCKS, out);
    CUIBM_LAUNCH_CHECK();
}

// ---- lockstep batches: the force of every case, BFORCE_BLOCKS blocks per case ----

__global__ void __launch_bounds__(BS) k_b_force(const BCase *__restrict__ tab, const double *__restrict__ u,
                                                const double *__restrict__ uold, const double *__restrict__ p,
                                                double *__restrict__ part) {
    const int c = blockIdx.x / BFORCE_BLOCKS, lb = blockIdx.x % BFORCE_BLOCKS;
    const BCase &C = tab[c];
    const long total = force_items(C.fncx, C.fncy);
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (long t = lb * (long)BS + threadIdx.x; t < total; t += (long)BFORCE_BLOCKS * BS)
        force_term(t, u + C.offUV, uold + C.offUV, p + C.offP, C.ghostUV, C.g, C.nu, C.dt, C.fI, C.fJ, C.fncx, C.fncy,
                   acc);
    __shared__ double sh[6][BS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int k = 0; k < 6; k++) {
        const double v = wsum(acc[k]);
        if (lane == 0) sh[k][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        part[(6 * c + k) * BFORCE_BLOCKS + lb] = (sh[k][0] + sh[k][1]) + (sh[k][2] + sh[k][3]);
    }
}

__global__ void k_b_force_finish(int ncases, const double *__restrict__ part, double *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 6 * ncases) return;
    double s = 0;
    for (int b = 0; b < BFORCE_BLOCKS; b++) s += part[t * BFORCE_BLOCKS + b];
    out[t] = s;
}

void launch_b_force(const BCase *tab, int ncases, const double *u, const double *uold, const double *p, double *part,
                    double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_b_force, dim3(ncases * BFORCE_BLOCKS), dim3(BS), 0, s, tab, u, uold, p, part);
    CUIBM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_b_force_finish, dim3((6 * ncases + BS - 1) / BS), dim3(BS), 0, s, ncases, part, out);
    CUIBM_LAUNCH_CHECK();
}

}  // namespace cuibm
