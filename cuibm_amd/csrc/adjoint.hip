// Discrete adjoint of the Luo external-IBM time step for a static body (no reference
// counterpart: the reference has no adjoint).
//
// Objective J = (1 / (n1 - n0 + 1)) sum_{n=n0..n1} C_D^n with C_D = 2 Fx, the
// control-volume drag of luo_base::calculateForce; parameters: the per-point body
// velocities (ub_k, vb_k) and the angular velocity of the solid nodes' rigid rotation.
// With fixed tags, one step maps (u, N, xp) -> (u', N', xp') through
//   xp' = Robin(xp, u)                          convective outflow
//   u1  = S(G(u))            + body terms       ghost-node interpolation, solid nodes
//   ustar = H u1             + body terms       hybrid-node interpolation
//   rhs1 = f(u1, Nold = N, xp'), N' = N(u1, xp')  (quadratic in u1: convection)
//   uh0 = LHS1^-1 rhs1,  uh1 = (1 - A) uh0 + A ustar
//   uh2 = G(uh1)             + body terms
//   p0  = LHS2^-1 (D uh2 + d(xp')),  p1 = Wp p0   (pressure blending and ghost values)
//   u2  = uh2 - dt Grad p1,  u' = S(G(u2)) + body terms
//   Fx  = F(u', uold = u1, p1)
// and the adjoint runs these backwards: transposed solves with LHS1^T (Jacobi) and LHS2^T
// (the forward V-cycle conjugated by the cell areas), transposed interpolation stencils as gathers
// through inverse maps built once per tagging (no atomics), the transposed divergence and
// gradient, and the force gradient. The transposed Jacobian of f and N is taken from the
// forward kernel itself: both are quadratic in u1 and linear in xp', so the central
// difference (f(u1 + d) - f(u1 - d)) / 2 is exactly the Jacobian times d; with d the
// indicator of one of 18 node colours (i mod 3, j mod 3 per component), no row of the
// stencil touches two nodes of one colour, and each node gathers its column from the
// rows within one cell of it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>

#include "solver.hpp"

namespace cuibm {

namespace {

int blocks_for(long n) { return (int)std::min<long>((n + BS - 1) / BS, 65536); }

// transposed interpolation stencil: for each target node (a corner of some stencil
// node), the stencil slots that read it; out[target] += sum w * lam[node of slot]
struct TMap {
    int nrows = 0;
    DArray<int> row, ptr, src;
    DArray<double> w;
};

void build_tmap(const std::vector<int> &node, const std::vector<int> &idx, const std::vector<double> &coef,
                double sign, TMap &m, hipStream_t s) {
    std::vector<std::pair<int, int>> e;  // (target, slot)
    for (size_t t = 0; t < node.size(); t++)
        for (int l = 0; l < 4; l++)
            if (idx[4 * t + l] >= 0) e.push_back({idx[4 * t + l], (int)(4 * t + l)});
    std::sort(e.begin(), e.end());
    std::vector<int> row, ptr{0}, src;
    std::vector<double> w;
    for (size_t k = 0; k < e.size(); k++) {
        if (k == 0 || e[k].first != e[k - 1].first) {
            if (k > 0) ptr.push_back((int)src.size());
            row.push_back(e[k].first);
        }
        src.push_back(node[e[k].second / 4]);
        w.push_back(sign * coef[e[k].second]);
    }
    if (!e.empty()) ptr.push_back((int)src.size());
    m.nrows = (int)row.size();
    if (row.empty()) {
        row.push_back(0);
        ptr.push_back(0);
        src.push_back(0);
        w.push_back(0.0);
    }
    m.row.upload(row, s);
    m.ptr.upload(ptr, s);
    m.src.upload(src, s);
    m.w.upload(w, s);
}

__global__ void k_tmap_gather(int nrows, const int *__restrict__ row, const int *__restrict__ ptr,
                              const int *__restrict__ src, const double *__restrict__ w,
                              const double *__restrict__ lam, double *__restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    double a = 0;
    for (int k = ptr[r]; k < ptr[r + 1]; k++) a += w[k] * lam[src[k]];
    out[row[r]] += a;
}

void tmap_gather(const TMap &m, const double *lam, double *out, hipStream_t s) {
    if (!m.nrows) return;
    hipLaunchKernelGGL(k_tmap_gather, dim3(nblk(m.nrows)), dim3(BS), 0, s, m.nrows, m.row.p(), m.ptr.p(), m.src.p(),
                       m.w.p(), lam, out);
    CUIBM_LAUNCH_CHECK();
}

// zero the listed rows of out; sens[t] += c[t] * lam[list[t]] (c null: 1; sens null: skipped)
__global__ void k_zero_list(int m, const int *__restrict__ list, const double *__restrict__ lam,
                            const double *__restrict__ c, double *__restrict__ sens, double *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const int i = list[t];
    if (sens) sens[t] += (c ? c[t] : 1.0) * lam[i];
    out[i] = 0.0;
}

// hybrid blending transposed: out[h] = (1 - a) lam[h] (in place on out), star[h] = a lam[h]
__global__ void k_blend_T(int m, const int *__restrict__ node, const double *__restrict__ alpha,
                          const double *__restrict__ lam, double *__restrict__ out, double *__restrict__ star) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const int h = node[t];
    const double a = alpha[t];
    star[h] = a * lam[h];
    out[h] = (1 - a) * lam[h];
}

// sens[t] += c[t] * lam[node[t]]
__global__ void k_sens(int m, const int *__restrict__ node, const double *__restrict__ c,
                       const double *__restrict__ lam, double *__restrict__ sens) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) sens[t] += c[t] * lam[node[t]];
}

// (-dt Grad)^T: lp[ip] += sum over the faces of cell ip of d u_face / d p[ip] * lu[face]
__global__ void k_grad_T(GridDev g, double dt, const double *__restrict__ lu, double *__restrict__ lp) {
    const int nx = g.nx, ny = g.ny, numU = (nx - 1) * ny;
    for (int ip = blockIdx.x * blockDim.x + threadIdx.x; ip < nx * ny; ip += gridDim.x * blockDim.x) {
        const int I = ip % nx, J = ip / nx;
        double a = 0;
        if (I < nx - 1) a += dt / (0.5 * g.dx[I + 1] + 0.5 * g.dx[I]) * lu[J * (nx - 1) + I];      // east face
        if (I > 0) a -= dt / (0.5 * g.dx[I] + 0.5 * g.dx[I - 1]) * lu[J * (nx - 1) + I - 1];      // west face
        if (J < ny - 1) a += dt / (0.5 * g.dy[J + 1] + 0.5 * g.dy[J]) * lu[numU + J * nx + I];    // north face
        if (J > 0) a -= dt / (0.5 * g.dy[J] + 0.5 * g.dy[J - 1]) * lu[numU + (J - 1) * nx + I];   // south face
        lp[ip] += a;
    }
}

// D^T of the divergence right-hand side (NavierStokes/kernels/intermediatePressure.cu:30-71):
// lu[face] += sum over the two cells of the face of d rhs2 / d u_face * lr[cell]
__global__ void k_div_T(GridDev g, const double *__restrict__ lr, double *__restrict__ lu) {
    const int nx = g.nx, ny = g.ny, numU = (nx - 1) * ny, numUV = numU + nx * (ny - 1);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < numUV; i += gridDim.x * blockDim.x) {
        if (i < numU) {
            const int I = i % (nx - 1), J = i / (nx - 1), ip = J * nx + I;
            lu[i] += -lr[ip] / g.dx[I] + lr[ip + 1] / g.dx[I + 1];
        } else {
            const int k = i - numU, I = k % nx, J = k / nx, ip = J * nx + I;
            lu[i] += -lr[ip] / g.dy[J] + lr[ip + nx] / g.dy[J + 1];
        }
    }
}

// colour of a velocity node: (I mod 3) + 3 (J mod 3), plus 9 for v
__device__ __forceinline__ int vel_colour(int i, int nx, int numU, int &I, int &J) {
    if (i < numU) {
        I = i % (nx - 1);
        J = i / (nx - 1);
        return (I % 3) + 3 * (J % 3);
    }
    const int k = i - numU;
    I = k % nx;
    J = k / nx;
    return 9 + (I % 3) + 3 * (J % 3);
}

// up = u + s d, um = u - s d with d the indicator of colour c
__global__ void k_colour_pm(GridDev g, const double *__restrict__ u, double *__restrict__ up, double *__restrict__ um,
                            int c, double s) {
    const int nx = g.nx, ny = g.ny, numU = (nx - 1) * ny, numUV = numU + nx * (ny - 1);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < numUV; i += gridDim.x * blockDim.x) {
        int I, J;
        const double d = vel_colour(i, nx, numU, I, J) == c ? s : 0.0;
        up[i] = u[i] + d;
        um[i] = u[i] - d;
    }
}

// sum over the velocity rows within one cell of (I, J) of w(row)
template <class F>
__device__ __forceinline__ double window_sum(int nx, int ny, int numU, int I, int J, F w) {
    double a = 0;
    for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++) {
            const int Ii = I + di, Jj = J + dj;
            if (Ii >= 0 && Ii < nx - 1 && Jj >= 0 && Jj < ny) a += w(Jj * (nx - 1) + Ii);  // u row
            if (Ii >= 0 && Ii < nx && Jj >= 0 && Jj < ny - 1) a += w(numU + Jj * nx + Ii);   // v row
        }
    return a;
}

// for every velocity node j of colour c:
// out[j] += sum_{rows i near j} ((Rp - Rm)[i] lr[i] + (Np - Nm)[i] lN[i]) / (2 s)
__global__ void k_colour_gather(GridDev g, int c, double s, const double *__restrict__ Rp,
                                const double *__restrict__ Rm, const double *__restrict__ Np,
                                const double *__restrict__ Nm, const double *__restrict__ lr,
                                const double *__restrict__ lN, double *__restrict__ out) {
    const int nx = g.nx, ny = g.ny, numU = (nx - 1) * ny, numUV = numU + nx * (ny - 1);
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < numUV; j += gridDim.x * blockDim.x) {
        int I, J;
        if (vel_colour(j, nx, numU, I, J) != c) continue;
        const double a = window_sum(nx, ny, numU, I, J, [&](int i) {
            return (Rp[i] - Rm[i]) * lr[i] + (Np[i] - Nm[i]) * lN[i];
        });
        out[j] += a / (2 * s);
    }
}

// the same for the x+ wall values xp (u part J in [0, ny), v part ny + J), colour J mod 3
// per part; xp[J] sits at (nx - 1, J)
__global__ void k_xp_pm(int ny, const double *__restrict__ xp, double *__restrict__ xpp, double *__restrict__ xpm,
                        int c, double s) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 2 * ny - 1) return;
    const int part = k < ny ? 0 : 1, J = k < ny ? k : k - ny;
    const double d = (3 * part + J % 3) == c ? s : 0.0;
    xpp[k] = xp[k] + d;
    xpm[k] = xp[k] - d;
}
__global__ void k_xp_gather(GridDev g, int c, double s, const double *__restrict__ Rp, const double *__restrict__ Rm,
                            const double *__restrict__ Np, const double *__restrict__ Nm,
                            const double *__restrict__ lr, const double *__restrict__ lN, double *__restrict__ out) {
    const int nx = g.nx, ny = g.ny, numU = (nx - 1) * ny;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 2 * ny - 1) return;
    const int part = k < ny ? 0 : 1, J = k < ny ? k : k - ny;
    if ((3 * part + J % 3) != c) return;
    // the rows that read xp[J] lie in the last columns, within one row of J
    const double a = window_sum(nx, ny, numU, nx - 2, J, [&](int i) {
        return (Rp[i] - Rm[i]) * lr[i] + (Np[i] - Nm[i]) * lN[i];
    });
    out[k] += a / (2 * s);
}

// the east cells (nx - 1, J) of the divergence read the wall value xp[J]:
// d rhs2 / d xp[J] = -1 / dx[nx - 1]
__global__ void k_div_xp_T(GridDev g, const double *__restrict__ lr, double *__restrict__ lxp) {
    const int J = blockIdx.x * blockDim.x + threadIdx.x;
    if (J < g.ny) lxp[J] += -lr[J * g.nx + g.nx - 1] / g.dx[g.nx - 1];
}

// Robin transposed: lx_in = (1 - beta) lxp; lu[last u of row J] += beta lxp[J],
// lu[last v of row J] += beta lxp[ny + J]
__global__ void k_robin_T(GridDev g, double dt, const double *__restrict__ lxp, double *__restrict__ lxo,
                          double *__restrict__ lu) {
    const int nx = g.nx, ny = g.ny, numU = (nx - 1) * ny;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const double beta = 1.0 * dt / g.dx[nx - 1];
    if (k < ny) {
        lxo[k] = (1 - beta) * lxp[k];
        lu[k * (nx - 1) + nx - 2] += beta * lxp[k];
    } else if (k < 2 * ny - 1) {
        const int J = k - ny;
        lxo[k] = (1 - beta) * lxp[k];
        lu[numU + J * nx + nx - 1] += beta * lxp[k];
    }
}

// out += a * x
__global__ void k_axpy(long n, double a, const double *__restrict__ x, double *__restrict__ out) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] += a * x[i];
}

// Gradient of c * Fx (FadlunModified/kernels/calculateForce.cu:32-118, the drag terms of
// luo_base::calculateForce) with respect to u (current velocity), uold and p, added into
// gu, guo and gp. One thread per node of the region around the control volume; every
// term the node appears in is evaluated where it is, so the sums need no atomics.
struct BoxArgs {
    int I, J, ncx, ncy;
};

__global__ void k_force_grad(GridDev g, double nu, double dt, BoxArgs b, double c, const double *__restrict__ u,
                             const int *__restrict__ tags, double *__restrict__ gu, double *__restrict__ guo,
                             double *__restrict__ gp) {
    const int nx = g.nx, ny = g.ny, numU = (nx - 1) * ny;
    const double *dx = g.dx, *dy = g.dy;
    const int I = b.I, J = b.J, ncx = b.ncx, ncy = b.ncy;
    // region: columns I - 2 .. I + ncx + 1, rows J - 2 .. J + ncy + 1, three fields
    const int W = ncx + 4, H = ncy + 4, per = W * H;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * per) return;
    const int f = t / per, r = t % per;
    const int col = I - 2 + r % W, row = J - 2 + r / W;
    auto U = [&](int ii, int jj) { return u[jj * (nx - 1) + ii]; };
    auto V = [&](int ii, int jj) { return u[numU + jj * nx + ii]; };
    if (f == 0) {  // u node (col, row)
        if (col < 0 || col >= nx - 1 || row < 0 || row >= ny) return;
        double a = 0, ao = 0;
        // dragLeftRight, rows J .. J + ncy - 1: left pair (I - 1, I), right pair (I + ncx - 1, I + ncx)
        if (row >= J && row < J + ncy) {
            const double dyj = dy[row];
            const double aL = U(I - 1, row) + U(I, row), aR = U(I + ncx, row) + U(I + ncx - 1, row);
            if (col == I - 1) a += aL / 2 * dyj + nu * dyj / dx[I];
            if (col == I) a += aL / 2 * dyj - nu * dyj / dx[I];
            if (col == I + ncx) a += -aR / 2 * dyj + nu * dyj / dx[I + ncx];
            if (col == I + ncx - 1) a += -aR / 2 * dyj - nu * dyj / dx[I + ncx];
        }
        // dragBottomTop, columns I - 1 + idx for idx = 0 .. ncx
        if (col >= I - 1 && col <= I - 1 + ncx) {
            const int idx = col - (I - 1);
            const double w = 0.5 * (dx[I + idx] + dx[I + idx - 1]);
            const int jt = J + ncy, jb = J;
            const double bT = V(I + idx, jt - 1) + V(I + idx - 1, jt - 1);
            const double bB = V(I + idx, jb - 1) + V(I + idx - 1, jb - 1);
            const double dyT = dy[J + ncy] + dy[J + ncy - 1], dyB = dy[J] + dy[J - 1];
            if (row == jt) a += -w * (0.25 * bT - nu / (2 * dyT));
            if (row == jt - 1) a += -w * (0.25 * bT + nu / (2 * dyT));
            if (row == jb) a += -w * (-0.25 * bB + nu / (2 * dyB));
            if (row == jb - 1) a += -w * (-0.25 * bB - nu / (2 * dyB));
        }
        // dragUnsteady: rows J .. J + ncy - 1, columns I - 1 .. I - 1 + ncx
        if (row >= J && row < J + ncy && col >= I - 1 && col <= I - 1 + ncx) {
            const int i = col - (I - 1);
            const int Iu = row * (nx - 1) + col;
            const double m = tags[Iu] == -1 ? 1.0 : 0.0;
            const double k = m * dy[row] * 0.5 * (dx[I + i] + dx[I - 1 + i]) / dt;
            a += -k;
            ao += k;
        }
        gu[row * (nx - 1) + col] += c * a;
        guo[row * (nx - 1) + col] += c * ao;
    } else if (f == 1) {  // v node (col, row)
        if (col < 0 || col >= nx || row < 0 || row >= ny - 1) return;
        double a = 0;
        // dragBottomTop: v row J + ncy - 1 (top) and J - 1 (bottom); node col is the first
        // v of the pair of term idx = col - I and the second of term idx = col - I + 1
        if (row == J + ncy - 1 || row == J - 1) {
            for (int second = 0; second < 2; second++) {
                const int idx = col - I + second;
                if (idx < 0 || idx > ncx) continue;
                const double w = 0.5 * (dx[I + idx] + dx[I + idx - 1]);
                const double dxw = dx[I + idx] + dx[I + idx - 1];
                const int cu = I - 1 + idx;
                // the node is v1 of the pair (second = 0) or v2: sg = -d(v1 - v2)/dv
                const double sg = second ? 1.0 : -1.0;
                if (row == J + ncy - 1) {
                    const double aT = U(cu, J + ncy) + U(cu, J + ncy - 1);
                    a += -w * (0.25 * aT + sg * nu / (2 * dxw));
                }
                if (row == J - 1) {
                    const double aB = U(cu, J) + U(cu, J - 1);
                    a += -w * (-0.25 * aB - sg * nu / (2 * dxw));
                }
            }
        }
        gu[numU + row * nx + col] += c * a;
    } else {  // p cell (col, row): dragLeftRight pressure, cells I and I + ncx of rows J .. J + ncy - 1
        if (col < 0 || col >= nx || row < 0 || row >= ny) return;
        if (row < J || row >= J + ncy) return;
        double a = 0;
        if (col == I + ncx) a += -dy[row];
        if (col == I) a += dy[row];
        gp[row * nx + col] += c * a;
    }
}

// t = r / w or t = r * w, elementwise; returns at once while the Krylov stop code is set
__global__ void k_wscale(long n, const double *__restrict__ r, const double *__restrict__ w, double *__restrict__ t,
                         int divide, const double *__restrict__ stop) {
    if (stop && *stop != 0.0) return;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        t[i] = divide ? r[i] / w[i] : r[i] * w[i];
}

// A preconditioner for LHS2^T from the forward one: with W the cell areas, W LHS2 is
// symmetric (each coupling is -dt dy_J / ((dx_I + dx_I+1) / 2) from either side, and the
// pins are diagonal), so LHS2^T = W LHS2 W^-1 and W M W^-1 approximates LHS2^-T wherever M
// approximates LHS2^-1. The forward V-cycle (stencil finest level, FP32 or FP64) then
// serves the transposed solve, whose operator stays the transposed CSR matrix.
class ConjugatedPrecond : public Preconditioner {
  public:
    ConjugatedPrecond(Preconditioner &M, const double *w, long n) : M_(M), w_(w), n_(n), t_(n), u_(n) {}
    void apply(const double *r, double *z, hipStream_t s) override {
        const int g = blocks_for(n_);
        hipLaunchKernelGGL(k_wscale, dim3(g), dim3(BS), 0, s, n_, r, w_, t_.p(), 1, stop_);
        CUIBM_LAUNCH_CHECK();
        M_.apply(t_.p(), u_.p(), s);
        hipLaunchKernelGGL(k_wscale, dim3(g), dim3(BS), 0, s, n_, u_.p(), w_, z, 0, stop_);
        CUIBM_LAUNCH_CHECK();
    }
    void set_stop(const double *stop) override {
        stop_ = stop;
        M_.set_stop(stop);
    }
    std::string name() const override { return "W " + M_.name() + " W^-1"; }
    double bytes_per_apply() const override { return M_.bytes_per_apply() + 2 * 24.0 * n_; }

  private:
    Preconditioner &M_;
    const double *w_;
    long n_;
    DArray<double> t_, u_;
};

__global__ void k_cell_area(GridDev g, double *__restrict__ w) {
    for (int ip = blockIdx.x * blockDim.x + threadIdx.x; ip < g.nx * g.ny; ip += gridDim.x * blockDim.x)
        w[ip] = g.dx[ip % g.nx] * g.dy[ip / g.nx];
}

}  // namespace

struct luoIBM::Adjoint {
    DCsr<double> L1T, L2T;
    std::unique_ptr<JacobiPrecond> M1;
    std::unique_ptr<ConjugatedPrecond> M2;
    DArray<double> area;  // cell areas (W)
    std::unique_ptr<Bicgstab> k1, k2;
    std::unique_ptr<CsrOp> op1, op2;
    TMap gnT, hnT, hnpT, gnpT;           // transposed stencils
    DArray<double> cGN, cHN;             // body-velocity coefficients per ghost / hybrid node
    DArray<double> sGN, sHN, sSolid;     // accumulated sensitivities to the per-node body velocities
};

luoIBM::~luoIBM() = default;

static HCsr transpose_csr(const HCsr &A) {
    HCsr T;
    T.n = A.ncols;
    T.ncols = A.n;
    T.rp.assign(T.n + 1, 0);
    for (int c : A.ci) T.rp[c + 1]++;
    for (int i = 0; i < T.n; i++) T.rp[i + 1] += T.rp[i];
    T.ci.resize(A.ci.size());
    T.v.resize(A.v.size());
    std::vector<int> pos(T.rp.begin(), T.rp.end() - 1);
    for (int r = 0; r < A.n; r++)
        for (int k = A.rp[r]; k < A.rp[r + 1]; k++) {
            const int q = pos[A.ci[k]]++;
            T.ci[q] = r;
            T.v[q] = A.v[k];
        }
    return T;
}

static HCsr download(const DCsr<double> &A, hipStream_t s) {
    HCsr h;
    h.n = A.n;
    h.ncols = A.ncols;
    h.rp = A.rp.to_host(s);
    h.ci = A.ci.to_host(s);
    h.v = A.v.to_host(s);
    return h;
}

void luoIBM::adjointRun(int nsteps, int n0, int n1, int every, bool backward, double *Jout, double *gub, double *gvb,
                        double *gomega, double *stats) {
    hipStream_t s = stream_;
    if (nsteps < 1 || n0 < 1 || n1 < n0 || n1 > nsteps || every < 1)
        throw Error("adjoint: need 1 <= n0 <= n1 <= nsteps and every >= 1", ERR_ARGUMENT);
    if (moving || db->VIV != 0) throw Error("adjoint: static bodies only", ERR_UNSUPPORTED);
    if (timeStep != 0) throw Error("adjoint: start from an initialised solver at time step 0", ERR_ARGUMENT);
    if (!perPointUB) throw Error("adjoint: set the body velocities first (cuibm_set_body_velocity)", ERR_ARGUMENT);
    const double c = 2.0 / (n1 - n0 + 1);  // dJ/dFx at a step of the window
    const long nuv = numUV, np = numP, nxp = 2 * ny - 1;
    using clk = std::chrono::steady_clock;
    auto secs = [](clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); };

    // ---- forward run with checkpoints of (u, uhat, N, p, xp)
    const int ncp = nsteps / every + 1;
    auto slot = [&](int k) { return (size_t)k * (3 * nuv + np + nxp); };
    DArray<double> cp((size_t)ncp * (3 * nuv + np + nxp));
    auto save = [&](double *dst) {
        CUIBM_HIP(hipMemcpyAsync(dst, u.p(), nuv * 8, hipMemcpyDeviceToDevice, s));
        CUIBM_HIP(hipMemcpyAsync(dst + nuv, uhat.p(), nuv * 8, hipMemcpyDeviceToDevice, s));
        CUIBM_HIP(hipMemcpyAsync(dst + 2 * nuv, N.p(), nuv * 8, hipMemcpyDeviceToDevice, s));
        CUIBM_HIP(hipMemcpyAsync(dst + 3 * nuv, pressure.p(), np * 8, hipMemcpyDeviceToDevice, s));
        CUIBM_HIP(hipMemcpyAsync(dst + 3 * nuv + np, bc[XPLUS].p(), nxp * 8, hipMemcpyDeviceToDevice, s));
    };
    auto restore = [&](const double *src) {
        CUIBM_HIP(hipMemcpyAsync(u.p(), src, nuv * 8, hipMemcpyDeviceToDevice, s));
        CUIBM_HIP(hipMemcpyAsync(uhat.p(), src + nuv, nuv * 8, hipMemcpyDeviceToDevice, s));
        CUIBM_HIP(hipMemcpyAsync(N.p(), src + 2 * nuv, nuv * 8, hipMemcpyDeviceToDevice, s));
        CUIBM_HIP(hipMemcpyAsync(pressure.p(), src + 3 * nuv, np * 8, hipMemcpyDeviceToDevice, s));
        CUIBM_HIP(hipMemcpyAsync(bc[XPLUS].p(), src + 3 * nuv + np, nxp * 8, hipMemcpyDeviceToDevice, s));
    };
    const auto t0 = clk::now();
    const double it1a = itTotal1, it2a = itTotal2;
    save(cp.p() + slot(0));
    double J = 0;
    for (int n = 1; n <= nsteps; n++) {
        stepTime();
        readForces();
        if (n >= n0 && n <= n1) J += c * forceX;
        if (n % every == 0) save(cp.p() + slot(n / every));
    }
    CUIBM_HIP(hipStreamSynchronize(s));
    if (Jout) *Jout = J;
    if (stats) {
        stats[0] = secs(t0);
        stats[2] = itTotal2 - it2a;
        stats[3] = itTotal1 - it1a;
    }
    if (!backward) return;

    // ---- set-up: transposed operators and stencils (fixed tags)
    const auto t1 = clk::now();
    if (!adj_) adj_.reset(new Adjoint);
    Adjoint &A = *adj_;
    {
        const HCsr h1 = transpose_csr(download(LHS1, s)), h2 = transpose_csr(download(LHS2, s));
        upload_csr(h1, A.L1T, s);
        upload_csr(h2, A.L2T, s);
        A.op1.reset(new CsrOp(A.L1T));
        A.op2.reset(new CsrOp(A.L2T));
        A.M1.reset(new JacobiPrecond(A.L1T, s));
        A.area.alloc(numP);
        hipLaunchKernelGGL(k_cell_area, dim3(blocks_for(numP)), dim3(BS), 0, s, grid(), A.area.p());
        CUIBM_LAUNCH_CHECK();
        A.M2.reset(new ConjugatedPrecond(*PC2, A.area.p(), numP));
        A.k1.reset(new Bicgstab(numUV));
        A.k2.reset(new Bicgstab(numP));
        auto host_st = [&](const Stencils &st, std::vector<int> &node, std::vector<int> &idx, std::vector<double> &cf) {
            node = st.n ? st.node.to_host(s) : std::vector<int>();
            node.resize(st.n);
            idx = st.n ? st.idx.to_host(s) : std::vector<int>();
            idx.resize(4 * st.n);
            cf = st.n ? st.coef.to_host(s) : std::vector<double>();
            cf.resize(4 * st.n);
        };
        std::vector<int> nd, ix;
        std::vector<double> cf;
        host_st(gnV, nd, ix, cf);
        build_tmap(nd, ix, cf, -1.0, A.gnT, s);
        std::vector<double> cg(std::max(gnV.n, 1), 0.0);
        for (int t = 0; t < gnV.n; t++) {
            double b = 0;
            for (int l = 0; l < 4; l++)
                if (ix[4 * t + l] < 0) b += cf[4 * t + l];
            cg[t] = 2 - b;
        }
        A.cGN.upload(cg, s);
        host_st(hnV, nd, ix, cf);
        build_tmap(nd, ix, cf, 1.0, A.hnT, s);
        std::vector<double> ch(std::max(hnV.n, 1), 0.0);
        for (int t = 0; t < hnV.n; t++)
            for (int l = 0; l < 4; l++)
                if (ix[4 * t + l] < 0) ch[t] += cf[4 * t + l];
        A.cHN.upload(ch, s);
        host_st(hnP, nd, ix, cf);
        build_tmap(nd, ix, cf, 1.0, A.hnpT, s);
        host_st(gnP, nd, ix, cf);
        build_tmap(nd, ix, cf, 1.0, A.gnpT, s);
        A.sGN.alloc(std::max(gnV.n, 1));
        A.sHN.alloc(std::max(hnV.n, 1));
        A.sSolid.alloc(std::max(nsolid, 1));
        launch_fill(A.sGN.p(), 0.0, std::max(gnV.n, 1), s);
        launch_fill(A.sHN.p(), 0.0, std::max(hnV.n, 1), s);
        launch_fill(A.sSolid.p(), 0.0, std::max(nsolid, 1), s);
    }
    CUIBM_HIP(hipStreamSynchronize(s));
    const double t_setup = secs(t1);
    double t_recompute = 0;

    // ---- adjoint state and work arrays
    DArray<double> lu(nuv), lN(nuv), lxp(nxp);
    DArray<double> lm(nuv), lu2(nuv), lp1(np), lq(np), lpst(np), lp0(np), lr2(np), luh(nuv), lst(nuv), lr1(nuv);
    DArray<double> lu1(nuv), lNo(nuv), lxpP(nxp);
    DArray<double> u1(nuv), up(nuv), um(nuv), Rp(nuv), Rm(nuv), Np(nuv), Nm(nuv), xpP(nxp), xpp(nxp), xpm(nxp);
    for (auto *a : {&lu, &lN, &lr1}) launch_fill(a->p(), 0.0, nuv, s);
    launch_fill(lr2.p(), 0.0, np, s);
    launch_fill(lxp.p(), 0.0, nxp, s);
    // per-step records of a segment: u_in, N_in, xp_in; and the state after the segment
    DArray<double> rec((size_t)(every + 1) * (2 * nuv + nxp));
    auto rslot = [&](int k) { return rec.p() + (size_t)k * (2 * nuv + nxp); };
    const GridDev gd = grid();
    const BoxArgs box{startI, startJ, numCellsX, numCellsY};
    const double *solidub = solidUB.p();
    auto zero = [&](DArray<double> &a, long n) { launch_fill(a.p(), 0.0, n, s); };
    auto copy = [&](double *dst, const double *src, long n) { launch_copy(dst, src, n, s); };
    // S^T G^T applied to lam: out = (G^T (S^T lam)), with the sensitivities of the solid
    // and ghost body values; `with_solid` selects S (set-inside) before G
    auto SG_T = [&](const double *lam, double *out, bool with_solid) {
        copy(lm.p(), lam, nuv);
        if (with_solid && nsolid)
            hipLaunchKernelGGL(k_zero_list, dim3(nblk(nsolid)), dim3(BS), 0, s, nsolid, solid.p(), lam, nullptr,
                               A.sSolid.p(), lm.p());
        copy(out, lm.p(), nuv);
        if (gnV.n)
            hipLaunchKernelGGL(k_zero_list, dim3(nblk(gnV.n)), dim3(BS), 0, s, gnV.n, gnV.node.p(), lm.p(),
                               A.cGN.p(), A.sGN.p(), out);
        tmap_gather(A.gnT, lm.p(), out, s);
    };
    double itA1 = 0, itA2 = 0;
    // the adjoint solves stop at the forward solves' tolerances and limits
    const double rtol1 = db->velocitySolve.tolerance, rtol2 = db->PoissonSolve.tolerance;
    const int maxit1 = db->velocitySolve.maxIterations, maxit2 = db->PoissonSolve.maxIterations;
    auto check = [&](const SolveStat &st, const char *what, int n) {
        if (!st.converged)
            throw Error(std::string("adjoint ") + what + " solve did not converge at time step " + std::to_string(n) +
                            " (" + std::to_string(st.iterations) + " iterations)",
                        ERR_NOT_CONVERGED);
    };
    for (int seg = (nsteps - 1) / every; seg >= 0; seg--) {
        const int c0 = seg * every, c1 = std::min(c0 + every, nsteps);
        // recompute the segment from its checkpoint, recording each step's inputs
        const auto tr = clk::now();
        restore(cp.p() + slot(seg));
        timeStep = c0;
        for (int n = c0 + 1; n <= c1; n++) {
            double *r = rslot(n - c0 - 1);
            copy(r, u.p(), nuv);
            copy(r + nuv, N.p(), nuv);
            copy(r + 2 * nuv, bc[XPLUS].p(), nxp);
            stepTime();
        }
        copy(rslot(c1 - c0), u.p(), nuv);  // u after the last step of the segment
        CUIBM_HIP(hipStreamSynchronize(s));
        t_recompute += secs(tr);
        for (int n = c1; n > c0; n--) {
            const double *uin = rslot(n - c0 - 1), *Nin = uin + nuv, *xpin = uin + 2 * nuv;
            const double *uout = rslot(n - c0);
            // forward quantities of step n: xp' and u1
            copy(xpP.p(), xpin, nxp);
            launch_robin(uin, xpP.p(), gd, dt, s);
            copy(u1.p(), uin, nuv);
            launch_gn_vel_apply(u1.p(), gnV, numU, bUB[0], bVB[0], s);
            launch_set_inside(u1.p(), solid.p(), nsolid, numU, bUB[0], bVB[0], s, solidub);
            // force: d(c Fx) / d(u', uold = u1, p1)
            zero(lu1, nuv);
            zero(lp1, np);
            copy(lxpP.p(), lxp.p(), nxp);
            if (n >= n0 && n <= n1) {
                const long m = 3l * (box.ncx + 4) * (box.ncy + 4);
                hipLaunchKernelGGL(k_force_grad, dim3(nblk(m)), dim3(BS), 0, s, gd, nu, dt, box, c, uout,
                                   ghostTagsUV.p(), lu.p(), lu1.p(), lp1.p());
                CUIBM_LAUNCH_CHECK();
            }
            // u' = S(G(u2)): lambda_u2
            SG_T(lu.p(), lu2.p(), true);
            // u2 = uh2 - dt Grad p1
            hipLaunchKernelGGL(k_grad_T, dim3(blocks_for(np)), dim3(BS), 0, s, gd, dt, lu2.p(), lp1.p());
            // p1 = GNp(blend(p0, HNp p0))
            copy(lq.p(), lp1.p(), np);
            if (gnP.n)
                hipLaunchKernelGGL(k_zero_list, dim3(nblk(gnP.n)), dim3(BS), 0, s, gnP.n, gnP.node.p(), lp1.p(),
                                   nullptr, nullptr, lq.p());
            tmap_gather(A.gnpT, lp1.p(), lq.p(), s);
            copy(lp0.p(), lq.p(), np);
            zero(lpst, np);
            if (hnP.n)
                hipLaunchKernelGGL(k_blend_T, dim3(nblk(hnP.n)), dim3(BS), 0, s, hnP.n, hnP.node.p(), hnP.extra.p(),
                                   lq.p(), lp0.p(), lpst.p());
            tmap_gather(A.hnpT, lpst.p(), lp0.p(), s);
            // p0 = LHS2^-1 rhs2
            // warm start: the previous (later) step's adjoint pressure, as the forward
            // solve starts from the previous pressure
            {
                const SolveStat st = A.k2->solve(*A.op2, lr2.p(), lp0.p(), *A.M2, rtol2, maxit2, s);
                check(st, "Poisson", n);
                itA2 += st.iterations;
            }
            // rhs2 = D uh2 + d(xp'): lambda_uh2 += D^T lr2; lambda_xp'[J] -= lr2[east cell] / dx
            hipLaunchKernelGGL(k_div_T, dim3(blocks_for(nuv)), dim3(BS), 0, s, gd, lr2.p(), lu2.p());
            hipLaunchKernelGGL(k_div_xp_T, dim3(nblk(ny)), dim3(BS), 0, s, gd, lr2.p(), lxpP.p());
            // uh2 = G(uh1): lambda_uh1
            SG_T(lu2.p(), luh.p(), false);
            // uh1 = (1 - A) uh0 + A ustar
            zero(lst, nuv);
            copy(lm.p(), luh.p(), nuv);
            if (hnV.n)
                hipLaunchKernelGGL(k_blend_T, dim3(nblk(hnV.n)), dim3(BS), 0, s, hnV.n, hnV.node.p(), hnV.extra.p(),
                                   lm.p(), luh.p(), lst.p());
            // uh0 = LHS1^-1 rhs1
            // warm start likewise from the later step's adjoint velocity
            {
                const SolveStat st = A.k1->solve(*A.op1, lr1.p(), luh.p(), *A.M1, rtol1, maxit1, s);
                check(st, "velocity", n);
                itA1 += st.iterations;
            }
            // rhs1 = f(u1, Nold, xp'), N' = N(u1, xp'): transposed Jacobians by colour probes
            BcDev wb = bcs();
            wb.xp = xpP.p();
            const double sc = 1.0;
            for (int col = 0; col < 18; col++) {
                hipLaunchKernelGGL(k_colour_pm, dim3(blocks_for(nuv)), dim3(BS), 0, s, gd, u1.p(), up.p(), um.p(),
                                   col, sc);
                launch_rhs1(up.p(), Nin, Np.p(), Rp.p(), nullptr, nullptr, gd, wb, nu, dt, fadlunRhs(), s);
                launch_rhs1(um.p(), Nin, Nm.p(), Rm.p(), nullptr, nullptr, gd, wb, nu, dt, fadlunRhs(), s);
                hipLaunchKernelGGL(k_colour_gather, dim3(blocks_for(nuv)), dim3(BS), 0, s, gd, col, sc, Rp.p(), Rm.p(),
                                   Np.p(), Nm.p(), lr1.p(), lN.p(), lu1.p());
            }
            for (int col = 0; col < 6; col++) {
                hipLaunchKernelGGL(k_xp_pm, dim3(nblk(nxp)), dim3(BS), 0, s, ny, xpP.p(), xpp.p(), xpm.p(), col, sc);
                BcDev wp = wb, wm = wb;
                wp.xp = xpp.p();
                wm.xp = xpm.p();
                launch_rhs1(u1.p(), Nin, Np.p(), Rp.p(), nullptr, nullptr, gd, wp, nu, dt, fadlunRhs(), s);
                launch_rhs1(u1.p(), Nin, Nm.p(), Rm.p(), nullptr, nullptr, gd, wm, nu, dt, fadlunRhs(), s);
                hipLaunchKernelGGL(k_xp_gather, dim3(nblk(nxp)), dim3(BS), 0, s, gd, col, sc, Rp.p(), Rm.p(), Np.p(),
                                   Nm.p(), lr1.p(), lN.p(), lxpP.p());
            }
            // lambda of Nold (the previous step's N): 0.5 dt lr1
            zero(lNo, nuv);
            hipLaunchKernelGGL(k_axpy, dim3(blocks_for(nuv)), dim3(BS), 0, s, nuv, 0.5 * dt, lr1.p(), lNo.p());
            // ustar = H u1 + body terms
            tmap_gather(A.hnT, lst.p(), lu1.p(), s);
            if (hnV.n)
                hipLaunchKernelGGL(k_sens, dim3(nblk(hnV.n)), dim3(BS), 0, s, hnV.n, hnV.node.p(), A.cHN.p(), lst.p(),
                                   A.sHN.p());
            // u1 = S(G(u_in)): lambda_u_in
            SG_T(lu1.p(), lu.p(), true);
            // xp' = Robin(xp_in, u_in)
            hipLaunchKernelGGL(k_robin_T, dim3(nblk(nxp)), dim3(BS), 0, s, gd, dt, lxpP.p(), lxp.p(), lu.p());
            copy(lN.p(), lNo.p(), nuv);
        }
    }
    // u^0 = S(u_init) + solid body values
    if (nsolid)
        hipLaunchKernelGGL(k_zero_list, dim3(nblk(nsolid)), dim3(BS), 0, s, nsolid, solid.p(), lu.p(), nullptr,
                           A.sSolid.p(), lm.p());
    CUIBM_HIP(hipStreamSynchronize(s));

    // ---- per-node sensitivities to the per-point velocities and the rotation
    const int npts = (int)bX.size();
    std::vector<double> gu(npts, 0.0), gv(npts, 0.0);
    auto spread = [&](const DArray<double> &sens, int m, const std::vector<int> &node, const std::vector<int> &seg,
                      const std::vector<double> &w) {
        if (!m) return;
        const std::vector<double> h = sens.to_host(s);
        for (int t = 0; t < m; t++) {
            const int k = seg[t], l = (k + 1) % npts;
            std::vector<double> &g = node[t] < numU ? gu : gv;
            g[k] += (1 - w[t]) * h[t];
            g[l] += w[t] * h[t];
        }
    };
    spread(A.sGN, gnV.n, gnNode, gnSeg, gnW);
    spread(A.sHN, hnV.n, hnNode, hnSeg, hnW);
    double gom = 0;
    if (nsolid) {
        const std::vector<double> h = A.sSolid.to_host(s), hx = gx.to_host(s), hy = gy.to_host(s);
        double cx = 0, cy = 0;
        for (int k = 0; k < npts; k++) {
            cx += bX[k];
            cy += bY[k];
        }
        cx /= npts;
        cy /= npts;
        for (int t = 0; t < nsolid; t++) {
            const int i = solidNode[t];
            if (i < numU) gom += h[t] * (-(hy[i / (nx - 1)] - cy));
            else gom += h[t] * (hx[(i - numU) % nx] - cx);
        }
    }
    for (int k = 0; k < npts; k++) {
        if (gub) gub[k] = gu[k];
        if (gvb) gvb[k] = gv[k];
    }
    if (gomega) *gomega = gom;
    if (stats) {
        stats[1] = secs(t1);
        stats[4] = itA2;
        stats[5] = itA1;
        stats[6] = t_setup;
        stats[7] = t_recompute;
    }
}

}  // namespace cuibm
