// Luo embedded method: matrix assembly with the interpolation folded in, and the
// per-step right-hand-side update, FP64.
//
// Restates (paths under src/solvers/NavierStokes/luo_iter/kernels/):
//   LHS1.cu:12-408   LHS1_mid_iter_X / LHS1_mid_iter_Y
//   LHS2.cu:12-224   LHS2_mid_iter
//   intermediateVelocity.cu:48-73, pressure.cu:36-46   update_rhs1_x/y, update_rhs2
// The reference writes COO entries at fixed slots and sorts them; here each row is
// built in registers, sorted by column, and written to a six-entry row buffer. The
// slot pattern's anomalies (a ghost row that writes fewer than five entries, a hybrid
// row whose tail slot is not written or written twice, an interpolation corner
// outside the 3x3 stencil) are counted, because the reference's result then depends
// on what the unwritten slots held from earlier steps.
#include <algorithm>

#include "embedded.hpp"

namespace cuibm {

namespace {

struct Row {
    int c[EMB_ROW];
    double v[EMB_ROW];
    int n;
    __device__ void add(int col, double val) {
        c[n] = col;
        v[n] = val;
        n++;
    }
};

__device__ void store_row(const EmbAsm &a, int i, Row &r) {
    for (int p = 1; p < r.n; p++)  // insertion sort by column (columns are distinct)
        for (int q = p; q > 0 && r.c[q] < r.c[q - 1]; q--) {
            const int tc = r.c[q];
            r.c[q] = r.c[q - 1];
            r.c[q - 1] = tc;
            const double tv = r.v[q];
            r.v[q] = r.v[q - 1];
            r.v[q - 1] = tv;
        }
    for (int k = 0; k < r.n; k++) {
        a.bc[EMB_ROW * i + k] = r.c[k];
        a.bv[EMB_ROW * i + k] = r.v[k];
    }
    a.cnt[i] = r.n;
}

__device__ void copy_plain(const EmbAsm &a, int i, Row &r) {
    for (int k = a.prp[i]; k < a.prp[i + 1]; k++) r.add(a.pci[k], a.pv[k]);
}

// hybrid row (LHS1.cu:33-100 / LHS2.cu:29-103): cns = N, E, S, W blended, cin = alpha *
// weights; the 5-point entries, then at most one tail entry (the last diagonal corner)
__device__ void hybrid_row(const EmbAsm &a, int h, int st, const double cns[4], const double cin[4], const int idx[4],
                           Row &r) {
    const int sidx[9] = {h + st - 1, h + st, h + st + 1, h - 1, h, h + 1, h - st - 1, h - st, h - st + 1};
    double sten[9] = {0, cns[0], 0, cns[3], 1, cns[1], 0, cns[2], 0};
    const bool used[9] = {false, true, false, true, true, true, false, true, false};
    for (int n = 0; n < 4; n++)
        for (int m = 0; m < 9; m++)
            if (sidx[m] == idx[n] && m != 4) sten[m] -= cin[n];
    for (int m = 0; m < 9; m++)
        if (used[m]) r.add(sidx[m], sten[m]);
    int writes = 0, dropped = 0, tcol = -1;
    double tval = 0;
    for (int n = 0; n < 4; n++) {
        bool inside = false;
        for (int m = 0; m < 9; m++) {
            if (sidx[m] != idx[n]) continue;
            inside = true;
            if (!used[m]) {
                writes++;
                tcol = idx[n];
                tval = -cin[n];
            }
        }
        if (!inside) dropped++;
    }
    if (writes > 0) r.add(tcol, tval);
    if (writes == 0) atomicAdd(&a.stats[1], 1);
    if (writes > 1) atomicAdd(&a.stats[2], 1);
    if (dropped) atomicAdd(&a.stats[3], dropped);
}

// ghost row (LHS1.cu:101-160 / LHS2.cu:104-170); sign +1 velocity, -1 pressure
__device__ void ghost_row(const EmbAsm &a, int g, int st, const double cin[4], const int idx[4], double sign, Row &r) {
    bool in[4];
    const int nsi[5] = {g + st, g + 1, g - st, g - 1, g};
    bool overlap[5] = {false, false, false, false, true};
    int counter = 0;
    for (int l = 0; l < 4; l++) {
        in[l] = a.ghost[idx[l]] > 0;
        if (in[l]) counter++;
        for (int n = 0; n < 5; n++)
            if (idx[l] == nsi[n]) overlap[n] = true;
    }
    r.add(g, 1.0);
    int written = 1;
    for (int i = 0; i < 4; i++)
        if (!in[i] && idx[i] != g) {
            r.add(idx[i], sign * cin[i]);
            written++;
        }
    int counter2 = 0;
    for (int i = 0; i < 5; i++) {
        if (counter2 >= counter) break;
        if (!overlap[i]) {
            r.add(nsi[i], 0.0);
            written++;
            counter2++;
        }
    }
    if (written < 5) atomicAdd(&a.stats[0], 5 - written);
}

__global__ void k_emb_lhs1(GridDev g, double nu, double dt, EmbAsm a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int nx = g.nx, ny = g.ny, numU = (nx - 1) * ny;
    const bool cu = i < numU;
    const int st = cu ? nx - 1 : nx, off = cu ? 0 : numU, ni = cu ? nx - 1 : nx, nj = cu ? ny : ny - 1;
    const int I = (i - off) % st, J = (i - off) / st;
    const double *dx = g.dx, *dy = g.dy;
    Row r;
    r.n = 0;
    double ns = 0;
    if (I == 0 || I == ni - 1 || J == 0 || J == nj - 1) {
        copy_plain(a, i, r);
    } else if (a.map[i] >= 0) {
        const int t = a.map[i];
        double cns[5];
        cns[0] = -dt * nu / (dy[J + 1] * (dy[J] + dy[J + 1]));
        cns[1] = -dt * nu / (dx[I] * (dx[I] + dx[I + 1]));
        cns[2] = -dt * nu / (dy[J] * (dy[J] + dy[J + 1]));
        cns[3] = -dt * nu / (dx[I] * (dx[I] + dx[I - 1]));
        cns[4] = 1 - cns[0] - cns[1] - cns[2] - cns[3];
        const double al = a.hyb.extra[t];
        double cin[4];
        int idx[4];
        for (int l = 0; l < 4; l++) {
            cns[l] = cns[l] * (1 - al) / cns[4];
            cin[l] = a.hyb.coef[4 * t + l] * al;
            idx[l] = a.hyb.idx0[4 * t + l];
        }
        hybrid_row(a, i, st, cns, cin, idx, r);
        ns = (1 - al) / cns[4];
    } else if (a.map[i] <= -2) {
        const int t = -a.map[i] - 2;
        double cin[4];
        int idx[4];
        for (int l = 0; l < 4; l++) {
            cin[l] = a.gho.coef[4 * t + l];
            idx[l] = a.gho.idx0[4 * t + l];
        }
        ghost_row(a, i, st, cin, idx, 1.0, r);
        ns = 0;
    } else {
        double temp, e, w, n, s;
        if (cu) {
            temp = 1 + 0.5 * dt * nu * (1 / (dx[I + 1] * (dx[I + 1] + dx[I]) * 0.5)) +
                   0.5 * dt * nu * (1 / (dx[I] * (dx[I + 1] + dx[I]) * 0.5)) +
                   0.5 * dt * nu * (1 / (dy[J] * (dy[J + 1] + dy[J]) * 0.5)) +
                   0.5 * dt * nu * (1 / (dy[J] * (dy[J - 1] + dy[J]) * 0.5));
            e = -0.5 * dt * nu * (1 / (dx[I + 1] * (dx[I + 1] + dx[I]) * 0.5)) / temp;
            w = -0.5 * dt * nu * (1 / (dx[I] * (dx[I + 1] + dx[I]) * 0.5)) / temp;
            n = -0.5 * dt * nu * (1 / (dy[J] * (dy[J + 1] + dy[J]) * 0.5)) / temp;
            s = -0.5 * dt * nu * (1 / (dy[J] * (dy[J - 1] + dy[J]) * 0.5)) / temp;
        } else {
            temp = 1 + 0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I + 1]) * 0.5)) +
                   0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I - 1]) * 0.5)) +
                   0.5 * dt * nu * (1 / (dy[J + 1] * (dy[J] + dy[J + 1]) * 0.5)) +
                   0.5 * dt * nu * (1 / (dy[J] * (dy[J] + dy[J + 1]) * 0.5));
            e = -0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I + 1]) * 0.5)) / temp;
            w = -0.5 * dt * nu * (1 / (dx[I] * (dx[I] + dx[I - 1]) * 0.5)) / temp;
            n = -0.5 * dt * nu * (1 / (dy[J + 1] * (dy[J] + dy[J + 1]) * 0.5)) / temp;
            s = -0.5 * dt * nu * (1 / (dy[J] * (dy[J] + dy[J + 1]) * 0.5)) / temp;
        }
        r.add(i + 1, e);
        r.add(i - 1, w);
        r.add(i + st, n);
        r.add(i - st, s);
        r.add(i, 1.0);
        ns = 1 / temp;
    }
    store_row(a, i, r);
    a.ns[i] = ns;
}

__global__ void k_emb_lhs2(GridDev g, double dt, EmbAsm a) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= a.n) return;
    const int nx = g.nx, ny = g.ny;
    const int I = ip % nx, J = ip / nx;
    const double *dx = g.dx, *dy = g.dy;
    Row r;
    r.n = 0;
    double ns = 0;
    if (I == 0 || I == nx - 1 || J == 0 || J == ny - 1) {
        copy_plain(a, ip, r);
    } else if (a.map[ip] >= 0) {
        const int t = a.map[ip];
        double cns[5];
        cns[0] = -dt / (dy[J] * (dy[J] + dy[J + 1]) * 0.5);
        cns[1] = -dt / (dx[I] * (dx[I] + dx[I + 1]) * 0.5);
        cns[2] = -dt / (dy[J] * (dy[J] + dy[J - 1]) * 0.5);
        cns[3] = -dt / (dx[I] * (dx[I] + dx[I - 1]) * 0.5);
        cns[4] = -cns[0] - cns[1] - cns[2] - cns[3];
        const double al = a.hyb.extra[t];
        double cin[4];
        int idx[4];
        for (int l = 0; l < 4; l++) {
            cns[l] = (1 - al) * cns[l] / cns[4];
            cin[l] = al * a.hyb.coef[4 * t + l];
            idx[l] = a.hyb.idx0[4 * t + l];
        }
        hybrid_row(a, ip, nx, cns, cin, idx, r);
        ns = (1 - al) / cns[4];
    } else if (a.map[ip] <= -2) {
        const int t = -a.map[ip] - 2;
        double cin[4];
        int idx[4];
        for (int l = 0; l < 4; l++) {
            cin[l] = a.gho.coef[4 * t + l];
            idx[l] = a.gho.idx0[4 * t + l];
        }
        ghost_row(a, ip, nx, cin, idx, -1.0, r);
        ns = 0;
    } else {
        double temp = 0;
        const double e = -dt / (dx[I] * (dx[I] + dx[I + 1]) * 0.5);
        temp += dt / (dx[I] * (dx[I] + dx[I + 1]) * 0.5);
        const double w = -dt / (dx[I] * (dx[I] + dx[I - 1]) * 0.5);
        temp += dt / (dx[I] * (dx[I] + dx[I - 1]) * 0.5);
        const double n = -dt / (dy[J] * (dy[J] + dy[J + 1]) * 0.5);
        temp += dt / (dy[J] * (dy[J] + dy[J + 1]) * 0.5);
        const double s = -dt / (dy[J] * (dy[J] + dy[J - 1]) * 0.5);
        temp += dt / (dy[J] * (dy[J] + dy[J - 1]) * 0.5);
        r.add(ip + 1, e);
        r.add(ip - 1, w);
        r.add(ip + nx, n);
        r.add(ip - nx, s);
        r.add(ip, temp);
        ns = 1;
    }
    store_row(a, ip, r);
    a.ns[ip] = ns;
}

__global__ void k_emb_compact(int n, const int *__restrict__ cnt, const int *__restrict__ bc,
                              const double *__restrict__ bv, const int *__restrict__ rp, int *__restrict__ ci,
                              double *__restrict__ v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = rp[i];
    for (int k = 0; k < cnt[i]; k++) {
        ci[b + k] = bc[EMB_ROW * i + k];
        v[b + k] = bv[EMB_ROW * i + k];
    }
}

// interp_rhs of one velocity row (LHS1.cu:78-99, 101-160)
__device__ double interp_rhs1(const EmbAsm &a, int i, const double *u, double ub) {
    const int m = a.map[i];
    if (m >= 0) {  // hybrid: the node's own corner (body intercept, value uB)
        const double al = a.hyb.extra[m];
        double s = 0;
        for (int l = 0; l < 4; l++)
            if (a.hyb.idx0[4 * m + l] == i) s += a.hyb.coef[4 * m + l] * al * ub;
        return s;
    }
    if (m <= -2) {  // ghost: 2 uB minus the explicit (ghost) corners at their current values
        const int t = -m - 2;
        double temp = 0;
        for (int l = 0; l < 4; l++) {
            const int j = a.gho.idx0[4 * t + l];
            if (a.ghost[j] > 0 || j == i) {
                const double q = a.gho.idx[4 * t + l] < 0 ? ub : u[j];
                temp -= a.gho.coef[4 * t + l] * q;
            }
        }
        return 2 * ub + temp;
    }
    return 0;
}

__global__ void k_emb_rhs1(double *__restrict__ rhs1, const double *__restrict__ u, EmbAsm a, int nx, int ny,
                           int vthreads, double uB, double vB) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int numU = (nx - 1) * ny;
    if (i >= a.n) return;
    double ub;
    if (i < numU) {
        const int I = i % (nx - 1), J = i / (nx - 1);
        if (I == 0 || I == nx - 2 || J == 0 || J == ny - 1) return;
        ub = uB;
    } else {
        const int ip = i - numU, I = ip % nx, J = ip / nx;
        if (ip >= vthreads) return;
        if (I == 0 || I == nx - 1 || J == 0 || J == ny - 2) return;
        ub = vB;
    }
    rhs1[i] = rhs1[i] * a.ns[i] + interp_rhs1(a, i, u, ub);
}

// interp_rhs of one pressure row (LHS2.cu:78-103, 104-170); replaced corners carry the
// Neumann datum -(n . duB/dt)
__global__ void k_emb_rhs2(double *__restrict__ rhs2, EmbAsm a, int nx, int ny, double dudt, double dvdt) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= a.n || ip >= (nx - 1) * ny) return;
    const int I = ip % nx, J = ip / nx;
    if (I == 0 || I == nx - 1 || J == 0 || J == ny - 1) return;
    const int m = a.map[ip];
    double interp = 0;
    if (m >= 0) {
        const double al = a.hyb.extra[m];
        for (int l = 0; l < 4; l++)
            if (a.hyb.idx0[4 * m + l] == ip) {
                const double q = -(a.hyb.nrm[8 * m + 2 * l] * dudt + a.hyb.nrm[8 * m + 2 * l + 1] * dvdt);
                interp += al * a.hyb.coef[4 * m + l] * q;
            }
    } else if (m <= -2) {
        const int t = -m - 2;
        double temp = 0;
        for (int l = 0; l < 4; l++) {
            const int j = a.gho.idx0[4 * t + l];
            if (a.ghost[j] > 0 || j == ip) {
                const double q = -(a.gho.nrm[8 * t + 2 * l] * dudt + a.gho.nrm[8 * t + 2 * l + 1] * dvdt);
                temp += a.gho.coef[4 * t + l] * q;
            }
        }
        const double matDClose = a.gho.extra[3 * t + 1] * dudt + a.gho.extra[3 * t + 2] * dvdt;
        interp = a.gho.extra[3 * t] * matDClose + temp;
    }
    rhs2[ip] = rhs2[ip] * a.ns[ip] + interp;
}

}  // namespace

void launch_emb_lhs1(const GridDev &g, double nu, double dt, const EmbAsm &a, hipStream_t s) {
    hipLaunchKernelGGL(k_emb_lhs1, dim3(nblk(a.n)), dim3(BS), 0, s, g, nu, dt, a);
    CUIBM_LAUNCH_CHECK();
}

void launch_emb_lhs2(const GridDev &g, double dt, const EmbAsm &a, hipStream_t s) {
    hipLaunchKernelGGL(k_emb_lhs2, dim3(nblk(a.n)), dim3(BS), 0, s, g, dt, a);
    CUIBM_LAUNCH_CHECK();
}

void launch_emb_compact(const EmbAsm &a, const int *rp, int *ci, double *v, hipStream_t s) {
    hipLaunchKernelGGL(k_emb_compact, dim3(nblk(a.n)), dim3(BS), 0, s, a.n, a.cnt, a.bc, a.bv, rp, ci, v);
    CUIBM_LAUNCH_CHECK();
}

void launch_emb_rhs1(double *rhs1, const double *u, const EmbAsm &a, const GridDev &g, double uB, double vB,
                     hipStream_t s) {
    // update_rhs1_y is launched with the u grid (luo_iter/intermediateVelocity.cu:197)
    const long numU = (long)(g.nx - 1) * g.ny;
    const int vthreads = (int)(BS * ((long)((numU - 0.5) / BS) + 1));
    hipLaunchKernelGGL(k_emb_rhs1, dim3(nblk(a.n)), dim3(BS), 0, s, rhs1, u, a, g.nx, g.ny, vthreads, uB, vB);
    CUIBM_LAUNCH_CHECK();
}

void launch_emb_rhs2(double *rhs2, const EmbAsm &a, const GridDev &g, double dudt, double dvdt, hipStream_t s) {
    hipLaunchKernelGGL(k_emb_rhs2, dim3(nblk(a.n)), dim3(BS), 0, s, rhs2, a, g.nx, g.ny, dudt, dvdt);
    CUIBM_LAUNCH_CHECK();
}

// ---- device row scan and row maps ----------------------------------------------------

constexpr int SCAN_BS = 256, SCAN_PER = 8, SCAN_TILE = SCAN_BS * SCAN_PER;

int scan_parts(int n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1; }

// inclusive scan of one thread's values, then across the block (Hillis-Steele on LDS)
__device__ int block_excl_scan(int v, int *sh, int &total) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < SCAN_BS; o <<= 1) {
        const int t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    total = sh[SCAN_BS - 1];
    return sh[threadIdx.x] - v;
}

// pass 1: the sum of each tile of SCAN_TILE counts
__global__ void __launch_bounds__(SCAN_BS) k_scan_tiles(const int *__restrict__ cnt, int n, int *__restrict__ part) {
    __shared__ int sh[SCAN_BS];
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER;
    int v = 0;
    for (int k = 0; k < SCAN_PER; k++)
        if (base + k < n) v += cnt[base + k];
    int total;
    block_excl_scan(v, sh, total);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// pass 2: exclusive scan of the tile sums in place (one block, looping over the tiles)
__global__ void __launch_bounds__(SCAN_BS) k_scan_parts(int *__restrict__ part, int np) {
    __shared__ int sh[SCAN_BS];
    int carry = 0;
    for (int b = 0; b < np; b += SCAN_BS) {
        const int v = b + (int)threadIdx.x < np ? part[b + threadIdx.x] : 0;
        int total;
        const int e = block_excl_scan(v, sh, total);
        if (b + (int)threadIdx.x < np) part[b + threadIdx.x] = carry + e;
        carry += total;
        __syncthreads();
    }
}

// pass 3: each tile's row pointers from its offset; the last tile writes rp[n]
__global__ void __launch_bounds__(SCAN_BS) k_scan_rows(const int *__restrict__ cnt, int n,
                                                       const int *__restrict__ part, int *__restrict__ rp) {
    __shared__ int sh[SCAN_BS];
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER;
    int v = 0;
    for (int k = 0; k < SCAN_PER; k++)
        if (base + k < n) v += cnt[base + k];
    int total;
    int off = part[blockIdx.x] + block_excl_scan(v, sh, total);
    if (blockIdx.x == 0 && threadIdx.x == 0) rp[0] = 0;
    for (int k = 0; k < SCAN_PER; k++)
        if (base + k < n) {
            off += cnt[base + k];
            rp[base + k + 1] = off;
        }
}

void launch_row_scan(const int *cnt, int *rp, int n, int *part, hipStream_t s) {
    if (n <= 0) return;
    const int nt = scan_parts(n) - 1;
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(SCAN_BS), 0, s, cnt, n, part);
    CUIBM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(SCAN_BS), 0, s, part, nt);
    CUIBM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_rows, dim3(nt), dim3(SCAN_BS), 0, s, cnt, n, part, rp);
    CUIBM_LAUNCH_CHECK();
}

__global__ void k_map_fill(int *__restrict__ map, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) map[i] = -1;
}
__global__ void k_map_scatter(int *__restrict__ map, const int *__restrict__ node, int m, int ghost) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < m; t += gridDim.x * blockDim.x)
        map[node[t]] = ghost ? -(t + 2) : t;
}

void launch_emb_map(int *map, int n, const int *ghost, int ng, const int *hybrid, int nh, hipStream_t s) {
    auto blocks = [](int m) { return dim3((unsigned)std::max(1, std::min((m + BS - 1) / BS, 65536))); };
    hipLaunchKernelGGL(k_map_fill, blocks(n), dim3(BS), 0, s, map, n);
    CUIBM_LAUNCH_CHECK();
    if (ng > 0) {
        hipLaunchKernelGGL(k_map_scatter, blocks(ng), dim3(BS), 0, s, map, ghost, ng, 1);
        CUIBM_LAUNCH_CHECK();
    }
    if (nh > 0) {  // after the ghost scatter: a node in both lists ends as a hybrid row
        hipLaunchKernelGGL(k_map_scatter, blocks(nh), dim3(BS), 0, s, map, hybrid, nh, 0);
        CUIBM_LAUNCH_CHECK();
    }
}

}  // namespace cuibm
