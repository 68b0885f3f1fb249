// Step stages: Euler move with absorption and deletion (A9), force and
// circulation sums (A10), compaction (A11) and collapse (A12/A13).
//
// These stages decide integers -- which vortex is absorbed, deleted, kept,
// merged and where each ends up -- so every multiply and add here is
// separately rounded and every decision uses the reference's expression.
// Sums use a fixed two-level order (per work-group, then over work-groups),
// so they are run-to-run deterministic.
#pragma clang fp contract(off)


#include "vx_internal.h"
#include "vx_prim.h"

namespace vx {

__device__ __forceinline__ double dist2(double a0, double a1, double b0, double b1) {
    return (a0 - b0) * (a0 - b0) + (a1 - b1) * (a1 - b1);
}

struct StepConsts {
    double dt, rc_x, rc_y, x_min, x_max, y_min, y_max, px_min, px_max, py_min, py_max;
};

// hitting(): src/kernel.cu:1357-1431.  Segment old -> new against every
// panel (after the profile-box reject); the crossing nearest the old position
// wins, the lowest panel index among equal distances.  A crossing set with no
// panel nearer than 5 (the reference would read panels[-1]) counts as no
// crossing.  The moving vortices whose segment box meets the profile box are
// selected in index order; one wave per candidate tests the panels lane,
// lane + 64, ... and reduces (distance, index) to the winner of the
// sequential scan.
__device__ __forceinline__ bool hit_box_reject(const StepConsts &k, double x1, double y1, double x2, double y2) {
    return ((x1 < k.px_min) && (x2 < k.px_min)) || ((x1 > k.px_max) && (x2 > k.px_max)) ||
           ((y1 < k.py_min) && (y2 < k.py_min)) || ((y1 > k.py_max) && (y2 > k.py_max));
}

struct HitCandidate {
    const double *x, *y, *g, *u, *v;
    StepConsts k;
    __device__ bool operator()(int32_t i) const {
        const double xi = x[i], yi = y[i];
        if (g[i] == 0) return false;
        const double xn = xi + u[i] * k.dt;
        const double yn = yi + v[i] * k.dt;
        return !hit_box_reject(k, xn, yn, xi, yi);
    }
};

// Squared distance from (x2, y2) to the crossing of the segment with panel
// i, or -1 if they do not cross.
__device__ __forceinline__ double panel_cross(const PanelsSoA &pan, int64_t i, double A, double B, double D,
                                              double x1, double y1, double x2, double y2) {
    const double porog = 1e-12;
    const double lx = pan.lx[i], ly = pan.ly[i], rx = pan.rx[i], ry = pan.ry[i];
    double r0 = A * lx + B * ly + D;
    double r1 = A * rx + B * ry + D;
    if (fabs(r0) < porog) r0 = 0.0;
    if (fabs(r1) < porog) r1 = 0.0;
    if (!(r0 * r1 <= 0)) return -1.0;
    const double A1 = ry - ly, B1 = lx - rx, D1 = ly * rx - lx * ry;
    double r2 = A1 * x1 + B1 * y1 + D1;
    double r3 = A1 * x2 + B1 * y2 + D1;
    if (fabs(r2) < porog) r2 = 0.0;
    if (fabs(r3) < porog) r3 = 0.0;
    if (!(r2 * r3 <= 0)) return -1.0;
    const double den = A * B1 - B * A1;
    const double X = (B * D1 - D * B1) / den;
    const double Y = (A1 * D - D1 * A) / den;
    return (x2 - X) * (x2 - X) + (y2 - Y) * (y2 - Y);
}

// Grid-stride over the candidates, one wave each; the count stays on the device.
__global__ __launch_bounds__(256) void k_hit_panels(const double *__restrict__ x, const double *__restrict__ y,
                                                    const double *__restrict__ u, const double *__restrict__ v,
                                                    PanelsSoA pan, StepConsts k, const int32_t *__restrict__ cand,
                                                    const int32_t *__restrict__ ncand, int32_t *__restrict__ hitp) {
    const int lane = threadIdx.x & 63;
    const int64_t nw = (int64_t)gridDim.x * 4;
    const int64_t total = *ncand;
    for (int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); t < total; t += nw) {
        const int64_t i = cand[t];
        const double x2 = x[i], y2 = y[i];
        const double x1 = x2 + u[i] * k.dt, y1 = y2 + v[i] * k.dt;
        const double A = y2 - y1, B = x1 - x2, D = y1 * x2 - x1 * y2;
        double best = 25.0;
        int bi = -1;
        for (int64_t p = lane; p < pan.P; p += 64) {
            const double dd2 = panel_cross(pan, p, A, B, D, x1, y1, x2, y2);
            if (dd2 >= 0.0 && dd2 < best) {
                best = dd2;
                bi = (int)p;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const double ob = __shfl_xor(best, o, 64);
            const int oi = __shfl_xor(bi, o, 64);
            if (oi >= 0 && (bi < 0 || ob < best || (ob == best && oi < bi))) {
                best = ob;
                bi = oi;
            }
        }
        if (lane == 0) hitp[i] = bi;
    }
}

// ---- A9: src/kernel.cu:1090-1143 -------------------------------------------
// hitp[i] holds the hit panel of every selected candidate (others are not
// read).  Per work-group partials: d_g, Fx, Fy, M sums, then min x, -max x,
// min y, -max y over the vortices alive after the move (the box the collapse
// grid covers).
__global__ __launch_bounds__(256) void k_step_move(double *__restrict__ x, double *__restrict__ y,
                                                   double *__restrict__ g, const double *__restrict__ u,
                                                   const double *__restrict__ v, Cnt cn, int64_t P,
                                                   PanelsSoA pan, StepConsts k, const int32_t *__restrict__ hitp,
                                                   double *__restrict__ part) {
    __shared__ double red[8][256];
    const int64_t n = cn.get();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double dg = 0.0, fx = 0.0, fy = 0.0, mm = 0.0;
    double bx0 = 1e300, bx1 = 1e300, by0 = 1e300, by1 = 1e300;
    if (i < n) {
        double xi = x[i], yi = y[i], gi = g[i];
        if (n >= P && i >= n - P) {
            fx = gi * (-yi);
            fy = gi * (xi);
            mm = gi * dist2(xi, yi, k.rc_x, k.rc_y);
        }
        const double xn = xi + u[i] * k.dt;
        const double yn = yi + v[i] * k.dt;
        if ((gi != 0) && !hit_box_reject(k, xn, yn, xi, yi)) {
            const int hp = hitp[i];
            if (hp >= 0) {
                const double cx = pan.cx[hp], cy = pan.cy[hp];
                fx -= gi * (-cy);
                fy -= gi * (cx);
                mm -= gi * dist2(cx, cy, k.rc_x, k.rc_y);
                dg = gi;
                gi = 0;
            }
        }
        xi = xn;
        yi = yn;
        if ((gi != 0) && ((xi > k.x_max) || (xi < k.x_min) || (yi > k.y_max) || (yi < k.y_min))) {
            xi = -2.0e+5;
            yi = -2.0e+5;
            gi = 0;
        }
        x[i] = xi;
        y[i] = yi;
        g[i] = gi;
        if (gi != 0) {
            bx0 = xi;
            bx1 = -xi;
            by0 = yi;
            by1 = -yi;
        }
    }
    red[0][threadIdx.x] = dg;
    red[1][threadIdx.x] = fx;
    red[2][threadIdx.x] = fy;
    red[3][threadIdx.x] = mm;
    red[4][threadIdx.x] = bx0;
    red[5][threadIdx.x] = bx1;
    red[6][threadIdx.x] = by0;
    red[7][threadIdx.x] = by1;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
            for (int q = 4; q < 8; ++q) red[q][threadIdx.x] = fmin(red[q][threadIdx.x], red[q][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 8) part[(int64_t)blockIdx.x * 8 + threadIdx.x] = red[threadIdx.x][0];
}

// A10: src/kernel.cu:1144-1157 -- d_g = sum, F = rho/dt sum, M = rho/(2 dt)
// sum; out[4..7] the live box (min x, max x, min y, max y).
__global__ __launch_bounds__(256) void k_sum4(const double *__restrict__ part, int64_t nb, double rho, double dt,
                                              double *__restrict__ out) {
    __shared__ double red[8][256];
    double a[8] = {0.0, 0.0, 0.0, 0.0, 1e300, 1e300, 1e300, 1e300};
    for (int64_t b = threadIdx.x; b < nb; b += 256) {
        for (int q = 0; q < 4; ++q) a[q] += part[b * 8 + q];
        for (int q = 4; q < 8; ++q) a[q] = fmin(a[q], part[b * 8 + q]);
    }
    for (int q = 0; q < 8; ++q) red[q][threadIdx.x] = a[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
            for (int q = 4; q < 8; ++q) red[q][threadIdx.x] = fmin(red[q][threadIdx.x], red[q][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = red[0][0];
        out[1] = red[1][0] * (rho / dt);
        out[2] = red[2][0] * (rho / dt);
        out[3] = red[3][0] * (rho / (2 * dt));
        out[4] = red[4][0];
        out[5] = -red[5][0];
        out[6] = red[6][0];
        out[7] = -red[7][0];
    }
}

// The count lives in dcount[1] or dcount[8]: each compaction reads the
// current slot and writes the new count to the other one, so no kernel reads
// a count that another thread of the same launch is replacing.
int64_t *step_dn(vx_case *c) { return c->dcount.p + c->dn_slot; }

// dcount[1] = n (the device count), dcount[3] = 0 (the collapse error flag,
// sticky over the passes of one step).
__global__ void k_step_init(int64_t *__restrict__ dcount, int64_t n) {
    dcount[1] = n;
    dcount[3] = 0;
}

// In device-count mode the count is already in dcount[1] and the error flag
// stays set over the steps of the sequence (cleared when it starts).
int launch_set_dn(vx_case *c, int64_t n) {
    if (c->dcount.ensure(16, c->stream)) return 1;
    if (!c->devcount) hipLaunchKernelGGL(k_step_init, dim3(1), dim3(1), 0, c->stream, c->dcount.p, n);
    VX_CHECK(hipGetLastError());
    c->dn_slot = 1;
    return 0;
}

int launch_step_move(vx_case *c) {
    const Cnt cn = cnt_n(c);
    const int64_t n = cn.ub;  // bound: grids and list sizes
    const int64_t nb = n > 0 ? (n + 255) / 256 : 1;
    hipStream_t st = c->stream;
    if (c->partial.ensure((size_t)(8 * nb), st) || c->red.ensure(8, st) ||
        c->misc.ensure((size_t)(n > 0 ? n : 1), st) || c->lists.ensure((size_t)(n > 0 ? n : 1), st) ||
        c->dcount.ensure(16, st))
        return 1;
    const vx_params &p = c->prm;
    StepConsts k{p.dt, p.rc_x, p.rc_y, p.x_min, p.x_max, p.y_min, p.y_max, p.px_min, p.px_max, p.py_min, p.py_max};
    int32_t *hitp = c->misc.p, *cand = c->lists.p;
    int32_t *ncand = (int32_t *)(c->dcount.p + 4);
    if (n > 0) {
        const HitCandidate pred{c->x.p, c->y.p, c->g.p, c->u.p, c->v.p, k};
        if (c->ptmp.ensure((size_t)prim::tiles_of(n) + 1, st)) return 1;
        VX_CHECK(prim::select_flags(prim::pred_flag(pred, cn), n, prim::WriteIndex{cand}, c->ptmp.p, ncand, st));
        hipLaunchKernelGGL(k_hit_panels, dim3(1024), dim3(256), 0, st, c->x.p, c->y.p, c->u.p, c->v.p, c->pan, k,
                           cand, ncand, hitp);
    }
    hipLaunchKernelGGL(k_step_move, dim3((unsigned)nb), dim3(256), 0, st, c->x.p, c->y.p, c->g.p, c->u.p, c->v.p, cn,
                       c->P, c->pan, k, hitp, c->partial.p);
    hipLaunchKernelGGL(k_sum4, dim3(1), dim3(256), 0, st, c->partial.p, nb, p.rho, p.dt, c->red.p);
    VX_CHECK(hipGetLastError());
    return 0;
}

// ---- tracer points: step_Kernel for the second component ---------------------
// src/unit_1.cu:1271-1340 move the passive and inertial points with
// step_Kernel and null force pointers: the same move, absorption (g = 0) and
// deletion as for vortices, without impulse terms or sums.
__global__ __launch_bounds__(256) void k_points_move(double *__restrict__ x, double *__restrict__ y,
                                                     double *__restrict__ g, const double *__restrict__ u,
                                                     const double *__restrict__ v, int64_t m, StepConsts k,
                                                     const int32_t *__restrict__ hitp) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double xi = x[i], yi = y[i], gi = g[i];
    const double xn = xi + u[i] * k.dt;
    const double yn = yi + v[i] * k.dt;
    if ((gi != 0) && !hit_box_reject(k, xn, yn, xi, yi) && hitp[i] >= 0) gi = 0;
    xi = xn;
    yi = yn;
    if ((gi != 0) && ((xi > k.x_max) || (xi < k.x_min) || (yi > k.y_max) || (yi < k.y_min))) {
        xi = -2.0e+5;
        yi = -2.0e+5;
        gi = 0;
    }
    x[i] = xi;
    y[i] = yi;
    g[i] = gi;
}

int launch_points_move(vx_case *c, int64_t m) {
    if (m <= 0) return 0;
    hipStream_t st = c->stream;
    if (c->misc.ensure((size_t)m, st) || c->lists.ensure((size_t)m, st) || c->dcount.ensure(16, st)) return 1;
    const vx_params &p = c->prm;
    StepConsts k{p.dt, p.rc_x, p.rc_y, p.x_min, p.x_max, p.y_min, p.y_max, p.px_min, p.px_max, p.py_min, p.py_max};
    int32_t *hitp = c->misc.p, *cand = c->lists.p;
    int32_t *ncand = (int32_t *)(c->dcount.p + 6);
    const HitCandidate pred{c->px.p, c->py.p, c->pg.p, c->pu.p, c->pv.p, k};
    if (c->ptmp.ensure((size_t)prim::tiles_of(m) + 1, st)) return 1;
    VX_CHECK(prim::select_flags(prim::pred_flag(pred), m, prim::WriteIndex{cand}, c->ptmp.p, ncand, st));
    hipLaunchKernelGGL(k_hit_panels, dim3(256), dim3(256), 0, st, c->px.p, c->py.p, c->pu.p, c->pv.p, c->pan, k,
                       cand, ncand, hitp);
    hipLaunchKernelGGL(k_points_move, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, c->px.p, c->py.p, c->pg.p,
                       c->pu.p, c->pv.p, m, k, hitp);
    VX_CHECK(hipGetLastError());
    return 0;
}

struct Grid {
    double x0, y0, inv_s;
    int64_t nx, ny, cells;
    int valid;
};

__device__ __forceinline__ void cell_of(const Grid &G, double x, double y, int64_t &cx, int64_t &cy, double &fx,
                                        double &fy) {
    fx = (x - G.x0) * G.inv_s;
    fy = (y - G.y0) * G.inv_s;
    if (!(fx >= 0.0)) fx = 0.0;  // also catches NaN
    if (!(fy >= 0.0)) fy = 0.0;
    cx = fx >= (double)(G.nx - 1) ? G.nx - 1 : (int64_t)fx;
    cy = fy >= (double)(G.ny - 1) ? G.ny - 1 : (int64_t)fy;
    fx -= (double)cx;
    fy -= (double)cy;
}

// A pool node for slot at the cell of (px, py): taken from the pool's count,
// linked in front of the cell's list.  The search that reads the lists runs
// in a later kernel.
__device__ __forceinline__ void grid_push(const GridPush &p, int32_t slot, double px, double py) {
    const Grid G = *(const Grid *)p.grid;
    if (!G.valid) return;
    const int32_t q = atomicAdd(p.pcount, 1);
    if ((int64_t)q >= p.cap) {
        atomicOr(p.err, 2);
        return;
    }
    int64_t cx, cy;
    double fx, fy;
    cell_of(G, px, py, cx, cy, fx, fy);
    p.pslot[q] = slot;
    const int32_t node = (int32_t)(p.base + q);
    p.next[node] = atomicExch(&p.head[cy * G.nx + cx], node);
}

// ---- A11: src/kernel.cu:1159-1178 ------------------------------------------
// Swap-with-last removal of |g| < 1e-12 in closed form.  With n1 live
// entries, the j-th hole of [0, n1) (ascending) takes the j-th live entry of
// [n1, n) (descending).  Every slot e of [n1, n) ends dead (g = 0) holding
// the position of src(e): e+1 when e+1 < n is dead, else the start of e's
// chain -- hole h_j where j-1 live entries of [n1, n) lie above e, or n1
// itself below the last live entry.  Gather to staging, then scatter.
// Grids cover the upper bound n_ub; the count n = *dn is read on the device
// and n1 goes to the other count slot.
__device__ __forceinline__ int32_t live_at(const double *__restrict__ g, int64_t i, int64_t n) {
    return (i < n && !(fabs(g[i]) < kDelt)) ? 1 : 0;
}

// The scan's input: live flags made on the fly.  With merged given and zero
// (a collapse pass that listed no candidate: it merged and killed nothing,
// so the compaction is the identity) every flag reads 0 and no vortex is read.
struct LiveFlag {
    const double *g;
    const int64_t *dn;
    const int64_t *merged;
    __device__ int32_t operator()(int64_t i) const { return (merged && *merged == 0) ? 0 : live_at(g, i, *dn); }
};

__device__ __forceinline__ bool compact_skip(const int64_t *merged) { return merged && *merged == 0; }

// The writer of the live-flag scan: s[i] = pos, the live entries before i;
// n1 = the count of live entries, written to the other count slot *dn1; the
// holes of [0, n1) by rank, the live entries of [n1, n) by rank from the top.
struct CompactLists {
    int32_t *s, *holes, *donors;
    const int64_t *dn, *merged;
    int64_t *dn1;
    __device__ void operator()(int64_t i, int32_t pos, int f, int32_t n1) const {
        if (compact_skip(merged)) {
            if (i == 0) *dn1 = *dn;
            return;
        }
        s[i] = pos;
        if (i == 0) *dn1 = n1;
        if (i < n1) {
            if (!f) holes[i - pos] = (int32_t)i;  // rank among the holes of [0, n1)
        } else if (f) {
            donors[n1 - pos - 1] = (int32_t)i;    // rank from the top among live entries of [n1, n)
        }
    }
};

__global__ void k_compact_gather(const double *__restrict__ x, const double *__restrict__ y,
                                 const double *__restrict__ g, const int32_t *__restrict__ s,
                                 const int64_t *__restrict__ dn, const int64_t *__restrict__ dn1,
                                 const int32_t *__restrict__ holes, const int32_t *__restrict__ donors,
                                 double *__restrict__ tx, double *__restrict__ ty, double *__restrict__ tg,
                                 const int64_t *__restrict__ merged) {
    const int64_t n = *dn;
    if (n <= 0 || compact_skip(merged)) return;
    const int64_t n1 = *dn1;
    const int64_t K = n1 > 0 ? (int64_t)(n1 - (n1 < n ? s[n1] : n1)) : 0;  // holes in [0, n1)
    const int64_t T = n - n1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < K) {
        const int64_t src = donors[t];
        tx[t] = x[src];
        ty[t] = y[src];
        tg[t] = g[src];
    } else if (t < K + T) {
        const int64_t e = n1 + (t - K);
        const int64_t above = n1 - ((int64_t)s[e] + live_at(g, e, n));  // live entries of [n1, n) above e
        const int64_t start = above < K ? (int64_t)holes[above] : n1;
        const int64_t src = (e + 1 < n && !live_at(g, e + 1, n)) ? e + 1 : start;
        tx[t] = x[src];
        ty[t] = y[src];
        tg[t] = 0.0;
    }
}

__global__ void k_compact_scatter(double *__restrict__ x, double *__restrict__ y, double *__restrict__ g,
                                  const int32_t *__restrict__ s, const int64_t *__restrict__ dn,
                                  const int64_t *__restrict__ dn1, const int32_t *__restrict__ holes,
                                  const double *__restrict__ tx, const double *__restrict__ ty,
                                  const double *__restrict__ tg, const int64_t *__restrict__ merged,
                                  GridPush gp) {
    const int64_t n = *dn;
    if (n <= 0 || compact_skip(merged)) return;
    const int64_t n1 = *dn1;
    const int64_t K = n1 > 0 ? (int64_t)(n1 - (n1 < n ? s[n1] : n1)) : 0;
    const int64_t T = n - n1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t dst;
    if (t < K)
        dst = holes[t];
    else if (t < K + T)
        dst = n1 + (t - K);
    else
        return;
    x[dst] = tx[t];
    y[dst] = ty[t];
    g[dst] = tg[t];
    // a filled hole holds a live vortex that its cell's list does not reach
    if (gp.grid && t < K) grid_push(gp, (int32_t)dst, tx[t], ty[t]);
}

int launch_compact(vx_case *c, int64_t n_ub, const int64_t *merged, const GridPush *gp) {
    if (n_ub <= 0) return 0;
    hipStream_t st = c->stream;
    if (c->scan.ensure((size_t)n_ub, st) || c->lists.ensure((size_t)(2 * n_ub), st) ||
        c->tx.ensure((size_t)(2 * n_ub), st) || c->ty.ensure((size_t)(2 * n_ub), st) ||
        c->tg.ensure((size_t)(2 * n_ub), st) || c->dcount.ensure(16, st))
        return 1;
    int64_t *dn = step_dn(c), *dn1 = c->dcount.p + (c->dn_slot == 1 ? 8 : 1);
    // exclusive scan of the live flags (ballot and popcount per wave), which
    // also lists the holes and the donors
    if (c->ptmp.ensure((size_t)prim::tiles_of(n_ub) + 1, st)) return 1;
    int32_t *holes = c->lists.p, *donors = c->lists.p + n_ub;
    VX_CHECK(prim::select_flags_counted(LiveFlag{c->g.p, dn, merged}, n_ub,
                                        CompactLists{c->scan.p, holes, donors, dn, merged, dn1}, c->ptmp.p,
                                        (int32_t *)(c->dcount.p + 13), st));
    const unsigned nb2 = (unsigned)((2 * n_ub + 255) / 256);
    hipLaunchKernelGGL(k_compact_gather, dim3(nb2), dim3(256), 0, st, c->x.p, c->y.p, c->g.p, c->scan.p, dn, dn1,
                       holes, donors, c->tx.p, c->ty.p, c->tg.p, merged);
    hipLaunchKernelGGL(k_compact_scatter, dim3(nb2), dim3(256), 0, st, c->x.p, c->y.p, c->g.p, c->scan.p, dn, dn1,
                       holes, c->tx.p, c->ty.p, c->tg.p, merged, gp ? *gp : GridPush{});
    VX_CHECK(hipGetLastError());
    c->dn_slot = c->dn_slot == 1 ? 8 : 1;
    return 0;
}

// ---- A12/A13 search: src/kernel.cu:1180-1235 -------------------------------
// COL_i = smallest j > i with
//   kind 1: g_i g_j < 0 and r_ij^2 < r2
//   kind 2: g_i g_j > 0 and r_ij^2 < r2 and |g_i + g_j| < max_ve_g
// Only vortices with g != 0 can pair.  They are linked into the cells of a
// dense grid over their box, cell size s > 2 sqrt(r2): a partner of i lies in
// i's cell or in the neighbour on the side of the cell's half that holds i,
// per axis, so four cells are searched.  The minimum j does not depend on the
// order of the cell lists.  The grid is set up on the device from the box.
// Collapse work arrays (per vortex slot unless noted): pc counts the
// candidates pointing at a vortex, pcd the ones not yet published, state
// holds the decision bits, first the smallest active candidate merging with
// it; cand lists the candidates in nseg segments of cap slots (segment s
// holds ncs[s] of them), rest the ones left to resolve (count *nrest).  The
// search fills kColSeg segments, one per residue of the work-group index
// mod kColSeg, so that no single counter takes every work-group's atomic;
// a candidate list from a device select is one segment.
constexpr int kColSeg = 32;

struct ColBufs {
    int32_t *pc, *pcd, *state, *first, *cand, *rest, *nrest, *ncs;
    int64_t cap;
    int nseg;
};

// Slots of the kColSeg segments for a search over n_ub vortices in 256-thread
// work-groups.
static int64_t col_seg_cap(int64_t n_ub) { return ((n_ub + 255) / 256 + kColSeg - 1) / kColSeg * 256; }

// lists holds cand, rest, pcd and state; a compaction's scan (c->scan) and its
// holes and donors (c->lists below 2 n_ub <= ncap + n_ub) leave pc, pcd,
// state and first alone, so a pass can leave them prepared for the next.
static int col_bufs(vx_case *c, int64_t n_ub, bool segmented, ColBufs &b) {
    hipStream_t st = c->stream;
    const int64_t cap = col_seg_cap(n_ub), ncap = kColSeg * cap;
    if (c->flags.ensure((size_t)n_ub, st) || c->lists.ensure((size_t)(ncap + 3 * n_ub), st) ||
        c->misc.ensure((size_t)n_ub, st) || c->dcount.ensure(16, st) || c->colseg.ensure(kColSeg, st))
        return 1;
    b.pc = c->flags.p;
    b.cand = c->lists.p;
    b.rest = c->lists.p + ncap;
    b.pcd = c->lists.p + ncap + n_ub;
    b.state = c->lists.p + ncap + 2 * n_ub;
    b.first = c->misc.p;
    b.ncs = c->colseg.p;
    b.nrest = (int32_t *)(c->dcount.p + 4);
    b.cap = segmented ? cap : n_ub;
    b.nseg = segmented ? kColSeg : 1;
    return 0;
}

// The candidate at slot t of the segmented list, or -1.
__device__ __forceinline__ int32_t col_cand(const ColBufs &cb, int64_t t) {
    if (t >= (int64_t)cb.nseg * cb.cap) return -1;
    const int64_t sg = t / cb.cap, j = t - sg * cb.cap;
    return j < cb.ncs[sg] ? cb.cand[t] : -1;
}

// Appends val of every flagged thread to out[*cnt, ...) with one atomic per
// 256-thread work-group; every thread of the work-group calls it.  The order
// within the list is not defined.
__device__ void block_append(bool flag, int32_t val, int32_t *__restrict__ out, int32_t *__restrict__ cnt) {
    __shared__ int wcnt[4];
    __shared__ int base;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t m = __ballot(flag);
    if (lane == 0) wcnt[w] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int tot = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        base = tot ? atomicAdd(cnt, tot) : 0;
    }
    __syncthreads();
    int off = 0;
    for (int q = 0; q < w; ++q) off += wcnt[q];
    if (flag) out[base + off + __popcll(m & ((1ull << lane) - 1))] = val;
}

// box = (min x, max x, min y, max y) of the g != 0 vortices; at most
// max_cells cells (a wider cell stays correct).
__device__ Grid grid_params(const double *__restrict__ box, double r2, int64_t max_cells) {
    Grid q{};
    const double x0 = box[0], x1 = box[1], y0 = box[2], y1 = box[3];
    q.valid = (r2 > 0.0) && (x0 <= x1) && (y0 <= y1);
    if (q.valid) {
        double s = 2.0 * sqrt(r2) * 1.001;
        const double w = x1 - x0, h = y1 - y0;
        while ((floor(w / s) + 1) * (floor(h / s) + 1) > (double)max_cells) s *= 1.25;
        q.x0 = x0;
        q.y0 = y0;
        q.inv_s = 1.0 / s;
        q.nx = (int64_t)floor(w / s) + 1;
        q.ny = (int64_t)floor(h / s) + 1;
        q.cells = q.nx * q.ny;
    }
    return q;
}

// Each work-group derives the grid from the box (the same result in every
// one), work-group 0 stores it, and the cell heads are cleared to -1.  With
// cb.pc set, the collapse arrays of [0, n_ub) are prepared as well (counts
// and states 0, first INT32_MAX, both list counts 0).
// With dn_prev set, a pass whose predecessor of the same kind left the count
// unchanged (*dn == *dn_prev: it merged nothing, so the state and with it
// the pairs are the same) gets an invalid grid: the search finds no pair, as
// the reference's loop stops when n no longer changes (src/unit_1.cu:1138-1216).
__global__ void k_grid_clear(int32_t *__restrict__ head, const double *__restrict__ box, double r2,
                             int64_t max_cells, Grid *__restrict__ G, int64_t n_ub, ColBufs cb,
                             const int64_t *__restrict__ dn, const int64_t *__restrict__ dn_prev,
                             int32_t *__restrict__ pool) {
    __shared__ Grid sq;
    if (threadIdx.x == 0) {
        sq = grid_params(box, r2, max_cells);
        if (dn_prev && *dn == *dn_prev) sq.valid = false;
        if (blockIdx.x == 0) {
            *G = sq;
            if (pool) *pool = 0;
        }
    }
    __syncthreads();
    const int64_t cells = sq.valid ? sq.cells : 0;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = t0; t < cells; t += stride) head[t] = -1;
    if (!cb.pc) return;
    for (int64_t t = t0; t < n_ub; t += stride) {
        cb.pc[t] = 0;
        cb.pcd[t] = 0;
        cb.state[t] = 0;
        cb.first[t] = INT32_MAX;
    }
    if (t0 < kColSeg) cb.ncs[t0] = 0;
    if (t0 == 0) *cb.nrest = 0;
}

__global__ void k_grid_insert(const double *__restrict__ x, const double *__restrict__ y,
                              const double *__restrict__ g, const int64_t *__restrict__ dn,
                              const Grid *__restrict__ Gp, int32_t *__restrict__ head, int32_t *__restrict__ next) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const Grid G = *Gp;
    if (!G.valid || i >= *dn || g[i] == 0) return;
    int64_t cx, cy;
    double fx, fy;
    cell_of(G, x[i], y[i], cx, cy, fx, fy);
    next[i] = atomicExch(&head[cy * G.nx + cx], (int32_t)i);
}

// With cb.pc set, each pair found also counts at its target (pc, pcd) and
// lists i as a candidate.
template <int KIND>
__global__ __launch_bounds__(256) void k_setka(const double *__restrict__ x, const double *__restrict__ y,
                                               const double *__restrict__ g, const int64_t *__restrict__ dn,
                                               const Grid *__restrict__ Gp, double r2, double max_ve_g,
                                               const int32_t *__restrict__ head, const int32_t *__restrict__ next,
                                               int32_t *__restrict__ col, ColBufs cb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t best = INT64_MAX;
    if (i < *dn) {
        const Grid G = *Gp;
        const double gi = g[i];
        if (G.valid && gi != 0) {
            const double r0 = x[i], r1 = y[i];
            int64_t cx, cy;
            double fx, fy;
            cell_of(G, r0, r1, cx, cy, fx, fy);
            const int64_t ox = fx < 0.5 ? cx - 1 : cx + 1, oy = fy < 0.5 ? cy - 1 : cy + 1;
            const int64_t xs[2] = {cx, ox}, ys[2] = {cy, oy};
            int32_t hd[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t xx = xs[q & 1], yy = ys[q >> 1];
                hd[q] = (xx >= 0 && xx < G.nx && yy >= 0 && yy < G.ny) ? head[yy * G.nx + xx] : -1;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                for (int32_t j = hd[q]; j >= 0; j = next[j]) {
                    if (j <= i || j >= best) continue;
                    const double gj = g[j];
                    bool hit;
                    if (KIND == 1)
                        hit = (gi * gj < 0) && (dist2(r0, r1, x[j], y[j]) < r2);
                    else
                        hit = (gi * gj > 0) && (dist2(r0, r1, x[j], y[j]) < r2) && (fabs(gi + gj) < max_ve_g);
                    if (hit) best = j;
                }
            }
        }
        col[i] = best == INT64_MAX ? -1 : (int32_t)best;
    }
    if (!cb.pc) return;  // the same for every thread
    const bool hit = best != INT64_MAX;
    if (hit) {
        atomicAdd(&cb.pc[best], 1);
        atomicAdd(&cb.pcd[best], 1);
    }
    const int sg = (int)(blockIdx.x % kColSeg);
    block_append(hit, (int32_t)i, cb.cand + sg * cb.cap, cb.ncs + sg);
}

// ---- one grid for the passes of a step ---------------------------------------
// The passes of a step can share one grid over the box of the vortices alive
// after the move (a cell of at least 2 sqrt(r2) keeps the four-cell search
// exact).  Node k of [0, no) stands for slot k; a merge's survivor and a
// compaction's filled hole get a pool node (id no + q, slot pslot[q]) at the
// cell of their new position (grid_push), so every live slot keeps a node at
// the cell of its current position.  Nodes whose slot has since taken other
// data change nothing: the search tests each slot it reaches by the slot's
// current circulation and position and keeps the smallest index found,
// whatever else the lists hold.
__global__ void k_grid_nodes(const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ g, const int64_t *__restrict__ dn,
                             const Grid *__restrict__ Gp, int64_t no, int32_t *__restrict__ head,
                             int32_t *__restrict__ next) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= no) return;
    const Grid G = *Gp;
    if (!G.valid || i >= *dn || g[i] == 0) return;
    int64_t cx, cy;
    double fx, fy;
    cell_of(G, x[i], y[i], cx, cy, fx, fy);
    next[i] = atomicExch(&head[cy * G.nx + cx], (int32_t)i);
}

// k_setka over the step's grid, recording only the hits (the collapse reads
// COL of the listed candidates alone).  With dn_prev, a pass whose
// predecessor of the same kind left the count unchanged searches nothing.
template <int KIND>
__global__ __launch_bounds__(256) void k_setka_nodes(const double *__restrict__ x, const double *__restrict__ y,
                                                     const double *__restrict__ g, const int64_t *__restrict__ dn,
                                                     const int64_t *__restrict__ dn_prev,
                                                     const Grid *__restrict__ Gp, double r2, double max_ve_g,
                                                     const int32_t *__restrict__ head,
                                                     const int32_t *__restrict__ next, int64_t no,
                                                     const int32_t *__restrict__ pslot, int32_t *__restrict__ col,
                                                     ColBufs cb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = *dn;
    int64_t best = INT64_MAX;
    if (i < n && !(dn_prev && n == *dn_prev)) {
        const Grid G = *Gp;
        const double gi = g[i];
        if (G.valid && gi != 0) {
            const double r0 = x[i], r1 = y[i];
            int64_t cx, cy;
            double fx, fy;
            cell_of(G, r0, r1, cx, cy, fx, fy);
            const int64_t ox = fx < 0.5 ? cx - 1 : cx + 1, oy = fy < 0.5 ? cy - 1 : cy + 1;
            const int64_t xs[2] = {cx, ox}, ys[2] = {cy, oy};
            int32_t hd[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t xx = xs[q & 1], yy = ys[q >> 1];
                hd[q] = (xx >= 0 && xx < G.nx && yy >= 0 && yy < G.ny) ? head[yy * G.nx + xx] : -1;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                for (int32_t kk = hd[q]; kk >= 0; kk = next[kk]) {
                    const int32_t j = kk < no ? kk : pslot[kk - no];
                    if (j <= i || j >= best) continue;
                    const double gj = g[j];
                    bool hit;
                    if (KIND == 1)
                        hit = (gi * gj < 0) && (dist2(r0, r1, x[j], y[j]) < r2);
                    else
                        hit = (gi * gj > 0) && (dist2(r0, r1, x[j], y[j]) < r2) && (fabs(gi + gj) < max_ve_g);
                    if (hit) best = j;
                }
            }
        }
    }
    const bool hit = best != INT64_MAX;
    if (hit) {
        col[i] = (int32_t)best;
        atomicAdd(&cb.pc[best], 1);
        atomicAdd(&cb.pcd[best], 1);
    }
    const int sg = (int)(blockIdx.x % kColSeg);
    block_append(hit, (int32_t)i, cb.cand + sg * cb.cap, cb.ncs + sg);
}

// Box of the vortices with g != 0 in [0, *dn) (min x, -max x, min y, -max y
// partials, then the box into out[0..3]).
__global__ __launch_bounds__(256) void k_live_box(const double *__restrict__ x, const double *__restrict__ y,
                                                  const double *__restrict__ g, const int64_t *__restrict__ dn,
                                                  double *__restrict__ part) {
    __shared__ double red[4][256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double b[4] = {1e300, 1e300, 1e300, 1e300};
    if (i < *dn && g[i] != 0) {
        b[0] = x[i];
        b[1] = -x[i];
        b[2] = y[i];
        b[3] = -y[i];
    }
    for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = b[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = fmin(red[q][threadIdx.x], red[q][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 4) part[(int64_t)blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_live_box_final(const double *__restrict__ part, int64_t nb,
                                                        double *__restrict__ out) {
    __shared__ double red[4][256];
    double b[4] = {1e300, 1e300, 1e300, 1e300};
    for (int64_t k = threadIdx.x; k < nb; k += 256)
        for (int q = 0; q < 4; ++q) b[q] = fmin(b[q], part[k * 4 + q]);
    for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = b[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = fmin(red[q][threadIdx.x], red[q][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = red[0][0];
        out[1] = -red[1][0];
        out[2] = red[2][0];
        out[3] = -red[3][0];
    }
}

int launch_live_box(vx_case *c) {
    const int64_t n = c->n > 0 ? c->n : 1;
    const int64_t nb = (n + 255) / 256;
    if (c->partial.ensure((size_t)(4 * nb), c->stream) || c->red.ensure(8, c->stream)) return 1;
    hipLaunchKernelGGL(k_live_box, dim3((unsigned)nb), dim3(256), 0, c->stream, c->x.p, c->y.p, c->g.p, step_dn(c),
                       c->partial.p);
    hipLaunchKernelGGL(k_live_box_final, dim3(1), dim3(256), 0, c->stream, c->partial.p, nb, c->red.p + 4);
    VX_CHECK(hipGetLastError());
    return 0;
}

int launch_setka(vx_case *c, int kind, int64_t n_ub, bool prepare, bool after_pass) {
    hipStream_t st = c->stream;
    if (c->col.ensure((size_t)(n_ub > 0 ? n_ub : 1), st)) return 1;
    if (n_ub <= 0) return 0;
    ColBufs cb{};
    if (prepare && col_bufs(c, n_ub, true, cb)) return 1;
    const unsigned nb = (unsigned)((n_ub + 255) / 256);
    const double ve2 = c->prm.ve_size * c->prm.ve_size;
    const double rr = kind == 1 ? c->prm.r_col_diff_sign : c->prm.r_col_same_sign;
    const double r2 = rr * rr * ve2;  // src/unit_1.cu:1228-1229
    // at most max(2^20, 8 n) cells of heads
    const int64_t max_cells = std::max<int64_t>(1 << 20, 8 * n_ub);
    if (c->vals.ensure((size_t)max_cells, st) || c->vals_alt.ensure((size_t)n_ub, st) ||
        c->gridp.ensure(sizeof(Grid), st))
        return 1;
    Grid *G = (Grid *)c->gridp.p;
    int64_t *dn = step_dn(c);
    // after a pass of the same kind, the other slot holds the count before it
    const int64_t *dn_prev = after_pass ? c->dcount.p + (c->dn_slot == 1 ? 8 : 1) : nullptr;
    hipLaunchKernelGGL(k_grid_clear, dim3(1024), dim3(256), 0, st, c->vals.p, c->red.p + 4, r2, max_cells, G, n_ub,
                       cb, (const int64_t *)dn, dn_prev, (int32_t *)nullptr);
    hipLaunchKernelGGL(k_grid_insert, dim3(nb), dim3(256), 0, st, c->x.p, c->y.p, c->g.p, dn, G, c->vals.p,
                       c->vals_alt.p);
    if (kind == 1)
        hipLaunchKernelGGL(k_setka<1>, dim3(nb), dim3(256), 0, st, c->x.p, c->y.p, c->g.p, dn, G, r2,
                           c->prm.max_ve_g, c->vals.p, c->vals_alt.p, c->col.p, cb);
    else
        hipLaunchKernelGGL(k_setka<2>, dim3(nb), dim3(256), 0, st, c->x.p, c->y.p, c->g.p, dn, G, r2,
                           c->prm.max_ve_g, c->vals.p, c->vals_alt.p, c->col.p, cb);
    VX_CHECK(hipGetLastError());
    return 0;
}

// ---- A12/A13 merge: src/kernel.cu:1237-1271 --------------------------------
// The sequential sweep (for i ascending: if COL_i >= 0 merge j = COL_i into i
// and kill j) is resolved without a sweep.  Candidate i is active when no
// active candidate i' < i has COL_i' = i; this is decided leaf-first over the
// forest of edges i -> COL_i (COL_i > i) in one work-group.  Among active
// candidates sharing a target j the smallest merges with j as it was, the
// others with the killed j (g = 0 at 1e10), exactly as the sweep does.
// Preparation and counting when COL comes from elsewhere than k_setka.
__global__ void k_col_prepare(int64_t n_ub, ColBufs cb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *cb.nrest = 0;
    if (i < kColSeg) cb.ncs[i] = 0;
    if (i >= n_ub) return;
    cb.pc[i] = 0;
    cb.pcd[i] = 0;
    cb.state[i] = 0;
    cb.first[i] = INT32_MAX;
}

__global__ __launch_bounds__(256) void k_col_count(const int32_t *__restrict__ col, const int64_t *__restrict__ dn,
                                                   ColBufs cb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t j = i < *dn ? col[i] : -1;
    if (j >= 0) {
        atomicAdd(&cb.pc[j], 1);
        atomicAdd(&cb.pcd[j], 1);
    }
}

// Candidates in index order: i < *dn with COL_i >= 0.
struct ColCandidate {
    const int32_t *col;
    const int64_t *dn;
    __device__ bool operator()(int32_t i) const { return i < *dn && col[i] >= 0; }
};

// The pass's resolution and merge, in one work-group, phase after phase
// (state bits: 1 active, 2 inactive, 4 killed by an active child, 16
// published; every access to state, first and pcd goes through L2 atomics,
// so no stale per-CU L1 line is read):
//  1. Leaves of the forest (no candidate points at them: pc 0) are active,
//     as nothing precedes them, and are published at once: their target is
//     marked killed, its first merger recorded and its pending count pcd
//     lowered.  pc itself is only read, so every leaf sees the counts from
//     before any publication.  The other candidates are listed in rest.
//  2. The rest are decided in rounds: a candidate is decided once pcd, its
//     count of unpublished candidates pointing at it, reaches 0.  Each round
//     decides at least the lowest undecided candidate (its children have
//     lower indices and were published in earlier rounds), so ncand + 1
//     rounds always suffice; the loop bound is a guard, reported through
//     *err (bit 1).
//  3. Each active candidate i merges with j = COL_i: the smallest active
//     candidate of j (first[j] == i) with j as it was, then kills j; the
//     others with the killed j (1e10, 1e10, 0).  No two threads touch one
//     slot: an active i is never the target of an active candidate (that
//     would have made it inactive), so x[i] is read and written by thread i
//     alone, and x[j] by j's first merger alone.  A kept grid gets a node
//     for each survivor at its new cell.
//  4. With reset, every slot the pass touched (the candidates and their
//     targets) gets its prepared work-array values back and the list counts
//     are zeroed, so the next pass's search needs no preparation.
// The candidates are the cb.nseg segments of the search's lists; *merged
// receives their count.
template <int KIND>
__global__ __launch_bounds__(1024) void k_col_pass(double *__restrict__ x, double *__restrict__ y,
                                                   double *__restrict__ g, const int32_t *__restrict__ col, ColBufs cb,
                                                   int64_t *__restrict__ merged, int32_t *__restrict__ err,
                                                   GridPush gp, int reset) {
    __shared__ int64_t segoff[kColSeg + 1];
    __shared__ int32_t nrest;
    __shared__ int undecided;
    if (threadIdx.x == 0) {
        int64_t k = 0;
        for (int q = 0; q < cb.nseg; ++q) {
            segoff[q] = k;
            k += cb.ncs[q];
        }
        segoff[cb.nseg] = k;
        *merged = k;
        nrest = 0;
    }
    __syncthreads();
    const int64_t K = segoff[cb.nseg];
    auto cand = [&](int64_t t) -> int32_t {
        int q = 0;
        while (segoff[q + 1] <= t) ++q;
        return cb.cand[q * cb.cap + (t - segoff[q])];
    };
    int32_t *__restrict__ state = cb.state, *__restrict__ first = cb.first, *__restrict__ pcd = cb.pcd;
    for (int64_t t = threadIdx.x; t < K; t += blockDim.x) {
        const int32_t i = cand(t);
        if (cb.pc[i] == 0) {
            const int32_t j = col[i];
            atomicOr(&state[i], 1 | 16);
            atomicOr(&state[j], 4);
            atomicMin(&first[j], i);
            atomicSub(&pcd[j], 1);
        } else {
            cb.rest[atomicAdd(&nrest, 1)] = i;
        }
    }
    __syncthreads();
    const int64_t ncand = nrest;
    for (int64_t round = 0;; ++round) {
        if (threadIdx.x == 0) undecided = 0;
        __syncthreads();
        for (int64_t t = threadIdx.x; t < ncand; t += blockDim.x) {
            const int32_t i = cb.rest[t];
            const int32_t s = atomicOr(&state[i], 0);
            if (s & 3) continue;
            if (atomicAdd(&pcd[i], 0) == 0)
                atomicOr(&state[i], (s & 4) ? 2 : 1);
            else
                undecided = 1;
        }
        __syncthreads();
        for (int64_t t = threadIdx.x; t < ncand; t += blockDim.x) {
            const int32_t i = cb.rest[t];
            const int32_t s = atomicOr(&state[i], 0);
            if ((s & 3) && !(s & 16)) {
                const int32_t j = col[i];
                if (s & 1) {
                    atomicOr(&state[j], 4);
                    atomicMin(&first[j], i);
                }
                atomicOr(&state[i], 16);
                atomicSub(&pcd[j], 1);
            }
        }
        __syncthreads();
        const int more = undecided;
        __syncthreads();
        if (!more) break;
        if (round > ncand + 1) {
            if (threadIdx.x == 0) atomicOr(err, 1);
            break;
        }
    }
    for (int64_t t = threadIdx.x; t < K; t += blockDim.x) {
        const int32_t i = cand(t);
        if (!(atomicOr(&state[i], 0) & 1)) continue;
        const int32_t j = col[i];
        const bool fst = atomicOr(&first[j], 0) == i;
        double xj = 1e+10, yj = 1e+10, gj = 0.0;
        if (fst) {
            xj = x[j];
            yj = y[j];
            gj = g[j];
        }
        const double xi = x[i], yi = y[i], gi = g[i];
        double nx, ny, ng;
        if (KIND == 1) {
            const double mn = fabs(gi) + fabs(gj);
            nx = (xi * fabs(gi) + xj * fabs(gj)) / fabs(mn);
            ny = (yi * fabs(gi) + yj * fabs(gj)) / fabs(mn);
            ng = gi + gj;
        } else {
            ng = gi + gj;
            nx = (xi * gi + xj * gj) / ng;
            ny = (yi * gi + yj * gj) / ng;
        }
        x[i] = nx;
        y[i] = ny;
        g[i] = ng;
        if (fst) {
            x[j] = 1e+10;
            y[j] = 1e+10;
            g[j] = 0.0;
        }
        if (gp.grid) grid_push(gp, i, nx, ny);  // the survivor may have left its cell
    }
    if (!reset) return;
    __syncthreads();
    for (int64_t t = threadIdx.x; t < K; t += blockDim.x) {
        const int32_t i = cand(t), j = col[i];
        cb.pc[i] = 0;
        pcd[i] = 0;
        state[i] = 0;
        first[i] = INT32_MAX;
        cb.pc[j] = 0;
        pcd[j] = 0;
        state[j] = 0;
        first[j] = INT32_MAX;
    }
    if (threadIdx.x < cb.nseg) cb.ncs[threadIdx.x] = 0;
    if (threadIdx.x == 0) *cb.nrest = 0;
}

int launch_collapse(vx_case *c, int kind, int64_t n_ub, bool prepared, const GridPush *gp, bool reset) {
    if (n_ub <= 0) return 0;
    hipStream_t st = c->stream;
    ColBufs cb{};
    if (col_bufs(c, n_ub, prepared, cb)) return 1;
    const unsigned nb = (unsigned)((n_ub + 255) / 256);
    if (!prepared) {
        hipLaunchKernelGGL(k_col_prepare, dim3(nb), dim3(256), 0, st, n_ub, cb);
        hipLaunchKernelGGL(k_col_count, dim3(nb), dim3(256), 0, st, c->col.p, step_dn(c), cb);
        const ColCandidate pred{c->col.p, step_dn(c)};
        if (c->ptmp.ensure((size_t)prim::tiles_of(n_ub) + 1, st)) return 1;
        VX_CHECK(prim::select_flags(prim::pred_flag(pred), n_ub, prim::WriteIndex{cb.cand}, c->ptmp.p, cb.ncs, st));
    }
    const GridPush push = gp ? *gp : GridPush{};
    int32_t *err = (int32_t *)(c->dcount.p + 3);
    if (kind == 1)
        hipLaunchKernelGGL(k_col_pass<1>, dim3(1), dim3(1024), 0, st, c->x.p, c->y.p, c->g.p, c->col.p, cb,
                           col_merged(c), err, push, (int)reset);
    else
        hipLaunchKernelGGL(k_col_pass<2>, dim3(1), dim3(1024), 0, st, c->x.p, c->y.p, c->g.p, c->col.p, cb,
                           col_merged(c), err, push, (int)reset);
    VX_CHECK(hipGetLastError());
    return 0;
}

// Below this many slots (the default of tune.col_share_min) a grid costs
// little more than its two launches, and the passes rebuild it: a pass that
// merges much would otherwise lengthen the lists the next pass walks (dead
// slots, survivors' second nodes).
constexpr int64_t kColGridShareMin = 1 << 17;

// The step's collapse (src/unit_1.cu:1138-1216): for each kind, n_col passes
// of search, resolution and merge, each followed by its compaction.  From
// tune.col_share_min slots on, the passes share a grid built after the move (see
// k_grid_nodes): one for both kinds when their radii are equal, else one per
// kind with cells for its own radius (cells sized for the larger radius would
// lengthen the other kind's list walks); below it every pass builds its own.
// A pass pushes at most one node per survivor and one per filled hole, two
// per slot, so the pool holds 2 n_ub nodes for each pass over a grid; a full
// pool is reported through the collapse error flag (bit 2).
int launch_step_collapse(vx_case *c, int64_t n_ub) {
    if (n_ub <= 0 || c->prm.n_col <= 0) return 0;
    hipStream_t st = c->stream;
    const double ve2 = c->prm.ve_size * c->prm.ve_size;
    const double r2d = c->prm.r_col_diff_sign * c->prm.r_col_diff_sign * ve2;  // src/unit_1.cu:1228-1229
    const double r2s = c->prm.r_col_same_sign * c->prm.r_col_same_sign * ve2;
    const bool keep = n_ub >= (c->tune.col_share_min > 0 ? c->tune.col_share_min : kColGridShareMin);
    const bool shared = keep && r2d == r2s;
    const int64_t passes = keep ? (shared ? 2 : 1) * c->prm.n_col : 1;  // passes over one grid
    const int64_t pcap = std::min<int64_t>(2 * passes * n_ub, (int64_t)INT32_MAX - n_ub - 1);
    const int64_t max_cells = std::max<int64_t>(1 << 20, 8 * n_ub);
    if (c->col.ensure((size_t)n_ub, st) || c->vals.ensure((size_t)max_cells, st) ||
        c->cnode.ensure((size_t)(n_ub + 2 * pcap), st) || c->gridp.ensure(sizeof(Grid), st) ||
        c->dcount.ensure(16, st))
        return 1;
    Grid *G = (Grid *)c->gridp.p;
    int32_t *head = c->vals.p, *next = c->cnode.p, *pslot = c->cnode.p + n_ub + pcap;
    int32_t *pcount = (int32_t *)(c->dcount.p + 12);
    const GridPush gp{G, head, next, pslot, pcount, (int32_t *)(c->dcount.p + 3), n_ub, pcap};
    const unsigned nb = (unsigned)((n_ub + 255) / 256);
    for (int kind = 1; kind <= 2; ++kind) {
        const double r2 = kind == 1 ? r2d : r2s;
        for (int64_t cc = 0; cc < c->prm.n_col; ++cc) {
            ColBufs cb{};
            if (col_bufs(c, n_ub, true, cb)) return 1;
            if (!keep || (cc == 0 && (kind == 1 || !shared))) {
                // heads cleared, pool emptied, the pass's work arrays prepared
                // and every live slot linked at its cell
                hipLaunchKernelGGL(k_grid_clear, dim3(1024), dim3(256), 0, st, head, c->red.p + 4, r2, max_cells,
                                   G, n_ub, cb, (const int64_t *)step_dn(c), (const int64_t *)nullptr, pcount);
                hipLaunchKernelGGL(k_grid_nodes, dim3(nb), dim3(256), 0, st, c->x.p, c->y.p, c->g.p, step_dn(c), G,
                                   n_ub, head, next);
            }  // else the work arrays were left prepared by the last pass (k_col_pass with reset)
            // after a pass of the same kind, the other slot holds the count before it
            const int64_t *dn = step_dn(c);
            const int64_t *dn_prev = cc > 0 ? c->dcount.p + (c->dn_slot == 1 ? 8 : 1) : nullptr;
            if (kind == 1)
                hipLaunchKernelGGL(k_setka_nodes<1>, dim3(nb), dim3(256), 0, st, c->x.p, c->y.p, c->g.p, dn,
                                   dn_prev, G, r2, c->prm.max_ve_g, head, next, n_ub, pslot, c->col.p, cb);
            else
                hipLaunchKernelGGL(k_setka_nodes<2>, dim3(nb), dim3(256), 0, st, c->x.p, c->y.p, c->g.p, dn,
                                   dn_prev, G, r2, c->prm.max_ve_g, head, next, n_ub, pslot, c->col.p, cb);
            VX_CHECK(hipGetLastError());
            // a rebuilt grid needs no pushes, and its clear prepares the work arrays
            const GridPush *push = keep ? &gp : nullptr;
            if (launch_collapse(c, kind, n_ub, true, push, keep) || launch_compact(c, n_ub, col_merged(c), push))
                return 1;
        }
    }
    return 0;
}
}  // namespace vx
