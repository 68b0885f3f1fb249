// Panel system: influence matrix (A1), its inverse (A2) and the per-step
// birth of vortices from it (A4).  All FP64; multiplies and adds are kept
// separately rounded so the Gauss-Jordan elimination and the newborn
// positions reproduce the FP64 restatement bit for bit.
#pragma clang fp contract(off)

#include "vx_internal.h"

namespace vx {

// ---- A1: src/unit_1.cu:256-373 --------------------------------------------
// L[i][j] = tau_i . vbar_ij, vbar_ij the average over panel i of the velocity
// of a unit sheet on panel j (three atan differences q1_k, three log ratios
// q2_k, vectors c_k); both panels reversed for the right-hand neighbour and
// q1_1 = q2_1 = 0 for either neighbour; L[i][i] = -0.5; row P = lengths,
// column P = 1, L[P][P] = 0.
__device__ __forceinline__ double dot2(double a0, double a1, double b0, double b1) { return a0 * b0 + a1 * b1; }

__global__ void k_panel_matrix(PanelsSoA pan, double *__restrict__ L) {
    const int64_t P = pan.P, m = P + 1;
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = blockIdx.y;
    if (j > P || i > P) return;
    if (i == P) {
        L[P * m + j] = j < P ? pan.len[j] : 0.0;
        return;
    }
    if (j == P) {
        L[i * m + P] = 1.0;
        return;
    }
    if (j == i) {
        L[i * m + j] = -0.5;
        return;
    }
    double a00 = pan.lx[i], a01 = pan.ly[i], a10 = pan.rx[i], a11 = pan.ry[i];
    double b00 = pan.lx[j], b01 = pan.ly[j], b10 = pan.rx[j], b11 = pan.ry[j];
    const bool next = (j == i + 1) || (j == 0 && i == P - 1);
    const bool prev = (j == i - 1) || (i == 0 && j == P - 1);
    if (next) {
        a10 = pan.lx[i]; a11 = pan.ly[i]; a00 = pan.rx[i]; a01 = pan.ry[i];
        b10 = pan.lx[j]; b11 = pan.ly[j]; b00 = pan.rx[j]; b01 = pan.ry[j];
    }
    const double dj0 = b10 - b00, dj1 = b11 - b01;
    const double di0 = a10 - a00, di1 = a11 - a01;
    const double p10 = a00 - b10, p11 = a01 - b11;
    const double p20 = a10 - b10, p21 = a11 - b11;
    const double s10 = a00 - b00, s11 = a01 - b01;
    const double s20 = a10 - b00, s21 = a11 - b01;
    const double z1 = p10 * p21 - p11 * p20;
    const double z2 = s10 * s21 - s11 * s20;
    const double z3 = s20 * p21 - s21 * p20;
    double q11 = 0.0, q21 = 0.0;
    if (!(next || prev)) {
        q11 = atan(dot2(di0, di1, p10, p11) / z1) - atan(dot2(di0, di1, p20, p21) / z1);
        q21 = 0.5 * log(dot2(p20, p21, p20, p21) / dot2(p10, p11, p10, p11));
    }
    const double q12 = atan(dot2(di0, di1, s20, s21) / z2) - atan(dot2(di0, di1, s10, s11) / z2);
    const double q13 = atan(dot2(dj0, dj1, p20, p21) / z3) - atan(dot2(dj0, dj1, s20, s21) / z3);
    const double q22 = 0.5 * log(dot2(s10, s11, s10, s11) / dot2(s20, s21, s20, s21));
    const double q23 = 0.5 * log(dot2(p20, p21, p20, p21) / dot2(s20, s21, s20, s21));
    const double djp1 = dot2(dj0, dj1, p10, p11);
    const double dis1 = dot2(di0, di1, s10, s11);
    const double didj = dot2(di0, di1, dj0, dj1);
    const double c10 = djp1 * di0 + dis1 * dj0 - didj * s10;
    const double c11 = djp1 * di1 + dis1 * dj1 - didj * s11;
    const double djj = dot2(dj0, dj1, dj0, dj1);
    const double dii = dot2(di0, di1, di0, di1);
    const double c20 = c10 + djj * di0, c21 = c11 + djj * di1;
    const double c30 = dii * dj0, c31 = dii * dj1;
    const double k = 1.0 / (2 * kPi * sqrt(djj) * dii);
    const double v0 = k * (q11 * c10 + q12 * c20 + q13 * c30 + (q21 * c11 + q22 * c21 + q23 * c31));
    const double v1 = k * (q11 * c11 + q12 * c21 + q13 * c31 - (q21 * c10 + q22 * c20 + q23 * c30));
    L[i * m + j] = dot2(v0, v1, pan.tx[i], pan.ty[i]);
}

int launch_panel_matrix(vx_case *c) {
    const int64_t m = c->P + 1;
    if (c->L.ensure((size_t)(m * m), c->stream)) return 1;
    hipLaunchKernelGGL(k_panel_matrix, dim3((unsigned)((m + 255) / 256), (unsigned)m), dim3(256), 0, c->stream,
                       c->pan, c->L.p);
    VX_CHECK(hipGetLastError());
    return 0;
}

// ---- A2: Gauss-Jordan inverse with partial pivoting ------------------------
// (src/unit_1.cu:115-254 eliminates with a look-ahead pivot choice; the
// standard choice -- largest |A[i][k]| over i >= k once columns < k are
// eliminated, first index on ties -- is used here because the look-ahead rule
// breaks down on the bordered system of thin closed bodies.)
// Per column k: pivot search (one work-group), row swap, column factors,
// rank-1 update of both A and the inverse over all columns.
__global__ __launch_bounds__(1024) void k_gj_pivot(const double *__restrict__ A, int64_t m, int64_t k,
                                                   int64_t *__restrict__ piv, int32_t *__restrict__ bad) {
    __shared__ double bv[1024];
    __shared__ int64_t bi[1024];
    double best = -1.0;
    int64_t besti = m;
    for (int64_t i = k + threadIdx.x; i < m; i += blockDim.x) {
        const double a = fabs(A[i * m + k]);
        if (a > best) { best = a; besti = i; }
    }
    bv[threadIdx.x] = best;
    bi[threadIdx.x] = besti;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const double ov = bv[threadIdx.x + s];
            const int64_t oi = bi[threadIdx.x + s];
            if (ov > bv[threadIdx.x] || (ov == bv[threadIdx.x] && oi < bi[threadIdx.x])) {
                bv[threadIdx.x] = ov;
                bi[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *piv = bi[0];
        if (!(bv[0] >= kDelt)) *bad = 1;
    }
}

__global__ void k_gj_swap_factor(double *__restrict__ A, double *__restrict__ I, int64_t m, int64_t k,
                                 const int64_t *__restrict__ piv) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t p = *piv;
    if (j >= m || p == k || p >= m) return;
    double t = A[k * m + j]; A[k * m + j] = A[p * m + j]; A[p * m + j] = t;
    t = I[k * m + j]; I[k * m + j] = I[p * m + j]; I[p * m + j] = t;
}

__global__ void k_gj_colfactor(const double *__restrict__ A, int64_t m, int64_t k, double *__restrict__ cf) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    cf[i] = (i == k) ? 0.0 : A[i * m + k] / A[k * m + k];
}

__global__ void k_gj_update(double *__restrict__ A, double *__restrict__ I, int64_t m, int64_t k,
                            const double *__restrict__ cf) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = blockIdx.y;
    if (j >= m || i == k) return;
    const double c = cf[i];
    A[i * m + j] = A[i * m + j] - c * A[k * m + j];
    I[i * m + j] = I[i * m + j] - c * I[k * m + j];
}

__global__ void k_gj_identity(double *__restrict__ I, int64_t m) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = blockIdx.y;
    if (j < m) I[i * m + j] = (i == j) ? 1.0 : 0.0;
}

__global__ void k_gj_scale(const double *__restrict__ A, double *__restrict__ I, int64_t m) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = blockIdx.y;
    if (j < m) I[i * m + j] = I[i * m + j] / A[i * m + i];
}

int launch_invert(vx_case *c) {
    const int64_t m = c->P + 1;
    DBuf<double> A;
    if (A.ensure((size_t)(m * m), c->stream) || c->M.ensure((size_t)(m * m), c->stream) ||
        c->red.ensure((size_t)m, c->stream) || c->dcount.ensure(4, c->stream)) {
        A.release();
        return 1;
    }
    int rc = 0;
    do {
        if (hipMemcpyAsync(A.p, c->L.p, sizeof(double) * m * m, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
            hipMemsetAsync(c->dcount.p, 0, sizeof(int64_t) * 4, c->stream) != hipSuccess) {
            set_error("invert: staging failed");
            rc = 1;
            break;
        }
        int64_t *piv = c->dcount.p;
        int32_t *bad = (int32_t *)(c->dcount.p + 1);
        const dim3 g2((unsigned)((m + 255) / 256), (unsigned)m);
        const unsigned g1 = (unsigned)((m + 255) / 256);
        hipLaunchKernelGGL(k_gj_identity, g2, dim3(256), 0, c->stream, c->M.p, m);
        for (int64_t k = 0; k < m; ++k) {
            hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(1024), 0, c->stream, A.p, m, k, piv, bad);
            hipLaunchKernelGGL(k_gj_swap_factor, dim3(g1), dim3(256), 0, c->stream, A.p, c->M.p, m, k, piv);
            hipLaunchKernelGGL(k_gj_colfactor, dim3(g1), dim3(256), 0, c->stream, A.p, m, k, c->red.p);
            hipLaunchKernelGGL(k_gj_update, g2, dim3(256), 0, c->stream, A.p, c->M.p, m, k, c->red.p);
        }
        hipLaunchKernelGGL(k_gj_scale, g2, dim3(256), 0, c->stream, A.p, c->M.p, m);
        if (hipGetLastError() != hipSuccess) {
            set_error("invert: launch failed");
            rc = 1;
            break;
        }
        int32_t hbad = 0;
        if (hipMemcpyAsync(&hbad, bad, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
            set_error("invert: read-back failed");
            rc = 1;
            break;
        }
        if (hbad) {
            set_error("panel matrix is singular (pivot below 1e-12)");
            rc = 1;
        }
    } while (0);
    (void)hipStreamSynchronize(c->stream);
    A.release();
    return rc;
}

// ---- A4: src/kernel.cu:849-873 ---------------------------------------------
// gsheet_i = sum_{j<P} M[i][j] R_j + M[i][P] d_g  (one wave64 per row,
// coalesced over the row, xor-shuffle reduction), then newborn i at the panel
// midpoint + offset * normal with g = (gsheet_i l_i + gsheet_next l_next)/2.
__global__ __launch_bounds__(256) void k_birth_gemv(const double *__restrict__ M, const double *__restrict__ R,
                                                   double d_g_host, const double *__restrict__ d_g_dev, int64_t P,
                                                   double *__restrict__ gs) {
    const double d_g = d_g_dev ? *d_g_dev : d_g_host;
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= P) return;
    const int64_t m = P + 1;
    const double *Mr = M + row * m;
    double acc = 0.0;
    for (int64_t j = lane; j < P; j += 64) acc += Mr[j] * R[j];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) gs[row] = acc + Mr[P] * d_g;
}

// Newborn positions, one per panel: the panel midpoint moved off along the
// normal (computed once per case; bp holds x at [0, P), y at [P, 2P)).
__global__ void k_birth_pos(PanelsSoA pan, double off_rel, double *__restrict__ bp) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pan.P) return;
    const double off = off_rel > 0.0 ? off_rel * pan.len[i] : 1e-6;
    bp[i] = (pan.lx[i] + pan.rx[i]) / 2 + off * pan.nx[i];
    bp[pan.P + i] = (pan.ly[i] + pan.ry[i]) / 2 + off * pan.ny[i];
}

int launch_birth_positions(vx_case *c) {
    hipLaunchKernelGGL(k_birth_pos, dim3((unsigned)((c->P + 255) / 256)), dim3(256), 0, c->stream, c->pan,
                       c->prm.birth_offset_rel, c->bpos);
    VX_CHECK(hipGetLastError());
    return 0;
}

__global__ void k_birth_place(const double *__restrict__ gs, PanelsSoA pan, const double *__restrict__ bp, Cnt cn,
                              int64_t cap, double *__restrict__ x, double *__restrict__ y, double *__restrict__ g) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = cn.get();
    if (i >= pan.P || n + i >= cap) return;
    const int64_t nx = pan.rp[i];
    x[n + i] = bp[i];
    y[n + i] = bp[pan.P + i];
    g[n + i] = 0.5 * (gs[i] * pan.len[i] + gs[nx] * pan.len[nx]);
}

// The sheet strengths of the last birth: the inverse's product, or with the
// FP32 LU the first P entries of the refined solution.
const double *birth_sheet(vx_case *c) { return c->lu ? lu_solution(c) : c->gsheet.p; }

// In device-count mode the absorbed circulation comes from red[0] (the
// previous step's sum) and the newborns go to the device count's slots,
// which then grows by P.
int launch_birth(vx_case *c, double d_g) {
    const int64_t P = c->P;
    const double *dgp = c->devcount ? c->red.p : nullptr;
    if (c->lu) {
        if (launch_lu_solve(c, c->rp.p, d_g, dgp)) return 1;
    } else {
        if (c->gsheet.ensure((size_t)P, c->stream)) return 1;
        hipLaunchKernelGGL(k_birth_gemv, dim3((unsigned)((P + 3) / 4)), dim3(256), 0, c->stream, c->M.p, c->rp.p,
                           d_g, dgp, P, c->gsheet.p);
    }
    hipLaunchKernelGGL(k_birth_place, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, c->stream, birth_sheet(c),
                       c->pan, c->bpos, cnt_n(c), c->cap, c->x.p, c->y.p, c->g.p);
    VX_CHECK(hipGetLastError());
    if (c->devcount && launch_count_add(c, c->dcount.p + 1, P)) return 1;
    return 0;
}

}  // namespace vx
