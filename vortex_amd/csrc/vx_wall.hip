// A8: wall diffusive velocity (src/kernel.cu:1031-1088, quadrature of
// I_0, I_3 at :1300-1346).
//
// For target i with characteristic distance d_i:
//   II0 = 2 pi d^2, II3 = 0; for each panel f in index order, r = |r_i - b_f|
//   with b_f the profile file's birth point:
//     0.1 l_f < r < 5 l_f : II0 -= d * RES0_f, II3 -= RES3_f  (midpoint rule
//                           with n_of_points nodes, normal -n_f)
//     r <= 0.1 l_f        : II0 = pi d^2, II3 = 2 n_f d (1 - e^{-l_f/2d}),
//                           and the panel loop stops there.
//   V_i += nu II3 / II0.
// Targets outside the box of birth points grown by 5 max(l) meet no panel and
// are skipped without touching the panel list.

#include "vx_internal.h"
#include "vx_prim.h"

namespace vx {

template <typename T>
__device__ __forceinline__ void quad_i0_i3(T ra0, T ra1, T rb0, T rb1, T n0, T n1, T rj0, T rj1, T dL, T dd,
                                           int64_t N, T &res0, T &res30, T &res31) {
    T s0 = 0, s30 = 0, s31 = 0;
    const T dR0 = (rb0 - ra0) / (T)N, dR1 = (rb1 - ra1) / (T)N;
    const T delt = dL / (T)N;
    for (int64_t k = 0; k < N; ++k) {
        const T rk0 = (ra0 + (T)k * dR0 + ra0 + (T)(k + 1) * dR0) / T(2);
        const T rk1 = (ra1 + (T)k * dR1 + ra1 + (T)(k + 1) * dR1) / T(2);
        const T e0 = (rj0 - rk0) / dd, e1 = (rj1 - rk1) / dd;
        const T e2 = e0 * e0 + e1 * e1;
        const T me = sqrt(e2);
        const T ex = exp(-me);
        s0 += (e0 * n0 + e1 * n1) / e2 * (me + T(1)) * ex * delt;
        s30 += n0 * ex * delt;
        s31 += n1 * ex * delt;
    }
    res0 = s0;
    res30 = s30;
    res31 = s31;
}

// One target per wave: the lanes take panels lane, lane + 64, ...  The first panel in
// index order with r <= 0.1 l decides the result alone (the sequential loop
// overwrites II0 and II3 there and stops); without one, the mid-range terms
// are summed, lane partials in FP64 and then across the wave.  FP32 works in
// coordinates relative to the target (differences taken in FP64), so that a
// target a fraction of a short panel away keeps FP32's relative precision.
template <typename T>
__device__ void wave_wall(const PanelsSoA &pan, int lane, double x, double y, T dd, int64_t npts, double &II0,
                          double &II30, double &II31) {
    constexpr bool F32 = sizeof(T) == 4;
    const double ox = F32 ? x : 0.0, oy = F32 ? y : 0.0;
    const T a0 = (T)(x - ox), a1 = (T)(y - oy);
    double c0 = 0.0, c30 = 0.0, c31 = 0.0;
    int64_t fb = pan.P;
    for (int64_t f = lane; f < pan.P; f += 64) {
        const T dL = (T)pan.len[f];
        const T q0 = F32 ? (T)(x - pan.bx[f]) : a0 - (T)pan.bx[f];
        const T q1 = F32 ? (T)(y - pan.by[f]) : a1 - (T)pan.by[f];
        const T r = sqrt(q0 * q0 + q1 * q1);
        if ((r < T(5) * dL) && (r > T(0.1) * dL)) {
            T R0, R30, R31;
            quad_i0_i3<T>((T)(pan.lx[f] - ox), (T)(pan.ly[f] - oy), (T)(pan.rx[f] - ox), (T)(pan.ry[f] - oy),
                          -(T)pan.nx[f], -(T)pan.ny[f], a0, a1, dL, dd, npts, R0, R30, R31);
            c0 += (double)((-dd) * R0);
            c30 += (double)R30;
            c31 += (double)R31;
        } else if (r <= T(0.1) * dL) {
            fb = f;
            break;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const int64_t ofb = __shfl_xor(fb, o, 64);
        fb = ofb < fb ? ofb : fb;
    }
    if (fb < pan.P) {
        const T dL = (T)pan.len[fb];
        const T ex = T(1) - exp(-dL / (T(2) * dd));
        II0 = (double)(T(kPi) * dd * dd);
        II30 = (double)(T(2) * (T)pan.nx[fb] * dd * ex);
        II31 = (double)(T(2) * (T)pan.ny[fb] * dd * ex);
        return;
    }
    for (int o = 32; o > 0; o >>= 1) {
        c0 += __shfl_xor(c0, o, 64);
        c30 += __shfl_xor(c30, o, 64);
        c31 += __shfl_xor(c31, o, 64);
    }
    II0 = (double)(T(2 * kPi) * dd * dd) + c0;
    II30 = -c30;
    II31 = -c31;
}

struct InWallBox {
    const double *x, *y;
    double4 box;
    __device__ bool operator()(int32_t i) const {
        const double xi = x[i], yi = y[i];
        return xi >= box.x && xi <= box.y && yi >= box.z && yi <= box.w;
    }
};

// One wave per target inside the box, selected in index order (grid-stride,
// the count stays on the device).  The velocity terms go to list slot t of
// wst (u term at 2t, v term at 2t + 1); k_wall_add adds them.
template <typename T>
__global__ __launch_bounds__(256) void k_diffusion_wall(const double *__restrict__ x,
                                                        const double *__restrict__ y,
                                                        const double *__restrict__ d,
                                                        const int32_t *__restrict__ list,
                                                        const int32_t *__restrict__ cnt, PanelsSoA pan,
                                                        double nu, int64_t npts, double *__restrict__ wst) {
    const int lane = threadIdx.x & 63;
    const int64_t nw = (int64_t)gridDim.x * 4;
    const int64_t total = *cnt;
    for (int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); t < total; t += nw) {
        const int64_t i = list[t];
        double II0, II30, II31;
        wave_wall<T>(pan, lane, x[i], y[i], (T)d[i], npts, II0, II30, II31);
        if (lane == 0) {
            wst[2 * t] = nu * II30 / II0;
            wst[2 * t + 1] = nu * II31 / II0;
        }
    }
}

__global__ __launch_bounds__(256) void k_wall_add(const int32_t *__restrict__ list, const int32_t *__restrict__ cnt,
                                                  const double *__restrict__ wst, double *__restrict__ u,
                                                  double *__restrict__ v) {
    const int64_t total = *cnt;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = list[t];
        u[i] += wst[2 * t];
        v[i] += wst[2 * t + 1];
    }
}

int launch_wall_side(vx_case *c, int64_t n_tgt, hipStream_t st) {
    return launch_wall_side(c, host_cnt(n_tgt), st);
}

int launch_wall_side(vx_case *c, Cnt cn_tgt, hipStream_t st) {
    const int64_t n_tgt = cn_tgt.ub;  // bound: list sizes and the selection's grid
    if (n_tgt <= 0 || c->P <= 0) return 0;
    // buffers grow on the main stream: a reallocation synchronises it, and
    // with it every earlier side-stream pass (joined into it)
    if (c->wlist.ensure((size_t)n_tgt, c->stream) || c->wstage.ensure((size_t)(2 * n_tgt), c->stream) ||
        c->dcount.ensure(16, c->stream))
        return 1;
    int32_t *cnt = (int32_t *)(c->dcount.p + 5);
    const InWallBox pred{c->x.p, c->y.p, c->wall_box};
    if (c->wtmp.ensure((size_t)prim::tiles_of(n_tgt) + 1, c->stream)) return 1;
    VX_CHECK(prim::select_flags(prim::pred_flag(pred, cn_tgt), n_tgt, prim::WriteIndex{c->wlist.p}, c->wtmp.p, cnt,
                                st));
    if (c->prm.precision == VX_PREC_MIXED)
        hipLaunchKernelGGL(k_diffusion_wall<float>, dim3(1024), dim3(256), 0, st, c->x.p, c->y.p, c->d.p,
                           c->wlist.p, cnt, c->pan, c->prm.nu, c->prm.n_of_points, c->wstage.p);
    else
        hipLaunchKernelGGL(k_diffusion_wall<double>, dim3(1024), dim3(256), 0, st, c->x.p, c->y.p, c->d.p,
                           c->wlist.p, cnt, c->pan, c->prm.nu, c->prm.n_of_points, c->wstage.p);
    VX_CHECK(hipGetLastError());
    c->wall_pending = true;
    return 0;
}

int launch_wall_add(vx_case *c) {
    if (!c->wall_pending) return 0;
    c->wall_pending = false;
    hipLaunchKernelGGL(k_wall_add, dim3(64), dim3(256), 0, c->stream, c->wlist.p, (const int32_t *)(c->dcount.p + 5),
                       c->wstage.p, c->u.p, c->v.p);
    VX_CHECK(hipGetLastError());
    return 0;
}

int launch_diffusion_wall(vx_case *c, int64_t n_tgt) {
    return launch_wall_side(c, n_tgt, c->stream) || launch_wall_add(c);
}

int launch_diffusion_wall(vx_case *c, Cnt n_tgt) { return launch_wall_side(c, n_tgt, c->stream) || launch_wall_add(c); }

}  // namespace vx
