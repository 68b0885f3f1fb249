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

template <typename T>
__global__ __launch_bounds__(256) void k_diffusion_wall(const double *__restrict__ x,
                                                        const double *__restrict__ y,
                                                        const double *__restrict__ d, int64_t n_tgt,
                                                        PanelsSoA pan, double nu, int64_t npts,
                                                        double4 box, double *__restrict__ u,
                                                        double *__restrict__ v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_tgt) return;
    const double xi = x[i], yi = y[i];
    if (!(xi >= box.x && xi <= box.y && yi >= box.z && yi <= box.w)) return;
    const T a0 = (T)xi, a1 = (T)yi, dd = (T)d[i];
    T II0 = T(2 * kPi) * dd * dd;
    T II30 = 0, II31 = 0;
    for (int64_t f = 0; f < pan.P; ++f) {
        const T dL = (T)pan.len[f];
        const T q0 = a0 - (T)pan.bx[f], q1 = a1 - (T)pan.by[f];
        const T r = sqrt(q0 * q0 + q1 * q1);
        if ((r < T(5) * dL) && (r > T(0.1) * dL)) {
            T R0, R30, R31;
            quad_i0_i3<T>((T)pan.lx[f], (T)pan.ly[f], (T)pan.rx[f], (T)pan.ry[f], -(T)pan.nx[f], -(T)pan.ny[f], a0,
                          a1, dL, dd, npts, R0, R30, R31);
            II0 += (-dd) * R0;
            II30 -= R30;
            II31 -= R31;
        } else if (r <= T(0.1) * dL) {
            const T ex = T(1) - exp(-dL / (T(2) * dd));
            II0 = T(kPi) * dd * dd;
            II30 = T(2) * (T)pan.nx[f] * dd * ex;
            II31 = T(2) * (T)pan.ny[f] * dd * ex;
            break;
        }
    }
    u[i] += nu * (double)II30 / (double)II0;
    v[i] += nu * (double)II31 / (double)II0;
}

int launch_diffusion_wall(vx_case *c, int64_t n_tgt) {
    if (n_tgt <= 0 || c->P <= 0) return 0;
    const unsigned nb = (unsigned)((n_tgt + 255) / 256);
    if (c->prm.precision == VX_PREC_MIXED)
        hipLaunchKernelGGL(k_diffusion_wall<float>, dim3(nb), dim3(256), 0, c->stream, c->x.p, c->y.p, c->d.p,
                           n_tgt, c->pan, c->prm.nu, c->prm.n_of_points, c->wall_box, c->u.p, c->v.p);
    else
        hipLaunchKernelGGL(k_diffusion_wall<double>, dim3(nb), dim3(256), 0, c->stream, c->x.p, c->y.p, c->d.p,
                           n_tgt, c->pan, c->prm.nu, c->prm.n_of_points, c->wall_box, c->u.p, c->v.p);
    VX_CHECK(hipGetLastError());
    return 0;
}

}  // namespace vx
