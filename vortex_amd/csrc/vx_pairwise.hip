// Pairwise O(N^2) stages: Biot-Savart velocity with the nearest-3 distance,
// vortex-vortex diffusive velocity, the birth right-hand side and the probe
// velocities.  Each is templated on the arithmetic type:
//   double  -- the reference formulas in FP64 (VX_PREC_F64);
//   float   -- FP32 pair arithmetic, FP32 partial sums over one 256-source
//              LDS tile, FP64 accumulation of the tile sums (VX_PREC_MIXED).
// Sources are staged through LDS in 256-wide tiles (one wave64-aligned load
// per thread, broadcast reads in the inner loop); each thread owns TPT
// targets spaced one block apart so that target loads and stores coalesce.
#include <algorithm>
#include <cfloat>
#include <mutex>
#include <string>


#include "vx_internal.h"
#include "vx_prim.h"

namespace vx {

template <typename T>
struct Src4 {
    T x, y, g, pad;
};

template <typename T>
struct Arith;

template <>
struct Arith<double> {
    static constexpr double kFar = 1e30;  // padding sources: squared distance stays finite
    __device__ static __forceinline__ double div(double a, double b) { return a / b; }
    __device__ static __forceinline__ double med3(double a, double b, double c) {
        return fmax(fmin(a, b), fmin(fmax(a, b), c));
    }
    __device__ static __forceinline__ double min_(double a, double b) { return fmin(a, b); }
    __device__ static __forceinline__ double max_(double a, double b) { return fmax(a, b); }
};

template <>
struct Arith<float> {
    static constexpr float kFar = 1e18f;
    __device__ static __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    __device__ static __forceinline__ float med3(float a, float b, float c) {
        return __builtin_amdgcn_fmed3f(a, b, c);
    }
    __device__ static __forceinline__ float min_(float a, float b) { return fminf(a, b); }
    __device__ static __forceinline__ float max_(float a, float b) { return fmaxf(a, b); }
};

constexpr int kTile = 256;

template <typename T>
__device__ __forceinline__ void stage_tile(Src4<T> *ls, const T *__restrict__ sx, const T *__restrict__ sy,
                                           const T *__restrict__ sg, int64_t base, int64_t n_src) {
    const int64_t j = base + threadIdx.x;
    Src4<T> s;
    if (j < n_src) {
        s.x = sx[j];
        s.y = sy[j];
        s.g = sg[j];
    } else {
        s.x = Arith<T>::kFar;
        s.y = Arith<T>::kFar;
        s.g = T(0);
    }
    s.pad = T(0);
    ls[threadIdx.x] = s;
}

// ---- A6: src/kernel.cu:875-983 -------------------------------------------
// u_i = V_inf + (1/2pi) sum_j g_j (-(y_i-y_j), x_i-x_j) / max(r_ij^2, eps^2), j != i
// d_i = mean of sqrt of the three smallest r_ij^2 over j != i (sentinel 1e5).
// The self pair contributes exactly 0 to the sums (dx = dy = 0); for d the
// four smallest r^2 including the self 0 are kept and the smallest one is
// dropped, which equals the three smallest over j != i.
template <typename T, int TPT>
__global__ __launch_bounds__(kTile) void k_biot_savart(
    const T *__restrict__ sx, const T *__restrict__ sy, const T *__restrict__ sg, Cnt cn_src,
    Cnt cn_tgt, T eps2, double vinf_x, double vinf_y, double *__restrict__ u,
    double *__restrict__ v, double *__restrict__ d) {
    __shared__ Src4<T> ls[kTile];
    using A = Arith<T>;
    const int64_t n_src = cn_src.get(), n_tgt = cn_tgt.get();
    const int64_t t0 = (int64_t)blockIdx.x * (kTile * TPT) + threadIdx.x;
    T ax[TPT], ay[TPT], m0[TPT], m1[TPT], m2[TPT], m3[TPT], su[TPT], sv[TPT];
    double au[TPT], av[TPT];
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int64_t i = t0 + (int64_t)k * kTile;
        ax[k] = i < n_tgt ? sx[i] : T(0);
        ay[k] = i < n_tgt ? sy[i] : T(0);
        m0[k] = m1[k] = m2[k] = m3[k] = T(1e5);
        su[k] = sv[k] = T(0);
        au[k] = av[k] = 0.0;
    }
    for (int64_t base = 0; base < n_src; base += kTile) {
        __syncthreads();
        stage_tile<T>(ls, sx, sy, sg, base, n_src);
        __syncthreads();
#pragma unroll 4
        for (int jj = 0; jj < kTile; ++jj) {
            const Src4<T> s = ls[jj];
#pragma unroll
            for (int k = 0; k < TPT; ++k) {
                const T dx = ax[k] - s.x;
                const T dy = ay[k] - s.y;
                const T r2 = dx * dx + dy * dy;
                m3[k] = A::med3(m2[k], m3[k], r2);
                m2[k] = A::med3(m1[k], m2[k], r2);
                m1[k] = A::med3(m0[k], m1[k], r2);
                m0[k] = A::min_(m0[k], r2);
                const T w = A::div(s.g, A::max_(r2, eps2));
                sv[k] += w * dx;
                su[k] -= w * dy;
            }
        }
        if constexpr (sizeof(T) == 4) {
#pragma unroll
            for (int k = 0; k < TPT; ++k) {
                au[k] += (double)su[k];
                av[k] += (double)sv[k];
                su[k] = sv[k] = T(0);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int64_t i = t0 + (int64_t)k * kTile;
        if (i >= n_tgt) continue;
        if constexpr (sizeof(T) == 8) {
            au[k] = (double)su[k];
            av[k] = (double)sv[k];
        }
        u[i] = au[k] / (2 * kPi) + vinf_x;
        v[i] = av[k] / (2 * kPi) + vinf_y;
        if (d) d[i] = (sqrt((double)m1[k]) + sqrt((double)m2[k]) + sqrt((double)m3[k])) / 3;
    }
}

// ---- A7: src/kernel.cu:985-1029 -------------------------------------------
// Over j with r_ij > 0.001 ve_size: I2 += -g_j (r_i - r_j)/r e^{-r/d_i},
// I1 += g_j e^{-r/d_i}; if |I1| > 1e-5: V_i += -nu I2 / (I1 d_i).
template <typename T, int TPT>
__global__ __launch_bounds__(kTile) void k_diffusion(
    const T *__restrict__ sx, const T *__restrict__ sy, const T *__restrict__ sg, Cnt cn_src,
    Cnt cn_tgt, const double *__restrict__ dd, double nu, double thr,
    double *__restrict__ u, double *__restrict__ v) {
    __shared__ Src4<T> ls[kTile];
    const int64_t n_src = cn_src.get(), n_tgt = cn_tgt.get();
    const int64_t t0 = (int64_t)blockIdx.x * (kTile * TPT) + threadIdx.x;
    T ax[TPT], ay[TPT], kd[TPT], s2x[TPT], s2y[TPT], s1[TPT];
    double a2x[TPT], a2y[TPT], a1[TPT];
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int64_t i = t0 + (int64_t)k * kTile;
        ax[k] = i < n_tgt ? sx[i] : T(0);
        ay[k] = i < n_tgt ? sy[i] : T(0);
        const double di = i < n_tgt ? dd[i] : 1.0;
        if constexpr (sizeof(T) == 4)
            kd[k] = (T)(-1.4426950408889634 / di);  // exp(-r/d) = exp2(r * kd)
        else
            kd[k] = (T)(1.0 / di);
        s2x[k] = s2y[k] = s1[k] = T(0);
        a2x[k] = a2y[k] = a1[k] = 0.0;
    }
    const T thr2 = (T)(thr * thr);
    for (int64_t base = 0; base < n_src; base += kTile) {
        __syncthreads();
        stage_tile<T>(ls, sx, sy, sg, base, n_src);
        __syncthreads();
#pragma unroll 4
        for (int jj = 0; jj < kTile; ++jj) {
            const Src4<T> s = ls[jj];
#pragma unroll
            for (int k = 0; k < TPT; ++k) {
                const T dx = ax[k] - s.x;
                const T dy = ay[k] - s.y;
                const T r2 = dx * dx + dy * dy;
                if constexpr (sizeof(T) == 4) {
                    const T r2c = r2 > thr2 ? r2 : T(1e30f);
                    const T rs = __builtin_amdgcn_rsqf(r2c);
                    const T dist = r2c * rs;
                    const T e = __builtin_amdgcn_exp2f(dist * kd[k]);
                    const T ge = s.g * e;
                    const T w = ge * rs;
                    s2x[k] += w * dx;
                    s2y[k] += w * dy;
                    s1[k] += ge;
                } else {
                    const T dist = sqrt(r2);
                    if (dist > (T)thr) {
                        const T w = s.g / dist;
                        const T e = exp(-dist * kd[k]);
                        s2x[k] += w * dx * e;
                        s2y[k] += w * dy * e;
                        s1[k] += s.g * e;
                    }
                }
            }
        }
        if constexpr (sizeof(T) == 4) {
#pragma unroll
            for (int k = 0; k < TPT; ++k) {
                a2x[k] += (double)s2x[k];
                a2y[k] += (double)s2y[k];
                a1[k] += (double)s1[k];
                s2x[k] = s2y[k] = s1[k] = T(0);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int64_t i = t0 + (int64_t)k * kTile;
        if (i >= n_tgt) continue;
        if constexpr (sizeof(T) == 8) {
            a2x[k] = s2x[k];
            a2y[k] = s2y[k];
            a1[k] = s1[k];
        }
        // I2 = -sum w (r_i - r_j) e, so -nu I2/(I1 d) = nu a2/(I1 d)
        if (fabs(a1[k]) > 1e-5) {
            const double di = dd[i];
            u[i] += nu * a2x[k] / (a1[k] * di);
            v[i] += nu * a2y[k] / (a1[k] * di);
        }
    }
}

// atan2 in FP32 for the panel right-hand side: octant reduction to
// t = min(|x|,|y|) / max(|x|,|y|) in [0, 1] and an odd polynomial of degree
// 15 for atan(t), fitted for this library by iteratively reweighted least
// squares on [0, 1] (maximum error 3.8e-8 before FP32 rounding, 1.2e-7 after).
// Signed zeros follow C's atan2 (x = -0 gives +-pi).
__device__ __forceinline__ float vx_atan2f(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float t = mn == 0.f ? 0.f : fminf(mn * __builtin_amdgcn_rcpf(mx), 1.f);
    const float s = t * t;
    float r = -0.004052795469760895f;
    r = __builtin_fmaf(r, s, 0.021855246275663376f);
    r = __builtin_fmaf(r, s, -0.05589873716235161f);
    r = __builtin_fmaf(r, s, 0.09640958160161972f);
    r = __builtin_fmaf(r, s, -0.13908010721206665f);
    r = __builtin_fmaf(r, s, 0.19946402311325073f);
    r = __builtin_fmaf(r, s, -0.3332984149456024f);
    r = __builtin_fmaf(r, s, 0.9999993443489075f);
    r *= t;
    if (ay > ax) r = 1.57079632679489662f - r;
    if (__builtin_signbit(x)) r = 3.14159265358979324f - r;
    return __builtin_copysignf(r, y);
}

// ---- A3: src/kernel.cu:779-847 --------------------------------------------
// For panel i with end points l, r, d = r - l:
//   y += g_j (alpha d + beta k x d), alpha = atan(s0.d/z) - atan(s1.d/z),
//   beta = 0.5 ln(|s1|^2/|s0|^2), s0 = l - r_j, s1 = r - r_j, z = d x s0.
// In FP32 alpha is evaluated as one atan2(-|d|^2 z, z^2 + (s0.d)(s1.d)),
// the same angle difference.  This is the panel average of the point
// vortices' conjugate velocity g/(2 pi i (z - z_j)), times 2 pi |d|^2.
//
// Sources far from the profile are summed as local expansions instead.
// The panels lie in the profile box, centre c, half-diagonal R_b, and in
// segments of at most 128 consecutive panels, each inside a circle of centre
// c_s and radius R_s.  For sources with |z_j - c| >= 2 R_b (global) and for
// z on a panel,
//   (u - i v)(z) = (i / 2 pi) sum_k S_k (z - c)^k,
//   S_k = sum_j g_j w_j^{k+1},  w_j = 1 / (z_j - c),
// truncated after kLoc terms, which leaves at most 2^{1-kLoc} of the far sum.
// It is shifted exactly to each segment centre (a polynomial of degree
// kLoc - 1 re-expanded: sum_{k>=m} S_k C(k, m) (c_s - c)^{k-m}).  Sources
// nearer than 2 R_b but at least 2 R_s from c_s add their own terms
// g_j w_{s,j}^{m+1} about c_s (the same bound).  The rest, nearer than 2 R_s,
// are summed pair by pair for that segment's panels.  The panel average of
// the segment's expansion is exact:
//   mean over [z_l, z_r] of (z - c_s)^k = Q_k / (k + 1),
//   Q_k = (a^{k+1} - b^{k+1}) / (a - b) = a Q_{k-1} + b^k,  a = z_r - c_s, b = z_l - c_s.
// The sources nearer than 2 R_b (with g != 0) are listed in index order.
constexpr int kLoc = 48;
constexpr int kRhsChunk = 512;

struct RhsNear {
    const double *x, *y, *g;
    double cx, cy, r2;
    __device__ bool operator()(int32_t j) const {
        const double dx = x[j] - cx, dy = y[j] - cy;
        return g[j] != 0 && dx * dx + dy * dy < r2;
    }
};

constexpr int kSegTile = 128;  // panels per segment, threads per pair work-group

// One work-group per (segment, chunk of kRhsChunk listed sources), a thread
// per panel of the segment.  Each tile of sources is staged with a flag for
// |z_j - c_s| < 2 R_s (FP64 positions, as the expansion kernel tests), the
// flagged ones compacted in LDS and summed pair by pair.
template <typename T>
__global__ __launch_bounds__(kSegTile) void k_right_part_partial(
    const T *__restrict__ sx, const T *__restrict__ sy, const T *__restrict__ sg, const double *__restrict__ x,
    const double *__restrict__ y, const int32_t *__restrict__ list, const int32_t *__restrict__ nlist,
    PanelsSoA pan, const RhsSeg *__restrict__ segs, double *__restrict__ part) {
    __shared__ Src4<T> ls[kSegTile];
    __shared__ int wcount[2];
    const RhsSeg S = segs[blockIdx.x];
    const int64_t c = blockIdx.y;
    const int64_t n_src = *nlist;
    const int64_t lo = c * kRhsChunk, hi = min(n_src, lo + kRhsChunk);
    if (lo >= n_src && lo > 0) return;  // a chunk k_right_part_final does not read
    const bool live = (int)threadIdx.x < S.count;
    const int64_t i = S.first + (live ? threadIdx.x : 0);
    // FP32: panel ends and sources relative to the segment centre, the panel
    // vector differenced in FP64 (absolute FP32 coordinates would leave the
    // short panels' vectors with relative errors of ulp(|z|) / length)
    constexpr bool F32 = sizeof(T) == 4;
    const double ox = F32 ? S.cx : 0.0, oy = F32 ? S.cy : 0.0;
    const T l0 = (T)(pan.lx[i] - ox), l1 = (T)(pan.ly[i] - oy);
    const T r0 = (T)(pan.rx[i] - ox), r1 = (T)(pan.ry[i] - oy);
    const T d0 = F32 ? (T)(pan.rx[i] - pan.lx[i]) : r0 - l0, d1 = F32 ? (T)(pan.ry[i] - pan.ly[i]) : r1 - l1;
    const T dd2 = d0 * d0 + d1 * d1;
    const double rn2 = 4.0 * S.r * S.r;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    T y0 = 0, y1 = 0;
    double a0 = 0.0, a1 = 0.0;
    for (int64_t base = lo; base < hi; base += kSegTile) {
        const int64_t j = base + threadIdx.x;
        bool near = false;
        Src4<T> q;
        if (j < hi) {
            const int32_t k = list[j];
            const double ex = x[k] - S.cx, ey = y[k] - S.cy;
            near = ex * ex + ey * ey < rn2;
            q.x = F32 ? (T)ex : sx[k];
            q.y = F32 ? (T)ey : sy[k];
            q.g = sg[k];
            q.pad = T(0);
        }
        const uint64_t m = __ballot(near);
        __syncthreads();  // the previous tile is read
        if (lane == 0) wcount[wv] = __popcll(m);
        __syncthreads();
        const int pos = (wv ? wcount[0] : 0) + __popcll(m & ((1ull << lane) - 1));
        if (near) ls[pos] = q;
        __syncthreads();
        const int cnt = wcount[0] + wcount[1];
        for (int jj = 0; jj < cnt; ++jj) {
            const Src4<T> s_ = ls[jj];
            const T s00 = l0 - s_.x, s01 = l1 - s_.y;
            const T s10 = r0 - s_.x, s11 = r1 - s_.y;
            const T z = d0 * s01 - d1 * s00;
            const T A = s00 * d0 + s01 * d1;
            const T B = s10 * d0 + s11 * d1;
            T alpha, beta;
            if constexpr (sizeof(T) == 4) {
                alpha = vx_atan2f(-dd2 * z, z * z + A * B);
                const T q0 = s00 * s00 + s01 * s01, q1 = s10 * s10 + s11 * s11;
                beta = 0.34657359027997264f * __builtin_amdgcn_logf(q1 * __builtin_amdgcn_rcpf(q0));
            } else {
                alpha = atan(A / z) - atan(B / z);
                beta = 0.5 * log((s10 * s10 + s11 * s11) / (s00 * s00 + s01 * s01));
            }
            y0 += s_.g * (alpha * d0 - beta * d1);
            y1 += s_.g * (alpha * d1 + beta * d0);
        }
        if constexpr (sizeof(T) == 4) {
            a0 += (double)y0;
            a1 += (double)y1;
            y0 = y1 = T(0);
        }
    }
    if constexpr (sizeof(T) == 8) {
        a0 = y0;
        a1 = y1;
    }
    if (live) {
        part[(c * 2 + 0) * pan.P + i] = a0;
        part[(c * 2 + 1) * pan.P + i] = a1;
    }
}

// Far sources: per work-group partial S_k (FP64), k < kLoc, grid-stride over
// the sources; spart[(blk * kLoc + k) * 2 + {0,1}].
__global__ __launch_bounds__(256) void k_rhs_far_partial(const double *__restrict__ x, const double *__restrict__ y,
                                                         const double *__restrict__ g, Cnt cn_src, double cx,
                                                         double cy, double r2, double *__restrict__ spart) {
    __shared__ double red[16][256];
    const int64_t n_src = cn_src.get();
    double sr[kLoc], si[kLoc];
#pragma unroll
    for (int k = 0; k < kLoc; ++k) sr[k] = si[k] = 0.0;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n_src; j += (int64_t)gridDim.x * blockDim.x) {
        const double gj = g[j];
        const double dx = x[j] - cx, dy = y[j] - cy;
        const double q2 = dx * dx + dy * dy;
        if (gj == 0 || !(q2 >= r2)) continue;
        const double wr = dx / q2, wi = -dy / q2;  // w = 1 / (z_j - c)
        double pr = gj * wr, pi = gj * wi;           // g w^{k+1}
#pragma unroll
        for (int k = 0; k < kLoc; ++k) {
            sr[k] += pr;
            si[k] += pi;
            const double npr = pr * wr - pi * wi;
            pi = pr * wi + pi * wr;
            pr = npr;
        }
    }
    // block sums, eight terms at a time, in a fixed order
#pragma unroll
    for (int k0 = 0; k0 < kLoc; k0 += 8) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            red[2 * q][threadIdx.x] = sr[k0 + q];
            red[2 * q + 1][threadIdx.x] = si[k0 + q];
        }
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if ((int)threadIdx.x < st)
                for (int q = 0; q < 16; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x < 16) spart[((int64_t)blockIdx.x * kLoc + k0) * 2 + threadIdx.x] = red[threadIdx.x][0];
    }
}

// S_k over the work-group partials: one work-group per real/imaginary part of
// each term, a strided pass per thread and a tree in LDS (fixed order).
__global__ __launch_bounds__(256) void k_rhs_far_final(const double *__restrict__ spart, int64_t nblk,
                                                       double *__restrict__ S) {
    __shared__ double r[256];
    const int t = blockIdx.x;
    double a = 0.0;
    for (int64_t b = threadIdx.x; b < nblk; b += 256) a += spart[b * 2 * kLoc + t];
    r[threadIdx.x] = a;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) r[threadIdx.x] += r[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) S[t] = r[0];
}

// Listed sources at least 2 R_s from a segment's centre: their terms
// g w^{m+1}, w = 1 / (z_j - c_s), for m < kLoc, per work-group (blockIdx.y
// of kSegBlocks for segment blockIdx.x, grid-stride over the list), summed in
// a fixed order; spart[((s * kSegBlocks + blk) * kLoc + m) * 2 + {0,1}].
constexpr int kSegBlocks = 32;

__global__ __launch_bounds__(256) void k_rhs_seg_partial(const double *__restrict__ x, const double *__restrict__ y,
                                                         const double *__restrict__ g,
                                                         const int32_t *__restrict__ list,
                                                         const int32_t *__restrict__ nlist,
                                                         const RhsSeg *__restrict__ segs,
                                                         double *__restrict__ spart) {
    __shared__ double red[16][256];
    const RhsSeg S = segs[blockIdx.x];
    const double rn2 = 4.0 * S.r * S.r;
    const int64_t n = *nlist;
    double sr[kLoc], si[kLoc];
#pragma unroll
    for (int k = 0; k < kLoc; ++k) sr[k] = si[k] = 0.0;
    for (int64_t t = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.y * blockDim.x) {
        const int32_t j = list[t];
        const double dx = x[j] - S.cx, dy = y[j] - S.cy;
        const double q2 = dx * dx + dy * dy;
        if (!(q2 >= rn2)) continue;
        const double wr = dx / q2, wi = -dy / q2;
        double pr = g[j] * wr, pi = g[j] * wi;
#pragma unroll
        for (int k = 0; k < kLoc; ++k) {
            sr[k] += pr;
            si[k] += pi;
            const double npr = pr * wr - pi * wi;
            pi = pr * wi + pi * wr;
            pr = npr;
        }
    }
    const int64_t blk = (int64_t)blockIdx.x * gridDim.y + blockIdx.y;
#pragma unroll
    for (int k0 = 0; k0 < kLoc; k0 += 8) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            red[2 * q][threadIdx.x] = sr[k0 + q];
            red[2 * q + 1][threadIdx.x] = si[k0 + q];
        }
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if ((int)threadIdx.x < st)
                for (int q = 0; q < 16; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x < 16) spart[(blk * kLoc + k0) * 2 + threadIdx.x] = red[threadIdx.x][0];
    }
}

// Each segment's coefficients about c_s: the global S shifted exactly,
//   sum_{k >= m} S_k C(k, m) (c_s - c)^{k - m},
// plus its own sources' partials (fixed order).  Thread m < kLoc.
__global__ void k_rhs_seg_final(const double *__restrict__ spart, const double *__restrict__ S, double cx,
                                double cy, const RhsSeg *__restrict__ segs, double *__restrict__ Ss) {
    const int m = threadIdx.x;
    if (m >= kLoc) return;
    const RhsSeg G = segs[blockIdx.x];
    const double er = G.cx - cx, ei = G.cy - cy;  // c_s - c
    double ar = 0.0, ai = 0.0;
    double bin = 1.0;              // C(k, m)
    double pr = 1.0, pi = 0.0;     // (c_s - c)^{k - m}
    for (int k = m; k < kLoc; ++k) {
        const double tr = pr * bin, ti = pi * bin;
        ar += S[2 * k] * tr - S[2 * k + 1] * ti;
        ai += S[2 * k] * ti + S[2 * k + 1] * tr;
        bin = bin * (double)(k + 1) / (double)(k + 1 - m);
        const double npr = pr * er - pi * ei;
        pi = pr * ei + pi * er;
        pr = npr;
    }
    for (int b = 0; b < kSegBlocks; ++b) {
        const double *q = spart + (((int64_t)blockIdx.x * kSegBlocks + b) * kLoc + m) * 2;
        ar += q[0];
        ai += q[1];
    }
    Ss[((int64_t)blockIdx.x * kLoc + m) * 2] = ar;
    Ss[((int64_t)blockIdx.x * kLoc + m) * 2 + 1] = ai;
}

// Sum of the chunk partials part[(c * 2 + q) * m + i] for one output i per
// work-group: a strided pass per thread, then a tree in LDS (fixed order).
__device__ __forceinline__ void chunk_sum2(const double *__restrict__ part, int64_t nchunk, int64_t m, int64_t i,
                                           double &a, double &b) {
    __shared__ double r[2][kTile];
    double x0 = 0.0, x1 = 0.0;
    for (int64_t c = threadIdx.x; c < nchunk; c += kTile) {
        x0 += part[(c * 2 + 0) * m + i];
        x1 += part[(c * 2 + 1) * m + i];
    }
    r[0][threadIdx.x] = x0;
    r[1][threadIdx.x] = x1;
    __syncthreads();
    for (int st = kTile / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            r[0][threadIdx.x] += r[0][threadIdx.x + st];
            r[1][threadIdx.x] += r[1][threadIdx.x + st];
        }
        __syncthreads();
    }
    a = r[0][0];
    b = r[1][0];
}

// One work-group per panel: the near chunk partials, plus the panel average
// of its segment's expansion (thread 0).  Only the chunks of the near list's
// length hold partials (the grid is sized for its bound); the others would
// add exact zeros.
__global__ __launch_bounds__(kTile) void k_right_part_final(const double *__restrict__ part, int64_t nchunk,
                                                            const int32_t *__restrict__ nlist, PanelsSoA pan,
                                                            const double *__restrict__ Ss,
                                                            const RhsSeg *__restrict__ segs,
                                                            const int32_t *__restrict__ seg_of, double vinf_x,
                                                            double vinf_y, double *__restrict__ R) {
    const int64_t i = blockIdx.x;
    const int64_t used = ((int64_t)*nlist + kRhsChunk - 1) / kRhsChunk;
    double y0, y1;
    chunk_sum2(part, used < nchunk ? used : nchunk, pan.P, i, y0, y1);
    if (threadIdx.x != 0) return;
    const int32_t sg = seg_of[i];
    const double cx = segs[sg].cx, cy = segs[sg].cy;
    const double *S = Ss + (int64_t)sg * 2 * kLoc;
    const double d0 = pan.rx[i] - pan.lx[i], d1 = pan.ry[i] - pan.ly[i];
    const double mnog = 1 / (2 * kPi * (d0 * d0 + d1 * d1));
    // W = sum_k (i S_k / 2 pi) Q_k / (k + 1), u = Re W, v = -Im W
    const double ar = pan.rx[i] - cx, ai = pan.ry[i] - cy;  // a = z_r - c_s
    const double br = pan.lx[i] - cx, bi = pan.ly[i] - cy;  // b = z_l - c_s
    double qr = 1.0, qi = 0.0;                               // Q_0
    double bkr = 1.0, bki = 0.0;                             // b^k
    double wr = 0.0, wi = 0.0;
#pragma unroll
    for (int k = 0; k < kLoc; ++k) {
        if (k > 0) {
            const double nbr = bkr * br - bki * bi;
            bki = bkr * bi + bki * br;
            bkr = nbr;
            const double nqr = ar * qr - ai * qi + bkr;
            qi = ar * qi + ai * qr + bki;
            qr = nqr;
        }
        const double sr = S[2 * k], si = S[2 * k + 1];
        const double inv = 1.0 / (double)(k + 1);
        wr += (sr * qr - si * qi) * inv;
        wi += (sr * qi + si * qr) * inv;
    }
    // times i / (2 pi)
    const double uf = -wi / (2 * kPi), vf = -(wr / (2 * kPi));
    R[i] = -((mnog * y0 + uf + vinf_x) * pan.tx[i] + (mnog * y1 + vf + vinf_y) * pan.ty[i]);
}

// ---- A19: src/kernel.cu:1434-1475, split over source chunks ---------------
__global__ __launch_bounds__(kTile) void k_probe_partial(const double *__restrict__ sx, const double *__restrict__ sy,
                                                         const double *__restrict__ sg, Cnt cn_src, int64_t chunk,
                                                         const double *__restrict__ px, const double *__restrict__ py,
                                                         int64_t m, double eps2, double *__restrict__ part) {
    __shared__ Src4<double> ls[kTile];
    using A = Arith<double>;
    const int64_t n_src = cn_src.get();
    const int64_t i = (int64_t)blockIdx.x * kTile + threadIdx.x;
    const int64_t c = blockIdx.y;
    const int64_t lo = c * chunk, hi = min(n_src, lo + chunk);
    const double a0 = i < m ? px[i] : 0.0, a1 = i < m ? py[i] : 0.0;
    double su = 0.0, sv = 0.0;
    for (int64_t base = lo; base < hi; base += kTile) {
        __syncthreads();
        const int64_t j = base + threadIdx.x;
        Src4<double> s;
        s.x = j < hi ? sx[j] : A::kFar;
        s.y = j < hi ? sy[j] : A::kFar;
        s.g = j < hi ? sg[j] : 0.0;
        s.pad = 0.0;
        ls[threadIdx.x] = s;
        __syncthreads();
        const int cnt = (int)min((int64_t)kTile, hi - base);
        for (int jj = 0; jj < cnt; ++jj) {
            const Src4<double> q = ls[jj];
            const double dx = a0 - q.x, dy = a1 - q.y;
            const double w = A::div(q.g, A::max_(dx * dx + dy * dy, eps2));
            sv += w * dx;
            su -= w * dy;
        }
    }
    if (i < m) {
        part[(c * 2 + 0) * m + i] = su;
        part[(c * 2 + 1) * m + i] = sv;
    }
}

// Mixed precision: each pair's separation is taken in FP64 and rounded once,
// so a probe next to a source keeps FP32's relative accuracy in the
// separation (FP32 positions would lose it in the coordinates' magnitude:
// 1e-7 of |x| against separations down to the core radius); kernel and sums
// in FP32 per tile, FP64 across tiles.
__global__ __launch_bounds__(kTile) void k_probe_partial_mixed(
    const double *__restrict__ sx, const double *__restrict__ sy, const double *__restrict__ sg, Cnt cn_src,
    int64_t chunk, const double *__restrict__ px, const double *__restrict__ py, int64_t m, float eps2,
    double *__restrict__ part) {
    __shared__ double lx[kTile], ly[kTile];
    __shared__ float lg[kTile];
    using A = Arith<float>;
    const int64_t n_src = cn_src.get();
    const int64_t i = (int64_t)blockIdx.x * kTile + threadIdx.x;
    const int64_t c = blockIdx.y;
    const int64_t lo = c * chunk, hi = min(n_src, lo + chunk);
    const double a0 = i < m ? px[i] : 0.0, a1 = i < m ? py[i] : 0.0;
    double au = 0.0, av = 0.0;
    for (int64_t base = lo; base < hi; base += kTile) {
        __syncthreads();
        const int64_t j = base + threadIdx.x;
        lx[threadIdx.x] = j < hi ? sx[j] : 0.0;
        ly[threadIdx.x] = j < hi ? sy[j] : 0.0;
        lg[threadIdx.x] = j < hi ? (float)sg[j] : 0.f;
        __syncthreads();
        const int cnt = (int)min((int64_t)kTile, hi - base);
        float su = 0.f, sv = 0.f;
        for (int jj = 0; jj < cnt; ++jj) {
            const float dx = (float)(a0 - lx[jj]), dy = (float)(a1 - ly[jj]);
            const float w = A::div(lg[jj], A::max_(dx * dx + dy * dy, eps2));
            sv += w * dx;
            su -= w * dy;
        }
        au += (double)su;
        av += (double)sv;
    }
    if (i < m) {
        part[(c * 2 + 0) * m + i] = au;
        part[(c * 2 + 1) * m + i] = av;
    }
}

// One work-group per probe.
__global__ __launch_bounds__(kTile) void k_probe_final(const double *__restrict__ part, int64_t nchunk, int64_t m,
                                                       double vinf_x, double vinf_y, double *__restrict__ u,
                                                       double *__restrict__ v) {
    const int64_t i = blockIdx.x;
    double a, b;
    chunk_sum2(part, nchunk, m, i, a, b);
    if (threadIdx.x != 0) return;
    u[i] = a / (2 * kPi) + vinf_x;
    v[i] = b / (2 * kPi) + vinf_y;
}

// ---------------------------------------------------------------------------
static bool mixed(const vx_case *c) { return c->prm.precision == VX_PREC_MIXED; }

// Split [0, n_src) into chunks of whole tiles so that tgt_blocks * chunks
// gives a few thousand work-groups.
static void split(int64_t n_src, int64_t tgt_blocks, int64_t *chunk, int64_t *nchunk) {
    const int64_t tiles = (n_src + kTile - 1) / kTile;
    int64_t want = (4096 + tgt_blocks - 1) / tgt_blocks;
    if (want < 1) want = 1;
    if (want > tiles) want = tiles > 0 ? tiles : 1;
    const int64_t tiles_per = (tiles + want - 1) / want;
    *chunk = (tiles_per > 0 ? tiles_per : 1) * kTile;
    *nchunk = n_src > 0 ? (n_src + *chunk - 1) / *chunk : 1;
}

int launch_biot_savart(vx_case *c, int64_t n_tgt, int64_t n_src, double vinf_x, double vinf_y) {
    return launch_biot_savart(c, host_cnt(n_tgt), host_cnt(n_src), vinf_x, vinf_y);
}

// Targets per thread of the direct kernels: the case's tuning (see
// vx_case_autotune) or the default, 4 in FP32 and 2 in FP64.  Every choice
// sums each target's sources in the same order, so the results do not depend
// on it.
// Default targets per thread of the direct kernels: 4 (FP32) or 2 (FP64),
// fewer while the grid would give the 256 CUs less than two work-groups each
// (the sums, and so the results, do not depend on it).
int direct_tpt_default(bool f32, int64_t n_tgt) {
    int tpt = f32 ? 4 : 2;
    while (tpt > 1 && (n_tgt + kTile * tpt - 1) / (kTile * tpt) < 512) tpt >>= 1;
    return tpt;
}

static int tpt_of(int want, bool f32, int64_t n_tgt) {
    if (want == 1 || want == 2 || want == 4 || want == 8) return want;
    return direct_tpt_default(f32, n_tgt);
}

template <typename T, int TPT>
static void bs_launch(vx_case *c, const T *sx, const T *sy, const T *sg, Cnt n_src, Cnt n_tgt, T eps2, double vinf_x,
                      double vinf_y) {
    const unsigned nb = (unsigned)((n_tgt.ub + kTile * TPT - 1) / (kTile * TPT));
    hipLaunchKernelGGL((k_biot_savart<T, TPT>), dim3(nb), dim3(kTile), 0, c->stream, sx, sy, sg, n_src, n_tgt, eps2,
                       vinf_x, vinf_y, c->u.p, c->v.p, c->d.p);
}

template <typename T>
static void bs_dispatch(vx_case *c, int tpt, const T *sx, const T *sy, const T *sg, Cnt n_src, Cnt n_tgt, T eps2,
                        double vinf_x, double vinf_y) {
    switch (tpt) {
        case 1: bs_launch<T, 1>(c, sx, sy, sg, n_src, n_tgt, eps2, vinf_x, vinf_y); break;
        case 2: bs_launch<T, 2>(c, sx, sy, sg, n_src, n_tgt, eps2, vinf_x, vinf_y); break;
        case 4: bs_launch<T, 4>(c, sx, sy, sg, n_src, n_tgt, eps2, vinf_x, vinf_y); break;
        default: bs_launch<T, 8>(c, sx, sy, sg, n_src, n_tgt, eps2, vinf_x, vinf_y); break;
    }
}

int launch_biot_savart(vx_case *c, Cnt n_tgt, Cnt n_src, double vinf_x, double vinf_y) {
    if (n_tgt.ub <= 0) return 0;
    tree_d_changed(c);  // d from direct sums: a tree's sorted copy no longer matches
    const double eps2 = c->prm.ve_size * c->prm.ve_size;
    if (mixed(c)) {
        if (launch_pack_f32(c, n_src)) return 1;
        bs_dispatch<float>(c, tpt_of(c->tune.tpt_bs, true, n_tgt.ub), c->xf.p, c->yf.p, c->gf.p, n_src, n_tgt,
                           (float)eps2, vinf_x, vinf_y);
    } else {
        bs_dispatch<double>(c, tpt_of(c->tune.tpt_bs, false, n_tgt.ub), c->x.p, c->y.p, c->g.p, n_src, n_tgt, eps2,
                            vinf_x, vinf_y);
    }
    VX_CHECK(hipGetLastError());
    return 0;
}

int launch_diffusion(vx_case *c, int64_t n_tgt, int64_t n_src) {
    return launch_diffusion(c, host_cnt(n_tgt), host_cnt(n_src));
}

template <typename T, int TPT>
static void diff_launch(vx_case *c, const T *sx, const T *sy, const T *sg, Cnt n_src, Cnt n_tgt, double thr) {
    const unsigned nb = (unsigned)((n_tgt.ub + kTile * TPT - 1) / (kTile * TPT));
    hipLaunchKernelGGL((k_diffusion<T, TPT>), dim3(nb), dim3(kTile), 0, c->stream, sx, sy, sg, n_src, n_tgt, c->d.p,
                       c->prm.nu, thr, c->u.p, c->v.p);
}

template <typename T>
static void diff_dispatch(vx_case *c, int tpt, const T *sx, const T *sy, const T *sg, Cnt n_src, Cnt n_tgt,
                          double thr) {
    switch (tpt) {
        case 1: diff_launch<T, 1>(c, sx, sy, sg, n_src, n_tgt, thr); break;
        case 2: diff_launch<T, 2>(c, sx, sy, sg, n_src, n_tgt, thr); break;
        case 4: diff_launch<T, 4>(c, sx, sy, sg, n_src, n_tgt, thr); break;
        default: diff_launch<T, 8>(c, sx, sy, sg, n_src, n_tgt, thr); break;
    }
}

int launch_diffusion(vx_case *c, Cnt n_tgt, Cnt n_src) {
    if (n_tgt.ub <= 0) return 0;
    const double thr = 0.001 * c->prm.ve_size;
    if (mixed(c)) {
        if (launch_pack_f32(c, n_src)) return 1;
        diff_dispatch<float>(c, tpt_of(c->tune.tpt_diff, true, n_tgt.ub), c->xf.p, c->yf.p, c->gf.p, n_src, n_tgt,
                             thr);
    } else {
        diff_dispatch<double>(c, tpt_of(c->tune.tpt_diff, false, n_tgt.ub), c->x.p, c->y.p, c->g.p, n_src, n_tgt,
                              thr);
    }
    VX_CHECK(hipGetLastError());
    return 0;
}

int launch_right_part(vx_case *c, int64_t n_src, double vinf_x, double vinf_y) {
    return launch_right_part(c, host_cnt(n_src), vinf_x, vinf_y);
}

int launch_right_part(vx_case *c, Cnt cn_src, double vinf_x, double vinf_y) {
    const int64_t P = c->P, K = c->nseg;
    hipStream_t st = c->stream;
    const int64_t n_src = cn_src.ub;  // bound: grids, list and partial sizes
    const int64_t nchunk = n_src > 0 ? (n_src + kRhsChunk - 1) / kRhsChunk : 1;
    constexpr int64_t kFarBlocks = 512;
    const size_t ns_far = (size_t)(2 * kLoc * kFarBlocks), ns_glob = 2 * kLoc,
                 ns_seg = (size_t)(2 * kLoc * kSegBlocks * K), ns_segS = (size_t)(2 * kLoc * K);
    if (c->partial.ensure((size_t)(2 * nchunk * P), st) || c->rp.ensure((size_t)P, st) ||
        c->lists.ensure((size_t)(n_src > 0 ? n_src : 1), st) || c->dcount.ensure(16, st) ||
        c->rhs_s.ensure(ns_far + ns_glob + ns_seg + ns_segS, st))
        return 1;
    const vx_params &p = c->prm;
    const double cx = 0.5 * (p.px_min + p.px_max), cy = 0.5 * (p.py_min + p.py_max);
    const double wx = p.px_max - p.px_min, wy = p.py_max - p.py_min;
    const double rnear2 = wx * wx + wy * wy;  // (2 R_b)^2
    int32_t *nlist = (int32_t *)(c->dcount.p + 7);
    double *S = c->rhs_s.p + ns_far, *spart = S + ns_glob, *Ss = spart + ns_seg;
    // sources nearer than 2 R_b, in index order
    {
        const RhsNear pred{c->x.p, c->y.p, c->g.p, cx, cy, rnear2};
        if (c->ptmp.ensure((size_t)prim::tiles_of(n_src) + 1, st)) return 1;
        VX_CHECK(prim::select_flags(prim::pred_flag(pred, cn_src), n_src, prim::WriteIndex{c->lists.p}, c->ptmp.p,
                                    nlist, st));
    }
    // pairs within 2 R_s of each segment
    if (mixed(c)) {
        if (launch_pack_f32(c, cn_src)) return 1;
        hipLaunchKernelGGL(k_right_part_partial<float>, dim3((unsigned)K, (unsigned)nchunk), dim3(kSegTile), 0, st,
                           c->xf.p, c->yf.p, c->gf.p, c->x.p, c->y.p, c->lists.p, nlist, c->pan, c->segs,
                           c->partial.p);
    } else {
        hipLaunchKernelGGL(k_right_part_partial<double>, dim3((unsigned)K, (unsigned)nchunk), dim3(kSegTile), 0, st,
                           c->x.p, c->y.p, c->g.p, c->x.p, c->y.p, c->lists.p, nlist, c->pan, c->segs,
                           c->partial.p);
    }
    // the global expansion (sources beyond 2 R_b), the segments' own
    // sources, and each segment's total about c_s
    hipLaunchKernelGGL(k_rhs_far_partial, dim3((unsigned)kFarBlocks), dim3(256), 0, st, c->x.p, c->y.p, c->g.p, cn_src,
                       cx, cy, rnear2, c->rhs_s.p);
    hipLaunchKernelGGL(k_rhs_far_final, dim3(2 * kLoc), dim3(256), 0, st, c->rhs_s.p, kFarBlocks, S);
    hipLaunchKernelGGL(k_rhs_seg_partial, dim3((unsigned)K, (unsigned)kSegBlocks), dim3(256), 0, st, c->x.p, c->y.p,
                       c->g.p, c->lists.p, nlist, c->segs, spart);
    hipLaunchKernelGGL(k_rhs_seg_final, dim3((unsigned)K), dim3(64), 0, st, spart, S, cx, cy, c->segs, Ss);
    VX_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_right_part_final, dim3((unsigned)P), dim3(kTile), 0, st, c->partial.p, nchunk, nlist, c->pan,
                       Ss, c->segs, c->seg_of, vinf_x, vinf_y, c->rp.p);
    VX_CHECK(hipGetLastError());
    return 0;
}

// Probe chunking: at most max_part doubles of split-K partials (2 per probe
// and chunk); max_part <= 0 lets the split choose.
void probe_split(int64_t m, int64_t n_src, int64_t max_part, int64_t *chunk, int64_t *nchunk) {
    const int64_t tb = (m + kTile - 1) / kTile;
    split(n_src, tb, chunk, nchunk);
    if (max_part > 0 && 2 * *nchunk * m > max_part) {
        const int64_t want = max_part / (2 * m) > 1 ? max_part / (2 * m) : 1;
        const int64_t tiles = (n_src + kTile - 1) / kTile;
        const int64_t tiles_per = (tiles + want - 1) / want;
        *chunk = (tiles_per > 0 ? tiles_per : 1) * kTile;
        *nchunk = n_src > 0 ? (n_src + *chunk - 1) / *chunk : 1;
    }
}

int launch_probe_to(vx_case *c, const double *px, const double *py, int64_t m, int64_t n_src, double vinf_x,
                    double vinf_y, double *pu, double *pv, double *part, int64_t max_part) {
    return launch_probe_to(c, px, py, m, host_cnt(n_src), vinf_x, vinf_y, pu, pv, part, max_part);
}

int launch_probe_to(vx_case *c, const double *px, const double *py, int64_t m, Cnt cn_src, double vinf_x,
                    double vinf_y, double *pu, double *pv, double *part, int64_t max_part) {
    if (m <= 0) return 0;
    const int64_t tb = (m + kTile - 1) / kTile;
    const int64_t n_src = cn_src.ub;  // the chunking follows the bound
    int64_t chunk, nchunk;
    probe_split(m, n_src, max_part, &chunk, &nchunk);
    const double eps2 = c->prm.ve_size * c->prm.ve_size;
    const dim3 grid((unsigned)tb, (unsigned)nchunk);
    if (mixed(c)) {
        hipLaunchKernelGGL(k_probe_partial_mixed, grid, dim3(kTile), 0, c->stream, c->x.p, c->y.p, c->g.p, cn_src,
                           chunk, px, py, m, (float)eps2, part);
    } else {
        hipLaunchKernelGGL(k_probe_partial, grid, dim3(kTile), 0, c->stream, c->x.p, c->y.p,
                           c->g.p, cn_src, chunk, px, py, m, eps2, part);
    }
    VX_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_probe_final, dim3((unsigned)m), dim3(kTile), 0, c->stream, part, nchunk, m, vinf_x, vinf_y,
                       pu, pv);
    VX_CHECK(hipGetLastError());
    return 0;
}

int launch_probe(vx_case *c, int64_t m, int64_t n_src, double vinf_x, double vinf_y) {
    if (m <= 0) return 0;
    int64_t chunk, nchunk;
    probe_split(m, n_src, 0, &chunk, &nchunk);
    if (c->partial.ensure((size_t)(2 * nchunk * m), c->stream)) return 1;
    return launch_probe_to(c, c->px.p, c->py.p, m, n_src, vinf_x, vinf_y, c->pu.p, c->pv.p, c->partial.p, 0);
}

// The probes of vx_case_set_probes, sources [0, ceil32(n)) of the current
// state, into c->prb_u / c->prb_v (own partials, so the step's split-K
// buffers are left alone).
int launch_step_probes(vx_case *c, double vinf_x, double vinf_y) {
    const int64_t m = c->prb_m;
    if (m <= 0) return 0;
    const Cnt s = cnt_s(c);
    int64_t chunk, nchunk;
    probe_split(m, s.ub, 0, &chunk, &nchunk);
    if (c->prb_part.ensure((size_t)(2 * nchunk * m), c->stream)) return 1;
    return launch_probe_to(c, c->prb_x.p, c->prb_y.p, m, s, vinf_x, vinf_y, c->prb_u.p, c->prb_v.p, c->prb_part.p,
                           0);
}

// ---- velocity probes without allocation ---------------------------------------
// vx_velocity_probe_noalloc keeps its split-K partials in this file-scope
// device array (one copy per GPU, held until the process ends) instead of
// the case's growable buffer.  Host code never uses the array's name as an
// address: its device address comes from hipGetSymbolAddress, once per
// device, and is kept in scratch_ptr.
constexpr int64_t kProbeScratch = int64_t(1) << 21;  // doubles (16 MiB)
static __device__ __align__(256) double scratch[kProbeScratch];

static std::mutex scratch_mu;
static double *scratch_ptr[256] = {};

static int probe_scratch(int dev, double **out) {
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (dev < 0 || dev >= 256) {
        set_error("bad argument: device index");
        return 2;
    }
    if (!scratch_ptr[dev]) {
        int cur = -1;
        VX_CHECK(hipGetDevice(&cur));
        if (cur != dev) VX_CHECK(hipSetDevice(dev));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        VX_CHECK(e);
        scratch_ptr[dev] = (double *)p;
    }
    *out = scratch_ptr[dev];
    return 0;
}

int launch_probe_noalloc(vx_case *c, const double *px, const double *py, int64_t m, double vinf_x, double vinf_y,
                         double *pu, double *pv) {
    if (m <= 0) return 0;
    if (2 * m > kProbeScratch) {
        set_error("bad argument: vx_velocity_probe_noalloc takes at most " + std::to_string(kProbeScratch / 2) +
                  " probes");
        return 2;
    }
    double *part = nullptr;
    if (const int rc = probe_scratch(c->device, &part)) return rc;
    const int64_t n_src = std::min<int64_t>((c->n + 31) / 32 * 32, c->cap);
    return launch_probe_to(c, px, py, m, n_src, vinf_x, vinf_y, pu, pv, part, kProbeScratch);
}

}  // namespace vx
