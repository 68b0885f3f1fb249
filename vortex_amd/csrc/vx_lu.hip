// Panel system in mixed precision: FP32 LU of the influence matrix with
// partial pivoting (blocked, right-looking; trailing updates as FP32 matrix
// products on MFMA), the triangular factors inverted blockwise (again MFMA
// products), and per birth an FP64-residual iterative refinement:
//
//     x = 0, r = b;  repeat: x += U^-1 L^-1 P r  (FP32 products),
//                            r = b - A x         (FP64, A = the FP64 matrix)
//     until ||r|| <= ||x|| ||A|| eps64 sqrt(n)   (LAPACK dsgesv's test)
//
// The reference inverts the matrix once in FP64 by Gauss-Jordan
// (src/unit_1.cu:115-254) and multiplies by the inverse at every birth
// (src/kernel.cu:849-873); the solution of the same system is what both
// deliver.  Everything stays on the device: the refinement's kernels read a
// convergence flag and return at once after the test passes, so a fixed
// number of them is queued per birth and no host wait is needed.
//
// Layout: the order n = P + 1 is padded to np, a multiple of 64, with an
// identity block (decoupled, never pivoted).  FP32 matrices are np x np
// row-major: F holds L (strict lower, unit diagonal) and U (upper) of P A,
// W the two inverses (strict lower: L^-1 below its unit diagonal; upper
// with diagonal: U^-1).
#include <cmath>

#include "vx_internal.h"
#include "vx_prim.h"

namespace vx {

constexpr int kLuB = 64;       // panel width, block size of the inverses, product tile
constexpr int kLuMaxIt = 6;    // refinement corrections queued per solve

typedef float floatx4 __attribute__((ext_vector_type(4)));

struct LuBufs {
    int64_t n = 0, np = 0;
    DBuf<float> F, W, Wl, Wu, T;
    DBuf<double> A;               // the FP64 matrix (n x n), for the residuals
    DBuf<double> x;               // solution (np)
    DBuf<float> rp, t;            // permuted FP32 residual, L^-1 P r (np)
    DBuf<int32_t> ipiv, perm, pinv;
    DBuf<unsigned long long> st;  // 0 done, 1 corrections, 2 |r| max bits, 3 |x| max bits, 4 singular,
                                  // 5 ||A||_inf bits
    DBuf<double> nrm;             // 1: last ||r|| / (||A|| ||x||)
    int32_t last_iters = 0;
};

// ---- FP32 matrix product on MFMA ---------------------------------------------
// C = alpha A B + beta Cin (Cin may be C) for row-major blocks with M, N
// multiples of 64 and K a multiple of 16.  One work-group of 4 waves per
// 64 x 64 tile of C, each wave a 32 x 32 quarter as 2 x 2 products of
// v_mfma_f32_16x16x4_f32 (A: lane = i + 16 k; B: lane = j + 16 k; C: lane =
// j + 16 (i / 4), register i % 4).  TRI = 1: A is lower triangular in its
// absolute indices (a_row0 + i, a_col0 + k) and the other triangle holds
// zeros, TRI = 2: upper; the K loop then skips whole zero steps.
template <int TRI>
__global__ __launch_bounds__(256) void k_gemm_f32(const float *__restrict__ A, int64_t lda, const float *__restrict__ B,
                                                  int64_t ldb, float *C, int64_t ldc, const float *Cin,
                                                  int64_t ldcin, int64_t K, float alpha, float beta, int64_t a_row0,
                                                  int64_t a_col0) {
    __shared__ float As[64][17];
    __shared__ float Bs[16][65];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t bm = (int64_t)blockIdx.y * 64, bn = (int64_t)blockIdx.x * 64;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    floatx4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
    int64_t k_lo = 0, k_hi = K;
    if (TRI == 1) {
        const int64_t lim = a_row0 + bm + 64 - a_col0;  // local k beyond the tile's last row are zero
        k_hi = lim < K ? (lim > 0 ? lim : 0) : K;
        k_hi = (k_hi + 15) / 16 * 16;
    } else if (TRI == 2) {
        const int64_t lim = a_row0 + bm - a_col0;  // local k before the tile's first row are zero
        k_lo = lim > 0 ? (lim < K ? lim : K) : 0;
        k_lo = k_lo / 16 * 16;
    }
    const int ar = tid >> 2, ac = (tid & 3) * 4;     // A tile: 64 rows x 16
    const int br = tid >> 4, bc = (tid & 15) * 4;    // B tile: 16 rows x 64
    for (int64_t k0 = k_lo; k0 < k_hi; k0 += 16) {
        const float *ap = A + (bm + ar) * lda + k0 + ac;
        const float *bp = B + (k0 + br) * ldb + bn + bc;
        const float4 av = *reinterpret_cast<const float4 *>(ap);
        const float4 bv = *reinterpret_cast<const float4 *>(bp);
        __syncthreads();
        As[ar][ac + 0] = av.x; As[ar][ac + 1] = av.y; As[ar][ac + 2] = av.z; As[ar][ac + 3] = av.w;
        Bs[br][bc + 0] = bv.x; Bs[br][bc + 1] = bv.y; Bs[br][bc + 2] = bv.z; Bs[br][bc + 3] = bv.w;
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; kk += 4) {
            const int kq = kk + (lane >> 4), q = lane & 15;
            const float a0 = As[wm + q][kq], a1 = As[wm + 16 + q][kq];
            const float b0 = Bs[kq][wn + q], b1 = Bs[kq][wn + 16 + q];
            acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
        }
    }
    for (int si = 0; si < 2; ++si)
        for (int sj = 0; sj < 2; ++sj)
            for (int r = 0; r < 4; ++r) {
                const int64_t i = bm + wm + si * 16 + 4 * (lane >> 4) + r;
                const int64_t j = bn + wn + sj * 16 + (lane & 15);
                const float old = beta != 0.f ? Cin[i * ldcin + j] : 0.f;
                C[i * ldc + j] = alpha * acc[si][sj][r] + beta * old;
            }
}

template <int TRI>
static int gemm(const float *A, int64_t lda, const float *B, int64_t ldb, float *C, int64_t ldc, const float *Cin,
                int64_t ldcin, int64_t M, int64_t N, int64_t K, float alpha, float beta, int64_t a_row0,
                int64_t a_col0, hipStream_t st) {
    if (M <= 0 || N <= 0 || K <= 0) return 0;
    if (M % 64 || N % 64 || K % 16) {
        set_error("internal: product sizes must be multiples of 64 (M, N) and 16 (K)");
        return 1;
    }
    hipLaunchKernelGGL(k_gemm_f32<TRI>, dim3((unsigned)(N / 64), (unsigned)(M / 64)), dim3(256), 0, st, A, lda, B,
                       ldb, C, ldc, Cin, ldcin, K, alpha, beta, a_row0, a_col0);
    VX_CHECK(hipGetLastError());
    return 0;
}

// ---- factorisation ---------------------------------------------------------------
// F = the FP32 matrix padded with an identity block.
__global__ void k_lu_load(const double *__restrict__ L, int64_t n, int64_t np, float *__restrict__ F) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = blockIdx.y;
    if (j >= np) return;
    F[i * np + j] = (i < n && j < n) ? (float)L[i * n + j] : (i == j ? 1.f : 0.f);
}

// ||A||_inf (largest row sum of magnitudes), one wave per row, into st[5] as
// the bits of a non-negative double (their integer order is the value order).
__global__ __launch_bounds__(256) void k_lu_norm(const double *__restrict__ L, int64_t n,
                                                 unsigned long long *__restrict__ st) {
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    double s = 0.0;
    for (int64_t j = lane; j < n; j += 64) s += fabs(L[i * n + j]);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) atomicMax(st + 5, (unsigned long long)__double_as_longlong(s));
}

// One work-group factors the panel of columns [k0, k0 + 64) over rows
// [k0, np): per column, the largest magnitude below the diagonal (ties: the
// first row) is swapped up within the panel, the column is divided by it and
// the rest of the panel is updated.  ipiv[col] = the row swapped with col.
// The panel is factored in a column-major copy Pt (64 x np, column c at
// Pt + c np), so that the pivot search and every column update read and
// write consecutive rows (the row-major matrix would put each thread on its
// own row, 4 np bytes from the next), then copied back.
__global__ __launch_bounds__(1024) void k_lu_panel(float *__restrict__ F, int64_t np, int64_t k0,
                                                   int32_t *__restrict__ ipiv, unsigned long long *__restrict__ st,
                                                   float *__restrict__ Pt) {
    __shared__ float bv[1024];
    __shared__ int bi[1024];
    __shared__ float prow[kLuB];
    const int tid = threadIdx.x;
    const int64_t rows = np - k0;
    for (int64_t e = tid; e < rows * kLuB; e += 1024) {
        const int64_t i = k0 + e / kLuB;
        const int c = (int)(e % kLuB);
        Pt[c * np + i] = F[i * np + k0 + c];
    }
    __syncthreads();
    for (int jj = 0; jj < kLuB; ++jj) {
        const int64_t col = k0 + jj;
        float *pc = Pt + (int64_t)jj * np;
        float best = -1.f;
        int besti = (int)col;
        for (int64_t i = col + tid; i < np; i += 1024) {
            const float v = fabsf(pc[i]);
            if (v > best) { best = v; besti = (int)i; }
        }
        bv[tid] = best;
        bi[tid] = besti;
        __syncthreads();
        for (int s = 512; s > 0; s >>= 1) {
            if (tid < s) {
                const float o = bv[tid + s];
                const int oi = bi[tid + s];
                if (o > bv[tid] || (o == bv[tid] && oi < bi[tid])) { bv[tid] = o; bi[tid] = oi; }
            }
            __syncthreads();
        }
        const int64_t p = bi[0];
        if (tid == 0) ipiv[col] = (int32_t)p;
        if (p != col && tid < kLuB) {
            float *q = Pt + (int64_t)tid * np;
            const float a = q[col], b = q[p];
            q[col] = b;
            q[p] = a;
        }
        __syncthreads();
        if (tid < kLuB) prow[tid] = Pt[(int64_t)tid * np + col];
        __syncthreads();
        const float piv = prow[jj];
        if (piv == 0.f) {
            if (tid == 0) st[4] = 1ull;  // singular: leave the column as it is
            __syncthreads();
            continue;
        }
        for (int64_t i = col + 1 + tid; i < np; i += 1024) {
            const float l = pc[i] / piv;
            pc[i] = l;
            for (int c = jj + 1; c < kLuB; ++c) Pt[(int64_t)c * np + i] -= l * prow[c];
        }
        __syncthreads();
    }
    for (int64_t e = tid; e < rows * kLuB; e += 1024) {
        const int64_t i = k0 + e / kLuB;
        const int c = (int)(e % kLuB);
        F[i * np + k0 + c] = Pt[c * np + i];
    }
}

// The panel's row swaps on every other column; block 0 also carries them
// into the row permutation.
__global__ __launch_bounds__(256) void k_lu_swap(float *__restrict__ F, int64_t np, int64_t k0,
                                                 const int32_t *__restrict__ ipiv, int32_t *__restrict__ perm) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int jj = 0; jj < kLuB; ++jj) {
            const int64_t a = k0 + jj, b = ipiv[a];
            const int32_t t = perm[a];
            perm[a] = perm[b];
            perm[b] = t;
        }
    if (c >= np || (c >= k0 && c < k0 + kLuB)) return;
    for (int jj = 0; jj < kLuB; ++jj) {
        const int64_t a = k0 + jj, b = ipiv[a];
        if (b != a) {
            const float u = F[a * np + c], v = F[b * np + c];
            F[a * np + c] = v;
            F[b * np + c] = u;
        }
    }
}

// U12 = L11^-1 A12 for the columns right of the panel, one column per thread
// held in registers, L11 in LDS.
__global__ __launch_bounds__(256) void k_lu_trsm(float *__restrict__ F, int64_t np, int64_t k0) {
    __shared__ float L11[kLuB][kLuB + 1];
    for (int e = threadIdx.x; e < kLuB * kLuB; e += 256) L11[e / kLuB][e % kLuB] = F[(k0 + e / kLuB) * np + k0 + e % kLuB];
    __syncthreads();
    const int64_t c = k0 + kLuB + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= np) return;
    float v[kLuB];
#pragma unroll
    for (int r = 0; r < kLuB; ++r) v[r] = F[(k0 + r) * np + c];
#pragma unroll
    for (int r = 1; r < kLuB; ++r) {
        float s = v[r];
#pragma unroll
        for (int q = 0; q < r; ++q) s -= L11[r][q] * v[q];
        v[r] = s;
    }
#pragma unroll
    for (int r = 1; r < kLuB; ++r) F[(k0 + r) * np + c] = v[r];
}

// Inverses of the diagonal blocks, one work-group of 64 threads per block,
// one column of the inverse per thread: W_l = L_jj^-1 (unit lower, with its
// unit diagonal), W_u = U_jj^-1 (upper).
__global__ __launch_bounds__(kLuB) void k_lu_inv_diag(const float *__restrict__ F, int64_t np,
                                                      float *__restrict__ Wl, float *__restrict__ Wu) {
    __shared__ float D[kLuB][kLuB + 1];
    const int64_t j0 = (int64_t)blockIdx.x * kLuB;
    const int c = threadIdx.x;
    for (int r = 0; r < kLuB; ++r) D[r][c] = F[(j0 + r) * np + j0 + c];
    __syncthreads();
    float x[kLuB];
    // lower, unit diagonal: forward substitution of e_c
#pragma unroll
    for (int r = 0; r < kLuB; ++r) {
        float s = r == c ? 1.f : 0.f;
#pragma unroll
        for (int q = 0; q < r; ++q) s -= D[r][q] * x[q];
        x[r] = r < c ? 0.f : s;
    }
#pragma unroll
    for (int r = 0; r < kLuB; ++r) Wl[(j0 + r) * np + j0 + c] = x[r];
    // upper: backward substitution of e_c
#pragma unroll
    for (int r = kLuB - 1; r >= 0; --r) {
        float s = r == c ? 1.f : 0.f;
#pragma unroll
        for (int q = r + 1; q < kLuB; ++q) s -= D[r][q] * x[q];
        x[r] = r > c ? 0.f : s / D[r][r];
    }
#pragma unroll
    for (int r = 0; r < kLuB; ++r) Wu[(j0 + r) * np + j0 + c] = x[r];
}

// W = strict lower part of Wl, upper part (with the diagonal) of Wu; the
// inverse permutation.
__global__ void k_lu_merge(const float *__restrict__ Wl, const float *__restrict__ Wu, int64_t np,
                           float *__restrict__ W) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = blockIdx.y;
    if (j >= np) return;
    W[i * np + j] = j < i ? Wl[i * np + j] : Wu[i * np + j];
}

__global__ void k_lu_perm_init(int32_t *__restrict__ perm, int64_t np) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) perm[i] = (int32_t)i;
}

__global__ void k_lu_pinv(const int32_t *__restrict__ perm, int64_t np, int32_t *__restrict__ pinv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) pinv[perm[i]] = (int32_t)i;
}

// ---- per-solve refinement ----------------------------------------------------------
__device__ __forceinline__ unsigned long long dbits(double v) { return (unsigned long long)__double_as_longlong(v); }
__device__ __forceinline__ double bitsd(unsigned long long v) { return __longlong_as_double((long long)v); }

__global__ void k_lu_begin(double *__restrict__ x, int64_t np, unsigned long long *__restrict__ st) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) x[i] = 0.0;
    if (i < 4) st[i] = 0ull;
}

// r = b - A x with b = [R; d_g], one wave per row, written permuted and in
// FP32 (rp[pinv[i]]); the largest |r_i| into st[2].
__global__ __launch_bounds__(256) void k_lu_resid(const double *__restrict__ A, const double *__restrict__ R,
                                                  double d_g_host, const double *__restrict__ d_g_dev, int64_t n,
                                                  int64_t np, const double *__restrict__ x,
                                                  const int32_t *__restrict__ pinv, float *__restrict__ rp,
                                                  unsigned long long *__restrict__ st) {
    if (st[0]) return;
    const double d_g = d_g_dev ? *d_g_dev : d_g_host;
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= np) return;
    if (i >= n) {
        if (lane == 0) rp[pinv[i]] = 0.f;
        return;
    }
    const double *a = A + i * n;
    double s = 0.0;
    for (int64_t j = lane; j < n; j += 64) s += a[j] * x[j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) {
        const double b = i < n - 1 ? R[i] : d_g;
        const double r = b - s;
        rp[pinv[i]] = (float)r;
        atomicMax(st + 2, dbits(fabs(r)));
    }
}

// Convergence test (LAPACK dsgesv): done when ||r|| <= ||x|| ||A|| eps sqrt(n);
// else one more correction is counted.  Resets the maxima for the next pass.
__global__ void k_lu_check(unsigned long long *__restrict__ st, double *__restrict__ nrm, int64_t n) {
    if (st[0]) return;
    const double r = bitsd(st[2]), xm = bitsd(st[3]), an = bitsd(st[5]);
    const double cte = an * 1.1102230246251565e-16 * sqrt((double)n);
    nrm[1] = (xm > 0.0 && an > 0.0) ? r / (xm * an) : r;
    if (r <= xm * cte) st[0] = 1ull;
    else st[1] += 1ull;
    st[2] = 0ull;
}

// t = L^-1 (P r): row i sums W's strict lower part (one wave per row).
__global__ __launch_bounds__(256) void k_lu_apply_l(const float *__restrict__ W, int64_t np,
                                                    const float *__restrict__ rp, float *__restrict__ t,
                                                    const unsigned long long *__restrict__ st) {
    if (st[0]) return;
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= np) return;
    const float *w = W + i * np;
    float s = 0.f;
    for (int64_t j = lane; j < i; j += 64) s += w[j] * rp[j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) t[i] = s + rp[i];
}

// x += U^-1 t: row i sums W's upper part; the largest |x_i| into st[3].
__global__ __launch_bounds__(256) void k_lu_apply_u(const float *__restrict__ W, int64_t np,
                                                    const float *__restrict__ t, double *__restrict__ x,
                                                    unsigned long long *__restrict__ st) {
    if (st[0]) return;
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= np) return;
    const float *w = W + i * np;
    float s = 0.f;
    for (int64_t j = i + lane; j < np; j += 64) s += w[j] * t[j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) {
        const double xi = x[i] + (double)s;
        x[i] = xi;
        atomicMax(st + 3, dbits(fabs(xi)));
    }
}

// ---- host side ------------------------------------------------------------------------
static LuBufs &lu_of(vx_case *c) {
    if (!c->lu) c->lu = new LuBufs();
    return *c->lu;
}

void lu_release(vx_case *c) {
    if (!c->lu) return;
    LuBufs &b = *c->lu;
    b.F.release(); b.W.release(); b.Wl.release(); b.Wu.release(); b.T.release(); b.A.release();
    b.x.release(); b.rp.release(); b.t.release(); b.ipiv.release(); b.perm.release(); b.pinv.release();
    b.st.release(); b.nrm.release();
    delete c->lu;
    c->lu = nullptr;
}

// Factor c->L (the FP64 panel matrix, (P+1)^2) and invert the factors.
int launch_lu_setup(vx_case *c) {
    LuBufs &b = lu_of(c);
    hipStream_t st = c->stream;
    const int64_t n = c->P + 1, np = (n + kLuB - 1) / kLuB * kLuB;
    b.n = n;
    b.np = np;
    const size_t nn = (size_t)(np * np);
    if (b.F.ensure(nn, st) || b.Wl.ensure(nn, st) || b.Wu.ensure(nn, st) || b.W.ensure(nn, st) ||
        b.T.ensure((size_t)(np * kLuB), st) || b.A.ensure((size_t)(n * n), st) || b.x.ensure((size_t)np, st) ||
        b.rp.ensure((size_t)np, st) || b.t.ensure((size_t)np, st) || b.ipiv.ensure((size_t)np, st) ||
        b.perm.ensure((size_t)np, st) || b.pinv.ensure((size_t)np, st) || b.st.ensure(8, st) || b.nrm.ensure(2, st))
        return 1;
    VX_CHECK(hipMemcpyAsync(b.A.p, c->L.p, sizeof(double) * n * n, hipMemcpyDeviceToDevice, st));
    VX_CHECK(prim::fill(b.st.p, 8, 0ull, st));
    VX_CHECK(prim::fill(b.Wl.p, (int64_t)nn, 0.f, st));
    VX_CHECK(prim::fill(b.Wu.p, (int64_t)nn, 0.f, st));
    const dim3 g2((unsigned)((np + 255) / 256), (unsigned)np);
    const unsigned g1 = (unsigned)((np + 255) / 256);
    hipLaunchKernelGGL(k_lu_load, g2, dim3(256), 0, st, c->L.p, n, np, b.F.p);
    hipLaunchKernelGGL(k_lu_norm, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, c->L.p, n, b.st.p);
    hipLaunchKernelGGL(k_lu_perm_init, dim3(g1), dim3(256), 0, st, b.perm.p, np);
    // right-looking LU: panel, swaps, block row, trailing product
    for (int64_t k0 = 0; k0 < np; k0 += kLuB) {
        hipLaunchKernelGGL(k_lu_panel, dim3(1), dim3(1024), 0, st, b.F.p, np, k0, b.ipiv.p, b.st.p, b.T.p);
        hipLaunchKernelGGL(k_lu_swap, dim3(g1), dim3(256), 0, st, b.F.p, np, k0, b.ipiv.p, b.perm.p);
        const int64_t rest = np - k0 - kLuB;
        if (rest > 0) {
            hipLaunchKernelGGL(k_lu_trsm, dim3((unsigned)((rest + 255) / 256)), dim3(256), 0, st, b.F.p, np, k0);
            float *A21 = b.F.p + (k0 + kLuB) * np + k0;
            float *A12 = b.F.p + k0 * np + k0 + kLuB;
            float *A22 = b.F.p + (k0 + kLuB) * np + k0 + kLuB;
            if (gemm<0>(A21, np, A12, np, A22, np, A22, np, rest, rest, kLuB, -1.f, 1.f, 0, 0, st)) return 1;
        }
    }
    VX_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_lu_pinv, dim3(g1), dim3(256), 0, st, b.perm.p, np, b.pinv.p);
    // inverses of the factors: diagonal blocks, then block columns
    //   L: from the last block column back, Wl[j+, j] = -(Wl[j+, j+] L[j+, j]) Wl[j, j]
    //   U: from the first forward,          Wu[:j, j] = -(Wu[:j, :j] U[:j, j]) Wu[j, j]
    hipLaunchKernelGGL(k_lu_inv_diag, dim3((unsigned)(np / kLuB)), dim3(kLuB), 0, st, b.F.p, np, b.Wl.p, b.Wu.p);
    for (int64_t j0 = np - kLuB; j0 >= 0; j0 -= kLuB) {
        const int64_t r0 = j0 + kLuB, m = np - r0;
        if (m <= 0) continue;
        if (gemm<1>(b.Wl.p + r0 * np + r0, np, b.F.p + r0 * np + j0, np, b.T.p, kLuB, b.T.p, kLuB, m, kLuB, m, 1.f,
                    0.f, r0, r0, st) ||
            gemm<0>(b.T.p, kLuB, b.Wl.p + j0 * np + j0, np, b.Wl.p + r0 * np + j0, np, b.Wl.p, np, m, kLuB, kLuB,
                    -1.f, 0.f, 0, 0, st))
            return 1;
    }
    for (int64_t j0 = kLuB; j0 < np; j0 += kLuB) {
        if (gemm<2>(b.Wu.p, np, b.F.p + j0, np, b.T.p, kLuB, b.T.p, kLuB, j0, kLuB, j0, 1.f, 0.f, 0, 0, st) ||
            gemm<0>(b.T.p, kLuB, b.Wu.p + j0 * np + j0, np, b.Wu.p + j0, np, b.Wu.p, np, j0, kLuB, kLuB, -1.f, 0.f,
                    0, 0, st))
            return 1;
    }
    hipLaunchKernelGGL(k_lu_merge, g2, dim3(256), 0, st, b.Wl.p, b.Wu.p, np, b.W.p);
    VX_CHECK(hipGetLastError());
    unsigned long long sing = 0;
    VX_CHECK(hipMemcpyAsync(&sing, b.st.p + 4, sizeof(sing), hipMemcpyDeviceToHost, st));
    VX_CHECK(hipStreamSynchronize(st));
    // the factors' staging is not needed per step
    b.F.release(); b.Wl.release(); b.Wu.release(); b.T.release();
    if (sing) {
        set_error("panel matrix is singular in FP32 (zero pivot)");
        return 1;
    }
    return 0;
}

// Solve A x = [R; d_g] into c->lu->x (first P entries: the sheet strengths);
// d_g from *d_g_dev when that is not null.
int launch_lu_solve(vx_case *c, const double *R, double d_g, const double *d_g_dev) {
    LuBufs &b = *c->lu;
    hipStream_t st = c->stream;
    const int64_t n = b.n, np = b.np;
    const unsigned gw = (unsigned)((np + 3) / 4);
    hipLaunchKernelGGL(k_lu_begin, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, b.x.p, np, b.st.p);
    hipLaunchKernelGGL(k_lu_resid, dim3(gw), dim3(256), 0, st, b.A.p, R, d_g, d_g_dev, n, np, b.x.p, b.pinv.p,
                       b.rp.p, b.st.p);
    hipLaunchKernelGGL(k_lu_check, dim3(1), dim3(1), 0, st, b.st.p, b.nrm.p, n);
    for (int it = 0; it < kLuMaxIt; ++it) {
        hipLaunchKernelGGL(k_lu_apply_l, dim3(gw), dim3(256), 0, st, b.W.p, np, b.rp.p, b.t.p, b.st.p);
        hipLaunchKernelGGL(k_lu_apply_u, dim3(gw), dim3(256), 0, st, b.W.p, np, b.t.p, b.x.p, b.st.p);
        hipLaunchKernelGGL(k_lu_resid, dim3(gw), dim3(256), 0, st, b.A.p, R, d_g, d_g_dev, n, np, b.x.p, b.pinv.p,
                           b.rp.p, b.st.p);
        hipLaunchKernelGGL(k_lu_check, dim3(1), dim3(1), 0, st, b.st.p, b.nrm.p, n);
    }
    VX_CHECK(hipGetLastError());
    return 0;
}

const double *lu_solution(vx_case *c) { return c->lu ? c->lu->x.p : nullptr; }

int lu_info(vx_case *c, int32_t *iters, int32_t *converged, double *resid) {
    if (!c->lu) {
        set_error("bad argument: the case has no FP32 LU (VX_FLAG_LU_REFINE)");
        return 2;
    }
    unsigned long long s[2] = {0, 0};
    double r = 0.0;
    VX_CHECK(hipMemcpyAsync(s, c->lu->st.p, sizeof(s), hipMemcpyDeviceToHost, c->stream));
    VX_CHECK(hipMemcpyAsync(&r, c->lu->nrm.p + 1, sizeof(r), hipMemcpyDeviceToHost, c->stream));
    VX_CHECK(hipStreamSynchronize(c->stream));
    if (iters) *iters = (int32_t)s[1];
    if (converged) *converged = (int32_t)s[0];
    if (resid) *resid = r;
    return 0;
}

}  // namespace vx
