// Morton-ordered tree for the two O(N^2) velocity stages at large N.
//
// Layout.  Particles that can influence anything (g != 0, or within 400 of
// the deletion box -- farther dead slots can never be among the three
// nearest, whose sentinel is r^2 = 1e5) are sorted by a 32-bit Morton key
// (16 bits per axis over their bounding box).  Leaves are runs of 64
// consecutive sorted particles; level l+1 pairs the nodes of level l, so every
// node is a contiguous range of the sorted order.  Each node keeps its
// bounding box, centre, radius (half diagonal) and the normalised complex
// multipole b_k = sum_j g_j ((z_j - c) / R)^k, k < ORDER (P2M at the leaves,
// binomial M2M above).
//
// Biot-Savart (src/kernel.cu:875-983 semantics).  One wave64 per target leaf
// walks the tree with a stack held one entry per lane.  A node is far when
// theta |c_node - c_leaf| > R_node + R_leaf and the closest possible pair is
// beyond the smoothing radius (|c_node - c_leaf| - R_node - R_leaf > eps, so
// max(r^2, eps^2) = r^2 there); far nodes are evaluated from their
// multipoles (M2P, Horner in R/(z - c)), near leaves pairwise.  The nearest-3
// distance comes from the pairwise part; a target whose third distance is not
// below its distance to every far node is recomputed exactly by a pruned
// search over the whole tree.
//
// Diffusion (src/kernel.cu:985-1029).  The weights e^{-r/d_i} are summed over
// the leaves whose boxes come within diff_cut * max(d) of the target leaf's
// box; every omitted source has e^{-r/d_i} < e^{-diff_cut} (diff_cut defaults
// to 20 in mixed precision and 30 in FP64).
//
// Mixed precision: pair and M2P arithmetic in FP32 on coordinates relative to
// the leaf centres (so close pairs keep FP32 relative accuracy), partial sums
// per source leaf / far node accumulated in FP64.  FP64: absolute FP64
// coordinates throughout.
#include <utility>
#include <vector>

#include "vx_internal.h"
#include "vx_prim.h"

namespace vx {

constexpr int kLeaf = 64;
constexpr double kTiny = 1e-300;

// Binomial coefficient C(n, k); folded to a constant wherever n and k are.
__host__ __device__ constexpr double binom(int n, int k) {
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (double)(n - k + i) / (double)i;
    return r;
}


// Traversal record of a node, one 32-byte load: centre, radius and either
// the children (a = child0 >= 0, b = child1) or, for a bucket, its particles
// (a = -1 - first, b = count).
struct NodeRec {
    double cx, cy, r;
    int32_t a, b;
};

struct TreeBufs {
    DBuf<uint32_t> keys, keys_alt;
    DBuf<int32_t> idx, perm, incl, child0, child1, first, count, parent, fb;
    DBuf<int32_t> rlo, rhi, rleft, rright, bstart, bscan, upper, uscan, toplist;
    DBuf<int32_t> gcs, gcs2, gflag, gscan, gfirst;  // target groups
    DBuf<int32_t> pcount;                          // included particles per block
    DBuf<int32_t> arr;                             // cut nodes (count in cnt[10])
    DBuf<int64_t> sizes;                           // group, bucket and upper-node counts
    DBuf<double> sx, sy, sg, ncx, ncy, nr, fbr2, hpart, sd;
    DBuf<float> fx, fy, fg;
    DBuf<double4> nbb, part, cbox, ftgt;
    DBuf<int2> fpairs;
    DBuf<double2> cd;
    DBuf<float2> cf;
    DBuf<NodeRec> rec;
    DBuf<unsigned long long> cnt;   // counters, zeroed by each build; slot 15 (int32): the key sort's error flag
    DBuf<unsigned char> tmp, tmp2;  // scan scratch of the build's two concurrent chains
    bool sd_fresh = false;          // sd holds d in sorted order, as the last Biot-Savart pass wrote it
    int64_t s = 0;                         // slots keyed and sorted by the last build
    int64_t m = 0, nleaf = 0, nnode = 0;  // nleaf: buckets, nodes [0, nleaf)
    int64_t ngroup = 0;                    // target groups (gfirst)
    int32_t root = 0;
    int order = 16;  // coefficients stored per node
    int terms = 16;  // of them used by the far field
    int64_t fcap = 4096;  // domain-scale diffusion targets listed at most (set per diffusion)
    unsigned long long *h_far = nullptr;  // pinned: the last diffusion's domain-scale target count
    hipEvent_t ev = nullptr;                       // read-back completion
};

struct TreeDev {
    int64_t m, nleaf;
    int32_t root;
    const int32_t *perm;
    const double *sx, *sy, *sg;
    const float *fx, *fy, *fg;
    const double *ncx, *ncy, *nr;
    const double4 *nbb;
    const int32_t *child0, *child1, *first, *count;
    const double2 *cd;
    const float2 *cf;
    const NodeRec *rec;
    const int32_t *gfirst;
    int64_t ngroup;
    const int32_t *parent, *bscan, *bstart;
    int64_t fcap;  // capacity of the domain-scale diffusion target list
};

static TreeDev view(const TreeBufs &t) {
    return TreeDev{t.m,     t.nleaf, t.root,     t.perm.p,   t.sx.p,    t.sy.p,    t.sg.p, t.fx.p, t.fy.p, t.fg.p,
                   t.ncx.p, t.ncy.p, t.nr.p,     t.nbb.p,    t.child0.p, t.child1.p, t.first.p, t.count.p,
                   t.cd.p,  t.cf.p,  t.rec.p, t.gfirst.p, t.ngroup, t.parent.p, t.bscan.p, t.bstart.p,
                   t.fcap};
}

// ---- wave helpers ----------------------------------------------------------
__device__ __forceinline__ double wave_min(double a) {
    for (int o = 32; o > 0; o >>= 1) a = fmin(a, __shfl_xor(a, o, 64));
    return a;
}
__device__ __forceinline__ double wave_max(double a) {
    for (int o = 32; o > 0; o >>= 1) a = fmax(a, __shfl_xor(a, o, 64));
    return a;
}
__device__ __forceinline__ double wave_sum(double a) {
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    return a;
}

__device__ __forceinline__ uint32_t spread16(uint32_t v) {
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

// ---- build -----------------------------------------------------------------
// Newborn slots [at, at + P) whose positions come from bp (x at [0, P), y at
// [P, 2P)) instead of the state arrays, which the birth may still be writing;
// they lie inside the inclusion margin, so they are included whatever their
// circulation.  P = 0: none.
struct Births {
    int64_t at, P;
    const double *bp;
    __device__ bool has(int64_t i) const { return i >= at && i - at < P; }
};

__global__ __launch_bounds__(256) void k_incl_box(const double *__restrict__ x, const double *__restrict__ y,
                                                  const double *__restrict__ g, int64_t s, double4 dbox,
                                                  int32_t *__restrict__ incl, double4 *__restrict__ part,
                                                  int32_t *__restrict__ pcount, Births B) {
    __shared__ double r[4][256];
    __shared__ int c[256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
    int in = 0;
    if (i < s) {
        const bool nb = B.has(i);
        const double xi = nb ? B.bp[i - B.at] : x[i], yi = nb ? B.bp[B.P + i - B.at] : y[i];
        if (nb) {
            in = 1;
        } else if (isfinite(xi) && isfinite(yi)) {
            const double gi = g[i];
            const double ex = fmax(0.0, fmax(dbox.x - xi, xi - dbox.y));
            const double ey = fmax(0.0, fmax(dbox.z - yi, yi - dbox.w));
            in = (gi != 0.0) || (ex * ex + ey * ey <= 400.0 * 400.0);
        }
        incl[i] = in;
        if (in) { x0 = x1 = xi; y0 = y1 = yi; }
    }
    r[0][threadIdx.x] = x0; r[1][threadIdx.x] = -x1; r[2][threadIdx.x] = y0; r[3][threadIdx.x] = -y1;
    c[threadIdx.x] = in;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            for (int q = 0; q < 4; ++q) r[q][threadIdx.x] = fmin(r[q][threadIdx.x], r[q][threadIdx.x + st]);
            c[threadIdx.x] += c[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = make_double4(r[0][0], -r[1][0], r[2][0], -r[3][0]);
        pcount[blockIdx.x] = c[0];
    }
}

// Box of the included particles and their count (cnt[0]) from the block partials.
// The included count also goes straight to the host: *hout, pinned and
// coherent, by a system-scope store (the host reads it after an event that
// follows this kernel).
__global__ __launch_bounds__(256) void k_box_final(const double4 *__restrict__ part, const int32_t *__restrict__ pcount,
                                                   int64_t nb, double4 *__restrict__ box,
                                                   unsigned long long *__restrict__ cnt, int64_t *__restrict__ hout) {
    __shared__ double r[4][256];
    __shared__ long long rc[256];
    double a[4] = {1e300, 1e300, 1e300, 1e300};
    long long k = 0;
    for (int64_t b = threadIdx.x; b < nb; b += 256) {
        const double4 p = part[b];
        a[0] = fmin(a[0], p.x); a[1] = fmin(a[1], -p.y); a[2] = fmin(a[2], p.z); a[3] = fmin(a[3], -p.w);
        k += pcount[b];
    }
    for (int q = 0; q < 4; ++q) r[q][threadIdx.x] = a[q];
    rc[threadIdx.x] = k;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            for (int q = 0; q < 4; ++q) r[q][threadIdx.x] = fmin(r[q][threadIdx.x], r[q][threadIdx.x + st]);
            rc[threadIdx.x] += rc[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cnt[0] = (unsigned long long)rc[0];
        __hip_atomic_store(hout, (int64_t)rc[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __threadfence_system();
        double4 b = make_double4(r[0][0], -r[1][0], r[2][0], -r[3][0]);
        if (!(b.x <= b.y)) b = make_double4(0.0, 1.0, 0.0, 1.0);
        box[0] = b;
    }
}

__global__ void k_tree_keys(const double *__restrict__ x, const double *__restrict__ y,
                            const int32_t *__restrict__ incl, int64_t s, const double4 *__restrict__ box,
                            uint32_t *__restrict__ keys, int32_t *__restrict__ vals, Births B) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s) return;
    vals[i] = (int32_t)i;
    if (!incl[i]) {
        keys[i] = 0xFFFFFFFFu;
        return;
    }
    const double4 b = box[0];
    const double wx = b.y - b.x, wy = b.w - b.z;
    const bool nb = B.has(i);
    const double xi = nb ? B.bp[i - B.at] : x[i], yi = nb ? B.bp[B.P + i - B.at] : y[i];
    double fx = wx > 0 ? (xi - b.x) / wx * 65534.0 : 0.0;
    double fy = wy > 0 ? (yi - b.z) / wy * 65534.0 : 0.0;
    fx = fmin(fmax(fx, 0.0), 65534.0);
    fy = fmin(fmax(fy, 0.0), 65534.0);
    keys[i] = spread16((uint32_t)fx) | (spread16((uint32_t)fy) << 1);  // at most 0xFFFFFFFC
}

__global__ void k_tree_gather(const double *__restrict__ x, const double *__restrict__ y,
                              const double *__restrict__ g, const int32_t *__restrict__ perm, int64_t m,
                              double *__restrict__ sx, double *__restrict__ sy, double *__restrict__ sg) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const int32_t i = perm[t];
    sx[t] = x[i];
    sy[t] = y[i];
    sg[t] = g[i];
}

// ---- binary radix tree over the sorted keys (Karras 2012) ---------------------
// delta(i, j): length of the common prefix of keys i and j, with the index as
// tie-breaker for equal keys; -1 outside [0, m).
__device__ __forceinline__ int rdelta(const uint32_t *__restrict__ k, int64_t m, int64_t i, int64_t j) {
    if (j < 0 || j >= m) return -1;
    const uint32_t a = k[i], b = k[j];
    if (a != b) return __clz(a ^ b);
    return 32 + __clz((uint32_t)i ^ (uint32_t)j);
}

// Internal node i covers [lo_i, hi_i]; children >= 0 are internal nodes,
// -(t+1) the single particle t.
__global__ void k_radix(const uint32_t *__restrict__ key, int64_t m, int32_t *__restrict__ lo,
                        int32_t *__restrict__ hi, int32_t *__restrict__ left, int32_t *__restrict__ right) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m - 1) return;
    const int d = (rdelta(key, m, i, i + 1) - rdelta(key, m, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = rdelta(key, m, i, i - d);
    int64_t lmax = 2;
    while (rdelta(key, m, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (rdelta(key, m, i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = rdelta(key, m, i, j);
    int64_t sp = 0;
    for (int64_t t = (l + 1) / 2;; t = (t + 1) / 2) {
        if (rdelta(key, m, i, i + (sp + t) * d) > dnode) sp += t;
        if (t == 1) break;
    }
    const int64_t gam = i + sp * d + (d < 0 ? -1 : 0);
    const int64_t a = i < j ? i : j, b = i < j ? j : i;
    lo[i] = (int32_t)a;
    hi[i] = (int32_t)b;
    left[i] = (a == gam) ? -(int32_t)(gam + 1) : (int32_t)gam;
    right[i] = (b == gam + 1) ? -(int32_t)(gam + 2) : (int32_t)(gam + 1);
}

__device__ __forceinline__ int64_t rsize(const int32_t *lo, const int32_t *hi, int32_t c) {
    return c >= 0 ? (int64_t)hi[c] - lo[c] + 1 : 1;
}
__device__ __forceinline__ int32_t rfirst(const int32_t *lo, int32_t c) { return c >= 0 ? lo[c] : -c - 1; }

// Buckets: nodes of at most kLeaf particles whose cell (common key prefix of
// at least dmin bits) is no coarser than about twice the typical bucket, and
// whose parent is not.  The depth condition keeps sparse regions from forming
// large buckets whose near field would cover much of the domain.  Marks the
// first particle of each bucket and the internal nodes kept above them.
__device__ __forceinline__ bool is_upper(const uint32_t *__restrict__ key, const int32_t *lo, const int32_t *hi,
                                         int32_t i, int dmin) {
    const int64_t sz = (int64_t)hi[i] - lo[i] + 1;
    if (sz > kLeaf) return true;
    const uint32_t a = key[lo[i]], b = key[hi[i]];
    const int pref = a == b ? 32 : __clz(a ^ b);
    return pref < dmin;
}

__global__ void k_bucket_mark(const uint32_t *__restrict__ key, const int32_t *__restrict__ lo,
                              const int32_t *__restrict__ hi, const int32_t *__restrict__ left,
                              const int32_t *__restrict__ right, int64_t m, int dmin, int32_t *__restrict__ bstart,
                              int32_t *__restrict__ upper) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m - 1) return;
    const bool up = is_upper(key, lo, hi, (int32_t)i, dmin);
    upper[i] = up ? 1 : 0;
    if (!up) return;
    const int32_t ch[2] = {left[i], right[i]};
    for (int q = 0; q < 2; ++q)
        if (ch[q] < 0 || !is_upper(key, lo, hi, ch[q], dmin)) bstart[rfirst(lo, ch[q])] = 1;
}

// Compact node array: buckets [0, nb) in particle order, then the kept
// internal nodes.  The buckets' first particles, then their counts.
__global__ void k_bucket_first(const int32_t *__restrict__ bstart, const int32_t *__restrict__ bscan, int64_t m,
                               int32_t *__restrict__ first) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m && bstart[t]) first[bscan[t]] = (int32_t)t;
}

// sz: the device counts of k_tree_counts (sz[1] buckets, sz[2] upper nodes).
__global__ void k_bucket_count(const int32_t *__restrict__ first, const int64_t *__restrict__ sz, int64_t m,
                               int32_t *__restrict__ count) {
    const int64_t nb = sz[1];
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) count[b] = (int32_t)((b + 1 < nb ? (int64_t)first[b + 1] : m) - first[b]);
}

// parent[0, nn) = -1 for the nn = sz[1] + sz[2] nodes.
__global__ void k_parent_init(const int64_t *__restrict__ sz, int32_t *__restrict__ parent) {
    const int64_t nn = sz[1] + sz[2];
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < nn; u += (int64_t)gridDim.x * blockDim.x)
        parent[u] = -1;
}

__global__ void k_upper_nodes(const int32_t *__restrict__ lo, const int32_t *__restrict__ hi,
                              const int32_t *__restrict__ left, const int32_t *__restrict__ right,
                              const int32_t *__restrict__ upper, const int32_t *__restrict__ uscan,
                              const int32_t *__restrict__ bscan, int64_t m, const int64_t *__restrict__ sz,
                              int32_t *__restrict__ first, int32_t *__restrict__ count, int32_t *__restrict__ child0,
                              int32_t *__restrict__ child1, int32_t *__restrict__ parent) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m - 1 || !upper[i]) return;
    const int64_t nb = sz[1];
    const int32_t u = (int32_t)(nb + uscan[i]);
    const int32_t ch[2] = {left[i], right[i]};
    int32_t id[2];
    for (int q = 0; q < 2; ++q) {
        const int32_t c = ch[q];
        id[q] = (c >= 0 && upper[c]) ? (int32_t)(nb + uscan[c]) : bscan[rfirst(lo, c)];  // bucket: its first particle
        parent[id[q]] = u;
    }
    child0[u] = id[0];
    child1[u] = id[1];
    first[u] = lo[i];
    count[u] = hi[i] - lo[i] + 1;
}

// One wave per bucket: box, centre, radius, bucket-relative FP32 copies, P2M.
// ORDER is a multiple of 8.
template <int ORDER>
__global__ __launch_bounds__(256) void k_bucket_build(const double *__restrict__ sx, const double *__restrict__ sy,
                                                      const double *__restrict__ sg, int64_t nb,
                                                      const int32_t *__restrict__ first,
                                                      const int32_t *__restrict__ count, float *__restrict__ fx,
                                                      float *__restrict__ fy, float *__restrict__ fg,
                                                      double *__restrict__ ncx, double *__restrict__ ncy,
                                                      double *__restrict__ nr, double4 *__restrict__ nbb,
                                                      int32_t *__restrict__ child0, int32_t *__restrict__ child1,
                                                      double2 *__restrict__ cd, float2 *__restrict__ cf) {
    __shared__ double2 tr_all[4][8][65];
    const int lane = threadIdx.x & 63;
    const int64_t b = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= nb) return;
    const int64_t t = (int64_t)first[b] + lane;
    const bool ok = lane < count[b];
    const double xi = ok ? sx[t] : 0.0, yi = ok ? sy[t] : 0.0, gi = ok ? sg[t] : 0.0;
    const double x0 = wave_min(ok ? xi : 1e300), x1 = wave_max(ok ? xi : -1e300);
    const double y0 = wave_min(ok ? yi : 1e300), y1 = wave_max(ok ? yi : -1e300);
    const double cx = 0.5 * (x0 + x1), cy = 0.5 * (y0 + y1);
    // radius: the farthest particle from the box centre (at most the half
    // diagonal; a tighter circle opens fewer nodes)
    const double R = sqrt(wave_max(ok ? (xi - cx) * (xi - cx) + (yi - cy) * (yi - cy) : 0.0));
    if (ok) {
        fx[t] = (float)(xi - cx);
        fy[t] = (float)(yi - cy);
        fg[t] = (float)gi;
    }
    const double Rs = fmax(R, kTiny);
    // idle lanes sit at the centre: their zero weight must not meet an
    // overflowing power ((0 - c)/R)^k
    const double zr = ok ? (xi - cx) / Rs : 0.0, zi = ok ? (yi - cy) / Rs : 0.0;
    // b_k = sum_j g_j ((z_j - c)/R)^k, eight terms at a time: each lane writes
    // its particle's terms to LDS (term-major, rows padded against bank
    // conflicts), lane (k, e) sums term k over the e-th eighth of the
    // particles, and three shuffles add the eighths.
    double2(*tr)[65] = tr_all[threadIdx.x >> 6];
    const int kk = lane & 7, e8 = lane >> 3;
    double pr = 1.0, pi = 0.0;  // ((z - c)/R)^k
    for (int k0 = 0; k0 < ORDER; k0 += 8) {
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            tr[k][lane] = make_double2(gi * pr, gi * pi);
            const double npr = pr * zr - pi * zi;
            pi = pr * zi + pi * zr;
            pr = npr;
        }
        __builtin_amdgcn_wave_barrier();
        double br = 0.0, bi = 0.0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const double2 q = tr[kk][8 * e8 + j];
            br += q.x;
            bi += q.y;
        }
#pragma unroll
        for (int o = 8; o < 64; o <<= 1) {
            br += __shfl_xor(br, o, 64);
            bi += __shfl_xor(bi, o, 64);
        }
        if (e8 == 0) {
            cd[b * ORDER + k0 + kk] = make_double2(br, bi);
            cf[b * ORDER + k0 + kk] = make_float2((float)br, (float)bi);
        }
    }
    if (lane == 0) {
        ncx[b] = cx;
        ncy[b] = cy;
        nr[b] = R;
        nbb[b] = make_double4(x0, x1, y0, y1);
        child0[b] = -1;
        child1[b] = -1;
    }
}

// Bottom-up pass (above the buckets) in two kernels, each building levels
// deepest first with a work-group barrier between them (build_levels below).
//   Cut nodes hold at most kCut buckets and have a parent (or are the root)
//   with more; their subtrees partition the buckets.
//   k_upper_local: a work-group per cut node builds the cut node's subtree.
//   k_upper_top: one work-group builds the nodes above the cut nodes.
constexpr int kCut = 256;

__device__ __forceinline__ int64_t bucket_lo(int32_t u, int64_t nb, const int32_t *__restrict__ first,
                                             const int32_t *__restrict__ bscan) {
    return u < nb ? (int64_t)u : (int64_t)bscan[first[u]];
}

__device__ __forceinline__ int64_t bucket_n(int32_t u, int64_t nb, int64_t m, const int32_t *__restrict__ first,
                                            const int32_t *__restrict__ count, const int32_t *__restrict__ bscan) {
    if (u < nb) return 1;
    const int32_t f = first[u];
    const int64_t e = (int64_t)f + count[u];
    return (e < m ? (int64_t)bscan[e] : nb) - bscan[f];
}

struct CutNode {
    const int32_t *first, *count, *bscan, *parent;
    const int64_t *sz;  // device counts: sz[1] buckets
    int64_t m;
    __device__ bool operator()(int32_t u) const {
        const int64_t nb = sz[1];
        if (bucket_n(u, nb, m, first, count, bscan) > kCut) return false;
        const int32_t p = parent[u];
        return p < 0 || bucket_n(p, nb, m, first, count, bscan) > kCut;
    }
};

// The cut nodes (CutNode) among the nn = sz[1] + sz[2] nodes, appended to
// cut with one atomic per work-group and round on *ncut (zero on entry);
// their order is arbitrary, each one is processed on its own.  Grid-stride
// over the device count, so the launch need not know it.
__global__ __launch_bounds__(256) void k_cut_list(CutNode pred, const int64_t *__restrict__ sz,
                                                  int32_t *__restrict__ cut, unsigned long long *__restrict__ ncut) {
    __shared__ int wcnt[4];
    __shared__ unsigned long long base;
    const int64_t nn = sz[1] + sz[2];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int64_t b0 = (int64_t)blockIdx.x * blockDim.x; b0 < nn; b0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t u = b0 + threadIdx.x;
        const bool hit = u < nn && pred((int32_t)u);
        const uint64_t mk = __ballot(hit);
        if (lane == 0) wcnt[w] = __popcll(mk);
        __syncthreads();
        if (threadIdx.x == 0) {
            const int tot = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
            base = tot ? atomicAdd(ncut, (unsigned long long)tot) : 0ull;
        }
        __syncthreads();
        if (hit) {
            unsigned long long k = base + __popcll(mk & ((1ull << lane) - 1));
            for (int q = 0; q < w; ++q) k += wcnt[q];
            cut[k] = (int32_t)u;
        }
        __syncthreads();  // wcnt and base are rewritten by the next round
    }
}

// Box union and M2M of node p from its two children (both complete), by a
// group of ORDER lanes (lane k writes coefficient k).  Radius: the smaller of
// two bounds on the particles' distance from the centre, the half diagonal
// and max over the children of |c - c_C| + R_C; both keep |delta|, rho <= 1.
//   b^P_k = sum over the two children of sum_{m<=k} C(k,m) delta^{k-m} w_m,
//   w_m = rho^m b^C_m,  delta = (c_C - c_P)/R_P,  rho = R_C/R_P,
// as a Horner sum in delta over m.  w_m is broadcast through the group's LDS
// slots ws[0, ORDER); bin[m][k] = C(k, m) (0 for m > k).  p < 0: no node.
template <int ORDER>
__device__ __forceinline__ void m2m_coop(int32_t p, int k, const int32_t *__restrict__ child0,
                                         const int32_t *__restrict__ child1, double *__restrict__ ncx,
                                         double *__restrict__ ncy, double *__restrict__ nr,
                                         double4 *__restrict__ nbb, double2 *__restrict__ cd,
                                         float2 *__restrict__ cf, double2 *ws, const double (*bin)[ORDER]) {
    const bool on = p >= 0;
    const int32_t c[2] = {on ? child0[p] : 0, on ? child1[p] : 0};
    double4 b = make_double4(0.0, 0.0, 0.0, 0.0);
    double cx = 0.0, cy = 0.0, R = 0.0, Rs = 1.0;
    if (on) {
        const double4 b0 = nbb[c[0]], b1 = nbb[c[1]];
        b = make_double4(fmin(b0.x, b1.x), fmax(b0.y, b1.y), fmin(b0.z, b1.z), fmax(b0.w, b1.w));
        cx = 0.5 * (b.x + b.y);
        cy = 0.5 * (b.z + b.w);
        double Rc = 0.0;
        for (int ch = 0; ch < 2; ++ch) {
            const double ex = ncx[c[ch]] - cx, ey = ncy[c[ch]] - cy;
            Rc = fmax(Rc, sqrt(ex * ex + ey * ey) + nr[c[ch]]);
        }
        R = fmin(0.5 * sqrt((b.y - b.x) * (b.y - b.x) + (b.w - b.z) * (b.w - b.z)), Rc);
        Rs = fmax(R, kTiny);
    }
    double accr = 0.0, acci = 0.0;
#pragma unroll 1
    for (int ch = 0; ch < 2; ++ch) {
        double dr = 0.0, di = 0.0, wr = 0.0, wi = 0.0;
        if (on) {
            const int32_t cc = c[ch];
            dr = (ncx[cc] - cx) / Rs;
            di = (ncy[cc] - cy) / Rs;
            const double rho = nr[cc] / Rs;
            double rk = 1.0, rb = rho;  // rho^k by binary powers
#pragma unroll
            for (int e = 0; (1 << e) < ORDER; ++e) {
                if ((k >> e) & 1) rk *= rb;
                rb *= rb;
            }
            const double2 bm = cd[(int64_t)cc * ORDER + k];
            wr = bm.x * rk;
            wi = bm.y * rk;
        }
        __builtin_amdgcn_wave_barrier();
        ws[k] = make_double2(wr, wi);
        __builtin_amdgcn_wave_barrier();
        double ar = 0.0, ai = 0.0;
#pragma unroll
        for (int m = 0; m < ORDER; ++m) {
            const double2 w = ws[m];
            const double cb = bin[m][k];
            const double nar = ar * dr - ai * di + cb * w.x;
            const double nai = ar * di + ai * dr + cb * w.y;
            ar = m <= k ? nar : ar;
            ai = m <= k ? nai : ai;
        }
        accr += ar;
        acci += ai;
    }
    __builtin_amdgcn_wave_barrier();
    if (!on) return;
    cd[(int64_t)p * ORDER + k] = make_double2(accr, acci);
    cf[(int64_t)p * ORDER + k] = make_float2((float)accr, (float)acci);
    if (k == 0) {
        ncx[p] = cx;
        ncy[p] = cy;
        nr[p] = R;
        nbb[p] = b;
    }
}

// Builds the internal nodes below root r (r included) that desc(u) accepts,
// with NW waves: the nodes are listed by depth (a breadth-first walk into
// accepted children, into lst), then built deepest level first, a node per
// group of ORDER lanes, with a work-group barrier between levels.  Node
// depths in the radix tree are below 64 (each level lengthens the common
// prefix of the 64-bit index-extended keys), so kLevCap level starts suffice.
constexpr int kLevCap = 72;

template <int ORDER, int NW, typename Desc>
__device__ void build_levels(int32_t r, Desc desc, int32_t *lst, int32_t *lev, int *ntot,
                             const int32_t *__restrict__ child0, const int32_t *__restrict__ child1,
                             double *__restrict__ ncx, double *__restrict__ ncy, double *__restrict__ nr,
                             double4 *__restrict__ nbb, double2 *__restrict__ cd, float2 *__restrict__ cf,
                             double2 *ws, const double (*bin)[ORDER]) {
    constexpr int G = 64 / ORDER;  // nodes per wave
    const int wave = threadIdx.x >> 6, gi = (threadIdx.x & 63) / ORDER, k = threadIdx.x % ORDER;
    __syncthreads();
    if (threadIdx.x == 0) {
        lst[0] = r;
        lev[0] = 0;
        lev[1] = 1;
        *ntot = 1;
    }
    __syncthreads();
    int nlev = 1;
    for (int d = 0; d + 2 < kLevCap; ++d) {
        const int s0 = lev[d], s1 = lev[d + 1];
        for (int t = s0 + threadIdx.x; t < s1; t += NW * 64) {
            const int32_t u = lst[t];
            const int32_t ch[2] = {child0[u], child1[u]};
            for (int q = 0; q < 2; ++q)
                if (desc(ch[q])) lst[atomicAdd(ntot, 1)] = ch[q];
        }
        __syncthreads();
        const int e = *ntot;
        if (e == s1) break;
        if (threadIdx.x == 0) lev[d + 2] = e;
        nlev = d + 2;
        __syncthreads();
    }
    for (int d = nlev - 1; d >= 0; --d) {
        const int s0 = lev[d], s1 = lev[d + 1];
        for (int base = s0 + wave * G; base < s1; base += NW * G) {
            const int idx = base + gi;
            m2m_coop<ORDER>(idx < s1 ? lst[idx] : -1, k, child0, child1, ncx, ncy, nr, nbb, cd, cf, ws, bin);
        }
        __syncthreads();
    }
}

template <int ORDER>
__device__ void fill_binomials(double (*bin)[ORDER]) {
    for (int t = threadIdx.x; t < ORDER * ORDER; t += blockDim.x) {
        const int m = t / ORDER, k = t % ORDER;
        bin[m][k] = m <= k ? binom(k, m) : 0.0;
    }
}

struct IsInternal {
    int64_t nb;
    __device__ bool operator()(int32_t u) const { return u >= nb; }
};

// A cut subtree's internal nodes held in LDS while it is built (slot = list
// index): children (an internal child's slot, or -1 - the id of a bucket,
// read from global memory), geometry and multipole coefficients.  About
// 86 KB at 16 coefficients.
template <int ORDER>
struct SubLds {
    int32_t lst[kCut];
    int32_t ch[kCut][2];
    int32_t lev[kLevCap];
    int ntot;
    double gx[kCut], gy[kCut], gr[kCut];
    double4 gb[kCut];
    double2 c[kCut][ORDER];
};

// m2m_coop with the children read from LDS (internal) or global memory
// (buckets) and the result kept in LDS slot idx; the same arithmetic.
template <int ORDER>
__device__ __forceinline__ void m2m_lds(SubLds<ORDER> &L, int idx, int k, const double *__restrict__ ncx,
                                        const double *__restrict__ ncy, const double *__restrict__ nr,
                                        const double4 *__restrict__ nbb, const double2 *__restrict__ cd,
                                        double2 *ws, const double (*bin)[ORDER]) {
    const bool on = idx >= 0;
    double ccx[2] = {0.0, 0.0}, ccy[2] = {0.0, 0.0}, crr[2] = {0.0, 0.0};
    double4 cb[2] = {make_double4(0.0, 0.0, 0.0, 0.0), make_double4(0.0, 0.0, 0.0, 0.0)};
    int32_t cs[2] = {0, 0};
    if (on) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            cs[q] = L.ch[idx][q];
            if (cs[q] >= 0) {
                ccx[q] = L.gx[cs[q]];
                ccy[q] = L.gy[cs[q]];
                crr[q] = L.gr[cs[q]];
                cb[q] = L.gb[cs[q]];
            } else {
                const int32_t bq = -1 - cs[q];
                ccx[q] = ncx[bq];
                ccy[q] = ncy[bq];
                crr[q] = nr[bq];
                cb[q] = nbb[bq];
            }
        }
    }
    double4 b = make_double4(0.0, 0.0, 0.0, 0.0);
    double cx = 0.0, cy = 0.0, R = 0.0, Rs = 1.0;
    if (on) {
        b = make_double4(fmin(cb[0].x, cb[1].x), fmax(cb[0].y, cb[1].y), fmin(cb[0].z, cb[1].z),
                         fmax(cb[0].w, cb[1].w));
        cx = 0.5 * (b.x + b.y);
        cy = 0.5 * (b.z + b.w);
        double Rc = 0.0;
        for (int ch = 0; ch < 2; ++ch) {
            const double ex = ccx[ch] - cx, ey = ccy[ch] - cy;
            Rc = fmax(Rc, sqrt(ex * ex + ey * ey) + crr[ch]);
        }
        R = fmin(0.5 * sqrt((b.y - b.x) * (b.y - b.x) + (b.w - b.z) * (b.w - b.z)), Rc);
        Rs = fmax(R, kTiny);
    }
    double accr = 0.0, acci = 0.0;
#pragma unroll 1
    for (int ch = 0; ch < 2; ++ch) {
        double dr = 0.0, di = 0.0, wr = 0.0, wi = 0.0;
        if (on) {
            dr = (ccx[ch] - cx) / Rs;
            di = (ccy[ch] - cy) / Rs;
            const double rho = crr[ch] / Rs;
            double rk = 1.0, rb = rho;  // rho^k by binary powers
#pragma unroll
            for (int e = 0; (1 << e) < ORDER; ++e) {
                if ((k >> e) & 1) rk *= rb;
                rb *= rb;
            }
            const double2 bm = cs[ch] >= 0 ? L.c[cs[ch]][k] : cd[(int64_t)(-1 - cs[ch]) * ORDER + k];
            wr = bm.x * rk;
            wi = bm.y * rk;
        }
        __builtin_amdgcn_wave_barrier();
        ws[k] = make_double2(wr, wi);
        __builtin_amdgcn_wave_barrier();
        double ar = 0.0, ai = 0.0;
#pragma unroll
        for (int m = 0; m < ORDER; ++m) {
            const double2 w = ws[m];
            const double cbm = bin[m][k];
            const double nar = ar * dr - ai * di + cbm * w.x;
            const double nai = ar * di + ai * dr + cbm * w.y;
            ar = m <= k ? nar : ar;
            ai = m <= k ? nai : ai;
        }
        accr += ar;
        acci += ai;
    }
    __builtin_amdgcn_wave_barrier();
    if (!on) return;
    L.c[idx][k] = make_double2(accr, acci);
    if (k == 0) {
        L.gx[idx] = cx;
        L.gy[idx] = cy;
        L.gr[idx] = R;
        L.gb[idx] = b;
    }
}

// Builds the nodes below root r that desc accepts (r included) in LDS with
// NW waves: listed by depth, built deepest level first, a node per group of
// ORDER lanes, then written out.  Children that desc does not accept come
// from global memory.  False (and nothing built) if more than kCut - 1
// nodes would be listed.
template <int ORDER, int NW, typename Desc>
__device__ bool build_sub_lds(int32_t r, Desc desc, SubLds<ORDER> &L, const int32_t *__restrict__ child0,
                              const int32_t *__restrict__ child1, double *__restrict__ ncx,
                              double *__restrict__ ncy, double *__restrict__ nr, double4 *__restrict__ nbb,
                              double2 *__restrict__ cd, float2 *__restrict__ cf, double2 *ws,
                              const double (*bin)[ORDER]) {
    constexpr int G = 64 / ORDER;  // nodes per wave
    const int wave = threadIdx.x >> 6, gi = (threadIdx.x & 63) / ORDER, k = threadIdx.x % ORDER;
    __shared__ int full;
    __syncthreads();
    if (threadIdx.x == 0) {
        L.lst[0] = r;
        L.lev[0] = 0;
        L.lev[1] = 1;
        L.ntot = 1;
        full = 0;
    }
    __syncthreads();
    int nlev = 1;
    for (int d = 0; d + 2 < kLevCap; ++d) {
        const int s0 = L.lev[d], s1 = L.lev[d + 1];
        for (int t = s0 + threadIdx.x; t < s1; t += NW * 64) {
            const int32_t u = L.lst[t];
            const int32_t chq[2] = {child0[u], child1[u]};
            for (int q = 0; q < 2; ++q) {
                if (desc(chq[q])) {
                    const int pos = atomicAdd(&L.ntot, 1);
                    if (pos < kCut - 1) {
                        L.lst[pos] = chq[q];
                        L.ch[t][q] = pos;
                    } else {
                        full = 1;
                    }
                } else {
                    L.ch[t][q] = -1 - chq[q];
                }
            }
        }
        __syncthreads();
        const int e = L.ntot;
        if (full || e == s1) break;
        if (threadIdx.x == 0) L.lev[d + 2] = e;
        nlev = d + 2;
        __syncthreads();
    }
    if (full) return false;  // the same for every thread
    for (int d = nlev - 1; d >= 0; --d) {
        const int s0 = L.lev[d], s1 = L.lev[d + 1];
        for (int base = s0 + wave * G; base < s1; base += NW * G) {
            const int idx = base + gi;
            m2m_lds<ORDER>(L, idx < s1 ? idx : -1, k, ncx, ncy, nr, nbb, cd, ws, bin);
        }
        __syncthreads();
    }
    const int tot = L.ntot;
    for (int t = threadIdx.x / ORDER; t < tot; t += NW * 64 / ORDER) {
        const int32_t u = L.lst[t];
        const double2 c = L.c[t][k];
        cd[(int64_t)u * ORDER + k] = c;
        cf[(int64_t)u * ORDER + k] = make_float2((float)c.x, (float)c.y);
        if (k == 0) {
            ncx[u] = L.gx[t];
            ncy[u] = L.gy[t];
            nr[u] = L.gr[t];
            nbb[u] = L.gb[t];
        }
    }
    return true;
}

// k_upper_local: a work-group per cut node builds its subtree (at most
// kCut - 1 internal nodes).  At 16 coefficients the subtree stays in LDS
// until it is complete (SubLds), so a level waits on LDS rather than on
// global memory; at 32 it is built through global memory (build_levels).
template <int ORDER>
__global__ __launch_bounds__(256) void k_upper_local(const int32_t *__restrict__ cut,
                                                     const unsigned long long *__restrict__ ncut,
                                                     int64_t nb, const int32_t *__restrict__ child0,
                                                     const int32_t *__restrict__ child1, double *__restrict__ ncx,
                                                     double *__restrict__ ncy, double *__restrict__ nr,
                                                     double4 *__restrict__ nbb, double2 *__restrict__ cd,
                                                     float2 *__restrict__ cf) {
    __shared__ double2 ws_all[4][64];
    __shared__ double bin[ORDER][ORDER];
    fill_binomials<ORDER>(bin);
    double2 *ws = ws_all[threadIdx.x >> 6] + ((threadIdx.x & 63) / ORDER) * ORDER;
    const int n = (int)*ncut;
    if constexpr (ORDER <= 16) {
        __shared__ SubLds<ORDER> L;
        for (int w = blockIdx.x; w < n; w += gridDim.x) {
            const int32_t r = cut[w];
            if (r < nb) continue;  // a bucket: nothing to build (the same for every thread)
            (void)build_sub_lds<ORDER, 4>(r, IsInternal{nb}, L, child0, child1, ncx, ncy, nr, nbb, cd, cf, ws, bin);
        }
    } else {
        __shared__ int32_t lst[kCut];
        __shared__ int32_t lev[kLevCap];
        __shared__ int ntot;
        for (int w = blockIdx.x; w < n; w += gridDim.x) {
            const int32_t r = cut[w];
            if (r < nb) continue;  // a bucket: nothing to build (the same for every thread)
            build_levels<ORDER, 4>(r, IsInternal{nb}, lst, lev, &ntot, child0, child1, ncx, ncy, nr, nbb, cd, cf,
                                   ws, bin);
        }
    }
}

// The nodes above the cut nodes: internal nodes holding more than kCut buckets.
struct IsTop {
    const int32_t *first, *count, *bscan;
    int64_t nb, m;
    __device__ bool operator()(int32_t u) const { return u >= nb && bucket_n(u, nb, m, first, count, bscan) > kCut; }
};

// k_upper_top: one work-group builds the nodes above the cut nodes (fewer
// than nb, listed in lst).
template <int ORDER>
__global__ __launch_bounds__(1024) void k_upper_top(int32_t root, int64_t nb, int64_t m,
                                                    const int32_t *__restrict__ first,
                                                    const int32_t *__restrict__ count,
                                                    const int32_t *__restrict__ bscan, int32_t *__restrict__ lst,
                                                    const int32_t *__restrict__ child0,
                                                    const int32_t *__restrict__ child1, double *__restrict__ ncx,
                                                    double *__restrict__ ncy, double *__restrict__ nr,
                                                    double4 *__restrict__ nbb, double2 *__restrict__ cd,
                                                    float2 *__restrict__ cf) {
    __shared__ int32_t lev[kLevCap];
    __shared__ int ntot;
    __shared__ double2 ws_all[16][64];
    __shared__ double bin[ORDER][ORDER];
    const IsTop top{first, count, bscan, nb, m};
    if (!top(root)) return;  // the root is a bucket or a cut node (the same for every thread)
    fill_binomials<ORDER>(bin);
    double2 *ws = ws_all[threadIdx.x >> 6] + ((threadIdx.x & 63) / ORDER) * ORDER;
    if constexpr (ORDER <= 16) {
        // in LDS when the nodes above the cut nodes fit (a few hundred at 2^20)
        __shared__ SubLds<ORDER> L;
        if (build_sub_lds<ORDER, 16>(root, top, L, child0, child1, ncx, ncy, nr, nbb, cd, cf, ws, bin)) return;
    }
    build_levels<ORDER, 16>(root, top, lst, lev, &ntot, child0, child1, ncx, ncy, nr, nbb, cd, cf, ws, bin);
}

// Group count, bucket count and upper-node count from the ends of their scans.
__global__ void k_tree_counts(const int32_t *__restrict__ gscan, const int32_t *__restrict__ gflag,
                              const int32_t *__restrict__ bscan, const int32_t *__restrict__ bstart,
                              const int32_t *__restrict__ uscan, const int32_t *__restrict__ upper, int64_t m, int big,
                              const int32_t *__restrict__ sort_err, int64_t *__restrict__ out,
                              int64_t *__restrict__ hout) {
    out[0] = (int64_t)gscan[m - 1] + gflag[m - 1];
    out[1] = big ? (int64_t)bscan[m - 1] + bstart[m - 1] : 1;
    out[2] = big ? (int64_t)uscan[m - 2] + upper[m - 2] : 0;
    out[3] = *sort_err;
    // and straight to the host (pinned, coherent; read after an event)
    for (int q = 0; q < 4; ++q) __hip_atomic_store(&hout[q], out[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __threadfence_system();
}

__global__ void k_pack_nodes(int64_t nn, const double *__restrict__ ncx, const double *__restrict__ ncy,
                             const double *__restrict__ nr, const int32_t *__restrict__ child0,
                             const int32_t *__restrict__ child1, const int32_t *__restrict__ first,
                             const int32_t *__restrict__ count, NodeRec *__restrict__ rec) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nn) return;
    NodeRec q;
    q.cx = ncx[u];
    q.cy = ncy[u];
    q.r = nr[u];
    const int32_t c0 = child0[u];
    q.a = c0 >= 0 ? c0 : -1 - first[u];
    q.b = c0 >= 0 ? child1[u] : count[u];
    rec[u] = q;
}

// ---- traversal helpers ------------------------------------------------------
// The tree above the buckets is at most 63 levels deep (each level lengthens
// the common prefix of the 64-bit keys extended by the index, and the index
// fits 31 bits), so a depth-first walk holds at most 64 entries; a push
// beyond 64 is dropped and counted as an error.
struct LaneStack {
    int slot = 0;  // this lane's stack entry
    int sp = 0;    // wave-uniform depth
    bool overflow = false;
    __device__ __forceinline__ void push(int lane, int node) {
        if (sp >= 64) {
            overflow = true;
            return;
        }
        if (lane == sp) slot = node;
        ++sp;
    }
    __device__ __forceinline__ int pop() {
        --sp;
        return __builtin_amdgcn_readlane(slot, sp);
    }
};

template <typename T>
struct PairMath;
template <>
struct PairMath<double> {
    __device__ static __forceinline__ double div(double a, double b) { return a / b; }
    __device__ static __forceinline__ double sqrt_bound(double a) { return sqrt(a); }
    __device__ static __forceinline__ double med3(double a, double b, double c) {
        return fmax(fmin(a, b), fmin(fmax(a, b), c));
    }
    __device__ static __forceinline__ double min_(double a, double b) { return fmin(a, b); }
};
template <>
struct PairMath<float> {
    __device__ static __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    // the hardware square root (1 ulp), for bounds that carry a margin
    __device__ static __forceinline__ float sqrt_bound(float a) { return __builtin_amdgcn_sqrtf(a); }
    __device__ static __forceinline__ float med3(float a, float b, float c) { return __builtin_amdgcn_fmed3f(a, b, c); }
    // min of two non-NaN values as one median (fminf adds a canonicalisation)
    __device__ static __forceinline__ float min_(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, -3.0e38f); }
};

// One source leaf staged into this wave's LDS slot: one coalesced load per
// lane, then broadcast reads in the pair loop.  The slot is private to the
// wave, whose LDS operations complete in order, so a scheduling barrier (no
// work-group barrier) separates the write from the reads and the reads from
// the next write.
template <typename T>
struct Src4T {
    T x, y, g, pad;
};

// Slots past a leaf's sources hold padding: no circulation, and far enough
// from any target to stay out of the four smallest distances (first bound
// 1e5).  FP32 slots are relative to their bucket's centre and the targets
// that read them lie within a few bucket radii of it: at (1000, 1000), r^2
// is about 2e6, and FP32 products of two squared distances stay finite.
// FP64 slots hold absolute positions: at (1e30, 1e30).  Waves that give a
// target several lanes read whole shares of the slot, padding included.
template <typename T>
__device__ __forceinline__ Src4T<T> pad_src() {
    Src4T<T> q;
    q.x = q.y = sizeof(T) == 4 ? T(1000) : T(1e30);
    q.g = q.pad = T(0);
    return q;
}

template <typename T>
__device__ __forceinline__ void stage_leaf(Src4T<T> *lds, int lane, int64_t j0, int cnt, const TreeDev &t) {
    __builtin_amdgcn_wave_barrier();
    Src4T<T> q = pad_src<T>();
    if (lane < cnt) {
        if constexpr (sizeof(T) == 4) {
            q.x = t.fx[j0 + lane];
            q.y = t.fy[j0 + lane];
            q.g = t.fg[j0 + lane];
        } else {
            q.x = t.sx[j0 + lane];
            q.y = t.sy[j0 + lane];
            q.g = t.sg[j0 + lane];
        }
    }
    lds[lane] = q;
    __builtin_amdgcn_wave_barrier();
}

// One source leaf staged like stage_leaf, keeping only the sources within
// sqrt(R2) of the box [bx0, bx1] x [by0, by1] (in the staged coordinates),
// packed to the front of the slot in leaf order, padding after them;
// returns their count.
template <typename T>
__device__ __forceinline__ int stage_leaf_near(Src4T<T> *lds, int lane, int64_t j0, int cnt, const TreeDev &t, T bx0,
                                               T bx1, T by0, T by1, T R2) {
    __builtin_amdgcn_wave_barrier();
    Src4T<T> q;
    bool keep = false;
    if (lane < cnt) {
        if constexpr (sizeof(T) == 4) {
            q.x = t.fx[j0 + lane];
            q.y = t.fy[j0 + lane];
            q.g = t.fg[j0 + lane];
        } else {
            q.x = t.sx[j0 + lane];
            q.y = t.sy[j0 + lane];
            q.g = t.sg[j0 + lane];
        }
        q.pad = T(0);
        const T ex = q.x < bx0 ? bx0 - q.x : (q.x > bx1 ? q.x - bx1 : T(0));
        const T ey = q.y < by0 ? by0 - q.y : (q.y > by1 ? q.y - by1 : T(0));
        keep = ex * ex + ey * ey < R2;
    }
    const uint64_t m = __ballot(keep);
    const int kept = (int)__popcll(m);
    if (keep) lds[__popcll(m & ((1ull << lane) - 1ull))] = q;
    if (lane >= kept) lds[lane] = pad_src<T>();
    __builtin_amdgcn_wave_barrier();
    return kept;
}

// v into the ascending four smallest m0 <= m1 <= m2 <= m3.
template <typename T>
__device__ __forceinline__ void near4_insert(T v, T &m0, T &m1, T &m2, T &m3) {
    using PM = PairMath<T>;
    m3 = PM::med3(m2, m3, v);
    m2 = PM::med3(m1, m2, v);
    m1 = PM::med3(m0, m1, v);
    m0 = PM::min_(m0, v);
}

// Pair loop over one staged leaf: Biot-Savart sums and, with TRACK, the four
// smallest squared distances (the target itself included).  Without CLAMP
// the weight is g / r2: the caller has shown r2 >= eps2 for every pair.
template <typename T, bool TRACK, bool CLAMP = true>
__device__ __forceinline__ void p2p_leaf(const Src4T<T> *lds, int cnt, T ox, T oy, T eps2, T &pu, T &pv, T &m0,
                                         T &m1, T &m2, T &m3) {
    using PM = PairMath<T>;
    int jj = 0;
    if constexpr (sizeof(T) == 4) {
        // two sources at a time with one quarter-rate reciprocal:
        // 1/a = b / (a b), 1/b = a / (a b).  For a target lane, a, b lie in
        // [eps^2, 4e6] (the included vortices lie within 400 of the domain,
        // padding slots about 1400 from the bucket centre), so the product
        // neither overflows nor underflows in FP32; lanes without a target
        // may produce inf or NaN, which are never stored.
        for (; jj + 1 < cnt; jj += 2) {
            const Src4T<T> q = lds[jj], p = lds[jj + 1];
            const T dx = ox - q.x, dy = oy - q.y, ex = ox - p.x, ey = oy - p.y;
            const T r2 = dx * dx + dy * dy, s2 = ex * ex + ey * ey;
            if constexpr (TRACK) {
                m3 = PM::med3(m2, m3, r2);
                m2 = PM::med3(m1, m2, r2);
                m1 = PM::med3(m0, m1, r2);
                m0 = PM::min_(m0, r2);
                m3 = PM::med3(m2, m3, s2);
                m2 = PM::med3(m1, m2, s2);
                m1 = PM::med3(m0, m1, s2);
                m0 = PM::min_(m0, s2);
            }
            const T a = CLAMP ? fmax(r2, eps2) : r2, b = CLAMP ? fmax(s2, eps2) : s2;
            const T rab = __builtin_amdgcn_rcpf(a * b);
            const T w = q.g * (b * rab), v = p.g * (a * rab);
            pv += w * dx;
            pu -= w * dy;
            pv += v * ex;
            pu -= v * ey;
        }
    }
    for (; jj < cnt; ++jj) {
        const Src4T<T> q = lds[jj];
        const T dx = ox - q.x, dy = oy - q.y;
        const T r2 = dx * dx + dy * dy;
        if constexpr (TRACK) {
            m3 = PM::med3(m2, m3, r2);
            m2 = PM::med3(m1, m2, r2);
            m1 = PM::med3(m0, m1, r2);
            m0 = PM::min_(m0, r2);
        }
        const T w = PM::div(q.g, CLAMP ? fmax(r2, eps2) : r2);
        pv += w * dx;
        pu -= w * dy;
    }
}

// ---- target groups ------------------------------------------------------------
// The targets of one wave: within each key-prefix cell of the bucket depth
// (a cell holds about four times 64 particles at the mean density), runs of
// 64 consecutive particles of the Morton order, so that nearly every lane has
// a target and a group never straddles a jump of the curve between cells.
// Box, centre and radius by wave reductions.
__global__ void k_group_cells(const uint32_t *__restrict__ key, int64_t m, int shift, int32_t *__restrict__ cs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t c = shift >= 32 ? 0u : key[i] >> shift;
    const bool start = i == 0 || (shift >= 32 ? 0u : key[i - 1] >> shift) != c;
    cs[i] = start ? (int32_t)i : 0;
}

// cs2 = running maximum of cs: the start of each particle's cell.
__global__ void k_group_flags(const int32_t *__restrict__ cs2, int64_t m, int32_t *__restrict__ gflag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) gflag[i] = ((i - cs2[i]) & 63) == 0 ? 1 : 0;
}

__global__ void k_group_first(const int32_t *__restrict__ gflag, const int32_t *__restrict__ gscan, int64_t m,
                              int32_t *__restrict__ gfirst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m && gflag[i]) gfirst[gscan[i]] = (int32_t)i;
}

// Lanes per target of a wave that holds nvalid targets: the largest power of
// two S with nvalid S <= 64.  Target k is held by lanes k + s 64 / S
// (s = 0 .. S-1), and lane share s sums the s-th of S equal runs of every
// staged source slot, so a wave of few targets takes S times fewer pair
// iterations than with one lane per target.
struct LaneSplit {
    int per;  // S
    int tpw;  // targets per wave, 64 / S
    int tl;   // this lane's target
    int sub;  // this lane's share
};

__device__ __forceinline__ LaneSplit lane_split(int nvalid, int lane) {
    LaneSplit L;
    int lg = 0;
    while (lg < 6 && nvalid << (lg + 1) <= 64) ++lg;
    L.per = 1 << lg;
    L.tpw = 64 >> lg;
    L.tl = lane & (L.tpw - 1);
    L.sub = lane >> (6 - lg);
    return L;
}

// The lane's share of a slot of cnt staged sources: ceil(cnt / S) entries
// from share * that (entries past cnt are padding).
__device__ __forceinline__ int share_len(const LaneSplit &L, int cnt) { return (cnt + L.per - 1) / L.per; }

struct TargetGroup {
    int64_t ti;
    bool ok;
    int nvalid;
    LaneSplit ls;
    double cx, cy, r;
    double4 box;
};

__device__ __forceinline__ TargetGroup target_group(const TreeDev &t, int64_t grp, int lane) {
    TargetGroup G;
    const int64_t f = t.gfirst[grp];
    const int64_t e = grp + 1 < t.ngroup ? (int64_t)t.gfirst[grp + 1] : t.m;
    G.nvalid = (int)(e - f);
    G.ls = lane_split(G.nvalid, lane);
    G.ti = f + G.ls.tl;
    G.ok = G.ls.tl < G.nvalid;
    const double xi = G.ok ? t.sx[G.ti] : 0.0, yi = G.ok ? t.sy[G.ti] : 0.0;
    const double x0 = wave_min(G.ok ? xi : 1e300), x1 = wave_max(G.ok ? xi : -1e300);
    const double y0 = wave_min(G.ok ? yi : 1e300), y1 = wave_max(G.ok ? yi : -1e300);
    G.box = make_double4(x0, x1, y0, y1);
    G.cx = 0.5 * (x0 + x1);
    G.cy = 0.5 * (y0 + y1);
    G.r = 0.5 * sqrt((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0));
    return G;
}

// Per-wave counts (the values held by lane 0 of each wave) summed over the
// work-group in LDS, then added once to the global counters: a counter that
// every wave of a launch updates serialises at one L2 address, and the
// launch does not end before the queue drains.  Called by every thread.
__device__ __forceinline__ void block_count_add(unsigned long long a, unsigned long long b,
                                                unsigned long long *ga, unsigned long long *gb) {
    __shared__ unsigned long long s[2];
    if (threadIdx.x == 0) s[0] = s[1] = 0ull;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        if (a) atomicAdd(&s[0], a);
        if (b) atomicAdd(&s[1], b);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s[0]) atomicAdd(ga, s[0]);
        if (gb && s[1]) atomicAdd(gb, s[1]);
    }
}

__device__ __forceinline__ float wave_sum_f(float a) {
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    return a;
}

__device__ __forceinline__ float readlane_f(float a, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a), l));
}

// M2L (FP32) on the matrix cores: far source nodes translated into the
// target bucket's local expansion about its centre c_T (radius R_T):
//   F(z) = sum_j g_j / (z - z_j) = sum_l L_l u^l,  u = (z - c_T) / R_T,
//   L_l += q (-beta)^l T_l,  T_l = sum_k C(k + l, l) a_k,  a_k = alpha^k b_k,
//   q = 1 / (c_T - c_S),  alpha = R_S q,  beta = R_T q,
// with b_k the node's multipole (about c_S, scaled by R_S).  It follows from
// 1/(D + w)^(k+1) = sum_l C(k + l, l) (-w)^l / D^(k+l+1), D = c_T - c_S,
// w = z - c_T.  The opening test keeps (R_S + R_T) / |D| below theta, so the
// truncation at P terms on both sides leaves about theta^P of the node's
// field.
//
// The contraction over k is a product with the fixed binomial matrix,
// T (P x 16 nodes) = B (P x P) A (P x 16 nodes), run as
// v_mfma_f32_16x16x4_f32 (operand A: lane i + 16 k holds A[i][k]; B: lane
// j + 16 k holds B[k][j]; D: lane j + 16 (i / 4), register i % 4 holds
// D[i][j]) over 16 nodes at a time, real and imaginary parts of a_k as two
// products.  Lane j + 16 g works on node j of the 16: it forms a_k for
// k = g mod 4 (the B operand of each 4-deep step), and receives T_l for
// l = 16 mt + 4 g + r (r = 0..3), which it weights by q (-beta)^l and adds
// to acc[mt][r].  The lanes' sums over nodes are combined once, by
// l2p_coeffs().  Rows l >= P of the binomial operand are zero.  gap takes
// the lane's least |c_T - c_S| - R_S - R_T.
template <int P>
struct M2LBinom {
    float v[P][P];  // v[k][l] = C(k + l, l)
};
template <int P>
constexpr M2LBinom<P> m2l_binom() {
    M2LBinom<P> b{};
    for (int k = 0; k < P; ++k)
        for (int l = 0; l < P; ++l) b.v[k][l] = (float)binom(k + l, l);
    return b;
}
__device__ const M2LBinom<16> kM2LBin16 = m2l_binom<16>();
__device__ const M2LBinom<20> kM2LBin20 = m2l_binom<20>();
__device__ const M2LBinom<32> kM2LBin32 = m2l_binom<32>();
template <int P>
__device__ __forceinline__ const M2LBinom<P> &m2l_bin();
template <>
__device__ __forceinline__ const M2LBinom<16> &m2l_bin<16>() { return kM2LBin16; }
template <>
__device__ __forceinline__ const M2LBinom<20> &m2l_bin<20>() { return kM2LBin20; }
template <>
__device__ __forceinline__ const M2LBinom<32> &m2l_bin<32>() { return kM2LBin32; }

typedef float m2l_f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <int P>
struct M2LShape {
    static constexpr int KS = P / 4;          // 4-deep steps over k
    static constexpr int MT = (P + 15) / 16;  // 16-row tiles over l
};

template <int P, int STRIDE>
__device__ __forceinline__ void m2l_mfma(const TreeDev &t, const int32_t *fq, int count, int lane, double cTx,
                                         double cTy, float rT, float2 (&acc)[M2LShape<P>::MT][4], float &gap) {
    constexpr int KS = M2LShape<P>::KS, MT = M2LShape<P>::MT;
    static_assert(P % 4 == 0 && P <= STRIDE, "terms: a multiple of 4 within the stored coefficients");
    const int j = lane & 15, g = lane >> 4;
    // binomial operand of this lane: A[i][kk] = C(k + l, l), k = 4 m + kk, l = 16 mt + i
    float ab[MT][KS];
    {
        const M2LBinom<P> &B = m2l_bin<P>();
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int m = 0; m < KS; ++m) {
                const int l = 16 * mt + j;
                ab[mt][m] = l < P ? B.v[4 * m + g][l] : 0.f;
            }
    }
    for (int base = 0; base < count; base += 16) {
        const int node = base + j < count ? fq[base + j] : -1;
        float2 q = make_float2(0.f, 0.f), al = q, nb = q;
        if (node >= 0) {
            const NodeRec S = t.rec[node];
            const float dx = (float)(cTx - S.cx), dy = (float)(cTy - S.cy);
            const float d2 = dx * dx + dy * dy;
            const float inv = __builtin_amdgcn_rcpf(d2);
            q = make_float2(dx * inv, -dy * inv);
            const float rs = fmaxf((float)S.r, 1e-30f);
            // nearest possible particle of the node: |c_T - c_S| - R_S - R_T
            gap = fminf(gap, __builtin_amdgcn_sqrtf(d2) - rs - rT);
            al = make_float2(rs * q.x, rs * q.y);
            nb = make_float2(-rT * q.x, -rT * q.y);
        }
        // a_k = alpha^k b_k for k = g, g + 4, ...
        const float2 a2 = cmulf(al, al), a3 = cmulf(a2, al), a4 = cmulf(a2, a2);
        float2 pw = g == 0 ? make_float2(1.f, 0.f) : (g == 1 ? al : (g == 2 ? a2 : a3));
        m2l_f4 Tr[MT], Ti[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            Tr[mt] = m2l_f4{0.f, 0.f, 0.f, 0.f};
            Ti[mt] = m2l_f4{0.f, 0.f, 0.f, 0.f};
        }
        const float2 *b = t.cf + (int64_t)(node >= 0 ? node : 0) * STRIDE + g;
#pragma unroll
        for (int m = 0; m < KS; ++m) {
            const float2 bk = node >= 0 ? b[4 * m] : make_float2(0.f, 0.f);
            const float2 ak = cmulf(pw, bk);
            pw = cmulf(pw, a4);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                Tr[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(ab[mt][m], ak.x, Tr[mt], 0, 0, 0);
                Ti[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(ab[mt][m], ak.y, Ti[mt], 0, 0, 0);
            }
        }
        // weights q (-beta)^l, l = 16 mt + 4 g + r
        const float2 n2 = cmulf(nb, nb), n4 = cmulf(n2, n2), n8 = cmulf(n4, n4), n12 = cmulf(n8, n4);
        float2 w = cmulf(q, g == 0 ? make_float2(1.f, 0.f) : (g == 1 ? n4 : (g == 2 ? n8 : n12)));
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float tr_ = Tr[mt][r], ti_ = Ti[mt][r];
                acc[mt][r].x += w.x * tr_ - w.y * ti_;
                acc[mt][r].y += w.x * ti_ + w.y * tr_;
                w = cmulf(w, nb);
            }
            if (mt + 1 < MT) w = cmulf(w, n12);  // from (-beta)^(16 mt + 4 g + 4) to the next tile's start
        }
    }
}

// The local coefficients: acc summed over the 16 lanes of each group g, so
// that every lane of group g holds L_l for l = 16 mt + 4 g + r in acc[mt][r].
template <int P>
__device__ __forceinline__ void l2p_coeffs(float2 (&acc)[M2LShape<P>::MT][4]) {
#pragma unroll
    for (int mt = 0; mt < M2LShape<P>::MT; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                acc[mt][r].x += __shfl_xor(acc[mt][r].x, o, 64);
                acc[mt][r].y += __shfl_xor(acc[mt][r].y, o, 64);
            }
}

// L_l of the combined coefficients, broadcast from lane 16 g.
template <int P>
__device__ __forceinline__ float2 local_coeff(const float2 (&acc)[M2LShape<P>::MT][4], int l) {
    const int mt = l >> 4, g = (l >> 2) & 3, r = l & 3;
    return make_float2(readlane_f(acc[mt][r].x, 16 * g), readlane_f(acc[mt][r].y, 16 * g));
}

// ---- Biot-Savart + nearest-3 --------------------------------------------------
// Work-groups are dispatched to the 8 XCDs in turn (work-group b to XCD b % 8).
// Each XCD gets a contiguous eighth of the Morton-ordered buckets instead, so
// the source buckets its waves re-read stay within one region and its L2.  The
// grid is padded to a multiple of 8 work-groups.
__device__ __forceinline__ int64_t xcd_block() {
    const int64_t q = (int64_t)(gridDim.x >> 3);
    return (int64_t)(blockIdx.x & 7u) * q + (int64_t)(blockIdx.x >> 3);
}

constexpr int kBsStack = 256;  // traversal stack entries per wave
constexpr int kFarQ = 512;     // FP32: far nodes queued per wave before any M2L

template <typename T, int ORDER, int STRIDE>
__device__ __forceinline__ void bs_wave(const TreeDev &t, double theta, double eps, double vinf_x, double vinf_y,
                                        double *__restrict__ u, double *__restrict__ v, double *__restrict__ d,
                                        double *__restrict__ sdo,
                                        int32_t *__restrict__ fb, double *__restrict__ fbr2,
                                        unsigned long long *__restrict__ nfb, int exact_all,
                                        unsigned long long &cpairs, unsigned long long &cfar,
                                        unsigned long long &cm2l) {
    using PM = PairMath<T>;
    constexpr bool F32 = sizeof(T) == 4;
    __shared__ Src4T<T> lds_all[4][kLeaf];
    __shared__ int32_t fq_all[4][kFarQ];  // FP32: far nodes waiting for M2L
    __shared__ int32_t stk_all[4][kBsStack];
    __shared__ int32_t lq_all[4][64];     // near leaves of the popped nodes
    const int lane = threadIdx.x & 63;
    Src4T<T> *lds = lds_all[threadIdx.x >> 6];
    int32_t *fq = fq_all[threadIdx.x >> 6];
    // targets: one bucket per wave (a compact node, which keeps the MAC's
    // opening radius small; runs of 64 particles of the Morton order fill
    // more lanes but open about 1.6 times as many pairs)
    const int leaf = __builtin_amdgcn_readfirstlane((int)(xcd_block() * 4 + (threadIdx.x >> 6)));
    if (leaf >= t.nleaf) return;
    const NodeRec L = t.rec[leaf];
    const int nvalid = L.b;
    // a bucket of at most 32 particles gives each target several lanes
    const LaneSplit ls = lane_split(nvalid, lane);
    const int64_t ti = (int64_t)(-1 - L.a) + ls.tl;
    const bool ok = ls.tl < nvalid;
    const double cAx = L.cx, cAy = L.cy, rA = L.r;
    T ax, ay;
    if constexpr (F32) {
        ax = ok ? t.fx[ti] : 0.f;  // relative to the bucket centre
        ay = ok ? t.fy[ti] : 0.f;
    } else {
        ax = ok ? t.sx[ti] : cAx;
        ay = ok ? t.sy[ti] : cAy;
    }
    const T eps2 = (T)(eps * eps);
    const double th2 = theta * theta;
    int64_t npairs = 0, nfar = 0;
    double su = 0.0, sv = 0.0;
    T fu = 0, fv = 0;  // far-field sums, folded into su, sv at each near bucket
    T m0 = T(1e5), m1 = T(1e5), m2 = T(1e5), m3 = T(1e5);
    T farmin = T(1e30);         // FP64: per target, nearest point of a far node
    double farmin_w = 1e30;     // FP32: per bucket, |c_T - c_S| - R_S - R_T
    float gap = 1e30f;          // FP32: this lane's share of it (m2l_mfma)
    int nq = 0;                 // FP32: far nodes queued
    float2 lacc[M2LShape<ORDER>::MT][4];  // FP32: local coefficients, partial (m2l_mfma)
#pragma unroll
    for (int mt = 0; mt < M2LShape<ORDER>::MT; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) lacc[mt][r] = make_float2(0.f, 0.f);
    const float rT = fmaxf((float)rA, 1e-30f);  // FP32: kTiny would round to zero
    // the own bucket first: its particles are usually the nearest, which
    // tightens the bound that lets most other leaves skip the nearest-3 updates
    {
        stage_leaf<T>(lds, lane, -1 - L.a, nvalid, t);
        T pu = 0, pv = 0;
        const int h = share_len(ls, nvalid);
        p2p_leaf<T, true>(lds + ls.sub * h, h, ax, ay, eps2, pu, pv, m0, m1, m2, m3);
        __builtin_amdgcn_wave_barrier();
        npairs += nvalid;
        su += (double)pu;
        sv += (double)pv;
    }
    // Traversal: a stack in the wave's LDS slot, popped up to 64 nodes at a
    // time, one per lane, which tests its node's MAC; far nodes are queued
    // (FP32: for M2L; FP64: evaluated one at a time), near leaves summed one
    // at a time, opened nodes' children pushed.  While the stack is within
    // 64 + 64 entries of its capacity, nodes are popped one at a time: a
    // depth-first descent adds at most one entry per level (the tree is at
    // most 63 levels deep), so the stack never outgrows kBsStack.
    int32_t *stk = stk_all[threadIdx.x >> 6];
    int32_t *lq = lq_all[threadIdx.x >> 6];
    int sp = 1;
    if (lane == 0) stk[0] = t.root;
    while (sp > 0 || (F32 && nq > 0)) {
        if constexpr (F32) {
            // queued far nodes go to M2L into the bucket's local expansion
            // after the traversal, 64 at a time; only a nearly full queue
            // (rare: a bucket meets about 150 far nodes at 2^20) is
            // translated before it ends
            if (sp == 0 || nq > kFarQ - 64) {
                const int nb_ = nq < 64 ? nq : 64;
                __builtin_amdgcn_wave_barrier();
                m2l_mfma<ORDER, STRIDE>(t, fq + nq - nb_, nb_, lane, cAx, cAy, rT, lacc, gap);
                __builtin_amdgcn_wave_barrier();
                cm2l += (unsigned long long)nb_;
                nq -= nb_;
                continue;
            }
        }
        const int k = sp + 64 + 64 <= kBsStack ? (sp < 64 ? sp : 64) : 1;
        __builtin_amdgcn_wave_barrier();
        const int node = lane < k ? stk[sp - 1 - lane] : -1;
        __builtin_amdgcn_wave_barrier();
        sp -= k;
        bool isfar = false, isleaf = false, isopen = false;
        int ca = 0, cb = 0;
        if (node >= 0 && node != leaf) {
            const NodeRec nd = t.rec[node];
            const double ddx = nd.cx - cAx, ddy = nd.cy - cAy;
            const double d2 = ddx * ddx + ddy * ddy;
            const double rs = nd.r + rA;
            // dist * theta > r_node + r_leaf and dist - r_node - r_leaf > eps
            isfar = d2 * th2 > rs * rs && d2 > (rs + eps) * (rs + eps);
            isleaf = !isfar && nd.a < 0;
            isopen = !isfar && !isleaf;
            ca = nd.a;
            cb = nd.b;
        }
        // children of the opened nodes, the first child on top
        const uint64_t om = __ballot(isopen);
        if (isopen) {
            const int pos = sp + 2 * __popcll(om & ((1ull << lane) - 1));
            stk[pos] = cb;
            stk[pos + 1] = ca;
        }
        sp += 2 * __popcll(om);
        // near leaves listed in LDS (summed below, one at a time)
        const uint64_t lm = __ballot(isleaf);
        if (isleaf) lq[__popcll(lm & ((1ull << lane) - 1))] = node;
        const int nl = __popcll(lm);
        const uint64_t fm = __ballot(isfar);
        nfar += __popcll(fm);
        if constexpr (F32) {
            if (isfar) fq[nq + __popcll(fm & ((1ull << lane) - 1))] = node;
            nq += __popcll(fm);
        } else {
            // M2P: S = (1/D) sum_k b_k (R/D)^k, D = z - c_node; u += Im S, v += Re S
            auto m2p = [&](int fnode) {
                const NodeRec nd = t.rec[fnode];
                const T Dx = ax - nd.cx, Dy = ay - nd.cy;
                const T den = Dx * Dx + Dy * Dy;
                const T inv = PM::div(T(1), den);
                const T qx = Dx * inv, qy = -Dy * inv;
                const T R = (T)fmax(nd.r, kTiny);
                const T tx = R * qx, ty = R * qy;
                const double2 *bm = t.cd + (int64_t)fnode * STRIDE;
                T sr = bm[ORDER - 1].x, si = bm[ORDER - 1].y;
#pragma unroll
                for (int q = ORDER - 2; q >= 0; --q) {
                    const T nr_ = sr * tx - si * ty + bm[q].x;
                    si = sr * ty + si * tx + bm[q].y;
                    sr = nr_;
                }
                fu += sr * qy + si * qx;
                fv += sr * qx - si * qy;
                farmin = fmin(farmin, PM::sqrt_bound(den) - (T)nd.r);
            };
            if (ls.per == 1) {
                for (uint64_t m = fm; m; m &= m - 1) m2p(__builtin_amdgcn_readlane(node, __builtin_ctzll(m)));
            } else {
                // a target's lanes take the far nodes in turn
                if (isfar) fq[__popcll(fm & ((1ull << lane) - 1))] = node;
                __builtin_amdgcn_wave_barrier();
                const int nf = __popcll(fm);
                for (int fi = ls.sub; fi < nf; fi += ls.per) m2p(fq[fi]);
            }
        }
        __builtin_amdgcn_wave_barrier();
        for (int li = 0; li < nl; ++li) {
            const int lnode = __builtin_amdgcn_readfirstlane(lq[li]);
            const NodeRec nd = t.rec[lnode];
            const double ddx = nd.cx - cAx, ddy = nd.cy - cAy;
            const int64_t j0 = -1 - nd.a;
            const int cnt = nd.b;
            T ox, oy;
            if constexpr (F32) {
                ox = ax - (float)ddx;  // target relative to the source bucket centre
                oy = ay - (float)ddy;
            } else {
                ox = ax;
                oy = ay;
            }
            stage_leaf<T>(lds, lane, j0, cnt, t);
            T pu = 0, pv = 0;
            npairs += cnt;
            const int h = share_len(ls, cnt);
            const Src4T<T> *sh = lds + ls.sub * h;
            // a pair at r2 >= m3 leaves the four smallest unchanged: skip the
            // updates when the leaf's circle lies beyond sqrt(m3) of every
            // target (a 1e-3 margin covers the FP32 distances)
            // likewise the cap max(r2, eps2) only matters for a leaf within
            // eps of some target
            const T cx_ = F32 ? ox : ax - (T)nd.cx, cy_ = F32 ? oy : ay - (T)nd.cy;
            const T dc2 = cx_ * cx_ + cy_ * cy_;
            const T reach = (T)nd.r + PM::sqrt_bound(m3), reache = (T)nd.r + (T)eps;
            const bool near3 = ok && !(dc2 > reach * reach * T(1.002));
            const bool neare = ok && !(dc2 > reache * reache * T(1.002));
            if (__ballot(near3))
                p2p_leaf<T, true>(sh, h, ox, oy, eps2, pu, pv, m0, m1, m2, m3);
            else if (__ballot(neare))
                p2p_leaf<T, false>(sh, h, ox, oy, eps2, pu, pv, m0, m1, m2, m3);
            else
                p2p_leaf<T, false, false>(sh, h, ox, oy, eps2, pu, pv, m0, m1, m2, m3);
            __builtin_amdgcn_wave_barrier();
            su += (double)pu + (double)fu;
            sv += (double)pv + (double)fv;
            fu = fv = 0;
        }
    }
    su += (double)fu;
    sv += (double)fv;
    // the target's shares: sums added in the same order on each of its
    // lanes, the four smallest distances merged, the far bound the least
    for (int o = ls.tpw; o < 64; o <<= 1) {
        su += __shfl_xor(su, o, 64);
        sv += __shfl_xor(sv, o, 64);
        const T q0 = __shfl_xor(m0, o, 64), q1 = __shfl_xor(m1, o, 64);
        const T q2 = __shfl_xor(m2, o, 64), q3 = __shfl_xor(m3, o, 64);
        near4_insert<T>(q0, m0, m1, m2, m3);
        near4_insert<T>(q1, m0, m1, m2, m3);
        near4_insert<T>(q2, m0, m1, m2, m3);
        near4_insert<T>(q3, m0, m1, m2, m3);
        if constexpr (!F32) farmin = fmin(farmin, __shfl_xor(farmin, o, 64));
    }
    if constexpr (F32) {
        for (int o = 32; o > 0; o >>= 1) gap = fminf(gap, __shfl_xor(gap, o, 64));
        farmin_w = (double)gap;
        // L2P: F = sum_l L_l u^l at u = (z - c_T) / R_T (ax, ay are relative
        // to the bucket centre); u += Im F, v += Re F
        l2p_coeffs<ORDER>(lacc);
        const float ur = ax / rT, ui = ay / rT;
        const float2 Lt = local_coeff<ORDER>(lacc, ORDER - 1);
        float Fr = Lt.x, Fi = Lt.y;
#pragma unroll
        for (int l = ORDER - 2; l >= 0; --l) {
            const float2 Ll = local_coeff<ORDER>(lacc, l);
            const float nr_ = Fr * ur - Fi * ui + Ll.x;
            Fi = Fr * ui + Fi * ur + Ll.y;
            Fr = nr_;
        }
        su += (double)Fi;
        sv += (double)Fr;
    }
    cpairs = (unsigned long long)(npairs * nvalid);
    cfar = (unsigned long long)(nfar * nvalid);
    if (!ok || ls.sub != 0) return;
    const int32_t i = t.perm[ti];
    u[i] = su / (2 * kPi) + vinf_x;
    v[i] = sv / (2 * kPi) + vinf_y;
    const double di = (sqrt((double)m1) + sqrt((double)m2) + sqrt((double)m3)) / 3;
    d[i] = di;
    sdo[ti] = di;  // in sorted order too, for the diffusion's far-target selection
    // exact only if no far node can hold a closer particle than the third
    // FP32 squared distances of close pairs carry up to ~1e-5 relative error,
    // hence the 1e-3 margins here and in the search bound
    const double fm = F32 ? farmin_w : (double)farmin;
    if (exact_all || (double)m3 >= fm * fm * (1.0 - 1e-3)) {
        const unsigned long long k = atomicAdd(nfb, 1ull);
        fb[k] = (int32_t)ti;
        fbr2[k] = (double)m3;
    }
}

// ORDER terms of coefficients stored STRIDE per node.
template <typename T, int ORDER, int STRIDE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && ORDER == 16 ? 6 : 1)))
void k_tree_bs(TreeDev t, double theta, double eps, double vinf_x, double vinf_y,
                                                 double *__restrict__ u, double *__restrict__ v,
                                                 double *__restrict__ d, double *__restrict__ sdo,
                                                 int32_t *__restrict__ fb, double *__restrict__ fbr2,
                                                 unsigned long long *__restrict__ nfb, int exact_all) {
    unsigned long long cp = 0, cf = 0, cl = 0;
    bs_wave<T, ORDER, STRIDE>(t, theta, eps, vinf_x, vinf_y, u, v, d, sdo, fb, fbr2, nfb, exact_all, cp, cf, cl);
    block_count_add(cp, cf, nfb + 1, nfb + 2);
    block_count_add(cl, 0ull, nfb + 7, nullptr);
}

// Exact nearest-3 for the flagged targets, FP64: the four smallest squared
// distances over all included particles (the target itself included), one
// 1024-thread work-group per target.  The bound starts at the fast pass's
// third distance (its three candidates are real particles), widened by 1e-3.
// The cut nodes partition the buckets: those whose box lies within the bound
// are listed in LDS.  The nearest one's buckets (a contiguous range) are
// scanned first, and the bound tightened to the fourth smallest found; then
// each wave takes listed cut nodes within the bound, tightening its own copy
// after each.  A bucket within the bound is read one particle per lane into
// the lane's private four smallest; the lanes' lists are merged at the end.
// Should the list overflow (or the tree be a single bucket), every other
// bucket is tested instead, each wave tightening its bound as it goes.
__device__ __forceinline__ double box_d2(const double4 &b, double ax, double ay) {
    const double ex = fmax(0.0, fmax(b.x - ax, ax - b.y));
    const double ey = fmax(0.0, fmax(b.z - ay, ay - b.w));
    return ex * ex + ey * ey;
}

__device__ __forceinline__ void near4_insert(double r2, double &m0, double &m1, double &m2, double &m3) {
    m3 = PairMath<double>::med3(m2, m3, r2);
    m2 = PairMath<double>::med3(m1, m2, r2);
    m1 = PairMath<double>::med3(m0, m1, r2);
    m0 = fmin(m0, r2);
}

// The four smallest of the lanes' sorted lists: four rounds of a wave
// minimum, the lowest lane holding it dropping its smallest each time.
__device__ __forceinline__ void near4_merge(int lane, double m0, double m1, double m2, double m3, double r[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double w = wave_min(m0);
        r[k] = w;
        const uint64_t hm = __ballot(m0 == w);
        if (lane == __builtin_ctzll(hm)) {
            m0 = m1;
            m1 = m2;
            m2 = m3;
            m3 = 1e5;
        }
    }
}

// The buckets in the wave's ballot `lm` (node ids in the lanes' `node`), four
// at a time, one particle per lane (a bucket holds at most 64): loads first,
// then the updates.
__device__ __forceinline__ void near4_buckets(const TreeDev &t, uint64_t lm, int node, int lane, double ax,
                                              double ay, double &m0, double &m1, double &m2, double &m3) {
    while (lm) {
        double px[4], py[4];
        bool pk[4];
#pragma unroll
        for (int u4 = 0; u4 < 4; ++u4) {
            pk[u4] = false;
            px[u4] = py[u4] = 0.0;
            if (lm) {
                const int nl = __builtin_amdgcn_readlane(node, __builtin_ctzll(lm));
                lm &= lm - 1;
                const int64_t j0 = t.first[nl];
                pk[u4] = lane < t.count[nl];
                if (pk[u4]) {
                    px[u4] = t.sx[j0 + lane];
                    py[u4] = t.sy[j0 + lane];
                }
            }
        }
#pragma unroll
        for (int u4 = 0; u4 < 4; ++u4) {
            if (pk[u4]) {
                const double dx = ax - px[u4], dy = ay - py[u4];
                near4_insert(dx * dx + dy * dy, m0, m1, m2, m3);
            }
        }
    }
}

constexpr int kFbList = 1024;

// The four smallest of the work-group's lanes (all 16 waves), in every
// thread's r; `wr` is LDS scratch, and a barrier ends the call.
__device__ __forceinline__ void near4_block(int lane, int w, double m0, double m1, double m2, double m3,
                                            double (*wr)[4], double r[4]) {
    near4_merge(lane, m0, m1, m2, m3, r);
    if (lane < 4) wr[w][lane] = lane == 0 ? r[0] : lane == 1 ? r[1] : lane == 2 ? r[2] : r[3];
    __syncthreads();
    near4_merge(lane, wr[lane >> 2][lane & 3], 1e5, 1e5, 1e5, r);
    __syncthreads();
}

__global__ __launch_bounds__(1024) void k_tree_nearest_fallback(TreeDev t, const int32_t *__restrict__ cut,
                                                                const unsigned long long *__restrict__ ncut,
                                                                const int32_t *__restrict__ fb,
                                                                const double *__restrict__ fbr2,
                                                                const unsigned long long *__restrict__ nfb,
                                                                double *__restrict__ d, double *__restrict__ sdo) {
    __shared__ int32_t lst[kFbList];
    __shared__ double lsd[kFbList];
    __shared__ int nl;
    __shared__ unsigned long long best;
    __shared__ double wr[16][4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t nq = (int64_t)*nfb;
    const int nc = t.nleaf == 1 ? 0 : (int)*ncut;  // one bucket: no bucket scan, no cut list
    for (int64_t q = blockIdx.x; q < nq; q += gridDim.x) {
        const int64_t ti = fb[q];
        const double ax = t.sx[ti], ay = t.sy[ti];
        double lim = fmin(fbr2[q] * (1.0 + 1e-3) + 1e-300, 1e5);
        if (threadIdx.x == 0) {
            nl = 0;
            best = ~0ull;
        }
        __syncthreads();
        // the cut nodes within the bound, and the nearest one (box distance
        // with its low mantissa bits replaced by the cut-list index: a
        // non-negative double orders as its bit pattern; any choice is exact)
        for (int k = threadIdx.x; k < nc; k += 1024) {
            const int32_t u = cut[k];
            const double b2 = box_d2(t.nbb[u], ax, ay);
            if (b2 <= lim) {
                const int p = atomicAdd(&nl, 1);
                if (p < kFbList) {
                    lst[p] = u;
                    lsd[p] = b2;
                }
                atomicMin(&best, ((unsigned long long)__double_as_longlong(b2) & ~0xffffffull) |
                                     (unsigned long long)(k & 0xffffff));
            }
        }
        __syncthreads();
        const int L = nl;
        double m0 = 1e5, m1 = 1e5, m2 = 1e5, m3 = 1e5, r[4];
        // the nearest cut node first, its buckets spread over the waves (16
        // each); its four smallest tighten the bound
        int32_t us = -1;
        int64_t ub0 = 0, ub1 = 0;
        if (t.nleaf > 1 && best != ~0ull) {
            us = cut[best & 0xffffffull];
            const int64_t f = t.first[us], e1 = f + t.count[us];
            ub0 = t.bscan[f];
            ub1 = e1 < t.m ? (int64_t)t.bscan[e1] : t.nleaf;
            for (int64_t bb = ub0 + w * 16; bb < ub1; bb += 256) {
                const int64_t b = bb + lane;
                const bool in = lane < 16 && b < ub1 && box_d2(t.nbb[b], ax, ay) <= lim;
                near4_buckets(t, __ballot(in), (int)b, lane, ax, ay, m0, m1, m2, m3);
            }
            near4_block(lane, w, m0, m1, m2, m3, wr, r);
            lim = fmin(lim, r[3]);
        }
        if (L <= kFbList && t.nleaf > 1) {
            // the others within the bound, one cut node per wave; each wave
            // tightens its own copy of the bound to the fourth smallest of
            // its lanes (four real distances bound the fourth smallest)
            for (int e = w; e < L; e += 16) {
                if (lst[e] == us || lsd[e] > lim) continue;
                const int32_t u = lst[e];
                const int64_t f = t.first[u], e1 = f + t.count[u];
                const int64_t b0 = t.bscan[f], b1 = e1 < t.m ? (int64_t)t.bscan[e1] : t.nleaf;
                uint64_t any = 0;
                for (int64_t bb = b0; bb < b1; bb += 64) {
                    const int64_t b = bb + lane;
                    const uint64_t bm = __ballot(b < b1 && box_d2(t.nbb[b], ax, ay) <= lim);
                    any |= bm;
                    near4_buckets(t, bm, (int)b, lane, ax, ay, m0, m1, m2, m3);
                }
                if (any) {
                    near4_merge(lane, m0, m1, m2, m3, r);
                    lim = fmin(lim, r[3]);
                }
            }
        } else {
            // every bucket outside the nearest cut node's range
            for (int64_t bb = (int64_t)w * 64; bb < t.nleaf; bb += 1024) {
                const int64_t b = bb + lane;
                const uint64_t bm = __ballot(b < t.nleaf && (b < ub0 || b >= ub1) && box_d2(t.nbb[b], ax, ay) <= lim);
                near4_buckets(t, bm, (int)b, lane, ax, ay, m0, m1, m2, m3);
                if (bm) {
                    near4_merge(lane, m0, m1, m2, m3, r);
                    lim = fmin(lim, r[3]);
                }
            }
        }
        near4_block(lane, w, m0, m1, m2, m3, wr, r);
        if (threadIdx.x == 0) {
            const double di = (sqrt(r[1]) + sqrt(r[2]) + sqrt(r[3])) / 3;
            d[t.perm[ti]] = di;
            sdo[ti] = di;
        }
    }
}

// ---- diffusion --------------------------------------------------------------
// Diffusion pair terms (src/kernel.cu:1011-1019): s1 += g e, (s2x, s2y) +=
// g e (dx, dy) / r with e = exp(-r / d), for r above the threshold.  FP32:
// kd = -log2(e) / d and a below-threshold r^2 is pushed to 1e30, whose
// weight underflows to zero; FP64: kd = 1 / d.
template <typename T>
struct DiffPair;
template <>
struct DiffPair<float> {
    template <bool GUARD>
    __device__ static __forceinline__ void add(float dx, float dy, float g, float kd, float thr2, float &s2x,
                                               float &s2y, float &s1) {
        const float r2 = dx * dx + dy * dy;
        const float r2c = (!GUARD || r2 > thr2) ? r2 : 1e30f;
        const float rs = __builtin_amdgcn_rsqf(r2c);
        const float e = __builtin_amdgcn_exp2f(r2c * rs * kd);
        const float ge = g * e;
        const float w = ge * rs;
        s2x += w * dx;
        s2y += w * dy;
        s1 += ge;
    }
};
// FP32 with the circulation folded into the exponent: lg = log2|g| of a
// source of known sign, h = 2^(r kd + lg) = |g| e^{-r/d}; NEG subtracts.
// One multiply less per pair than g * e; the weight carries the rounding of
// log2|g| (about 1e-6 relative).
template <bool GUARD, bool NEG>
__device__ __forceinline__ void diff_pair_log(float dx, float dy, float lg, float kd, float thr2, float &s2x,
                                              float &s2y, float &s1) {
    const float r2 = dx * dx + dy * dy;
    const float r2c = (!GUARD || r2 > thr2) ? r2 : 1e30f;
    const float rs = __builtin_amdgcn_rsqf(r2c);
    const float h = __builtin_amdgcn_exp2f(__builtin_fmaf(r2c * rs, kd, lg));
    const float w = h * rs;
    if constexpr (NEG) {
        s1 -= h;
        s2x -= w * dx;
        s2y -= w * dy;
    } else {
        s1 += h;
        s2x += w * dx;
        s2y += w * dy;
    }
}

template <>
struct DiffPair<double> {
    template <bool GUARD>
    __device__ static __forceinline__ void add(double dx, double dy, double g, double kd, double thr,
                                               double &s2x, double &s2y, double &s1) {
        const double dist = sqrt(dx * dx + dy * dy);
        if (!GUARD || dist > thr) {
            const double w = g / dist;
            const double e = exp(-dist * kd);
            s2x += w * dx * e;
            s2y += w * dy * e;
            s1 += g * e;
        }
    }
};

// Radius beyond which a diffusion target is summed by k_diff_far: 5 % of
// the root's radius.  At most t.fcap targets are listed (the most of kFarCap,
// one in 16 included vortices and a quarter more than the last diffusion
// found: sparse wakes have many); further ones stay in their group's wave.
constexpr int64_t kFarCap = 4096;
// A target goes to k_diff_far when kFarCut d (or its cut-off times d, if
// smaller) exceeds far_radius(): the same targets for both cut-offs, so the
// FP64 path's 30 d widens some group waves rather than listing a large part
// of a sparse wake.
constexpr double kFarCut = 18.0;
__device__ __forceinline__ double far_radius(const TreeDev &t) {
    const double4 b = t.nbb[t.root];
    const double rr = 0.5 * sqrt((b.y - b.x) * (b.y - b.x) + (b.w - b.z) * (b.w - b.z));
    return fmax(0.05 * rr, 1e-300);
}

__device__ __forceinline__ int64_t far_count(const unsigned long long *cnt, int64_t cap) {
    const unsigned long long h = cnt[2];
    return (int64_t)(h < (unsigned long long)cap ? h : (unsigned long long)cap);
}

// Targets are summed over the sources within cut * d_i of them (the omitted
// weight is bounded by (1 + cut) e^{-cut} of the near sum).  One wave per
// target group: a node is visited when the box around the group within the
// largest radius of its lanes reaches it and, lane by lane, when some lane's
// own radius reaches it.  A lane whose radius exceeds far_radius() is listed for
// k_diff_far instead and does not widen the wave's radius.  A group of at
// most 32 targets gives each several lanes (LaneSplit).  cnt[0] +=
// evaluated pairs, cnt[2] the list length.
template <typename T, bool CENSUS>
__device__ __forceinline__ void diffusion_wave(const TreeDev &t, double cut, double nu, double thr,
                                               const double *__restrict__ sd, double *__restrict__ u,
                                               double *__restrict__ v, unsigned long long *__restrict__ cnt,
                                               unsigned long long &cpairs, unsigned long long &cin) {
    constexpr bool F32 = sizeof(T) == 4;
    __shared__ Src4T<T> lds_all[4][kLeaf];
    __shared__ int32_t stk_all[4][kBsStack];
    __shared__ double4 lbox_all[4][64];  // the popped leaves: box,
    __shared__ double2 lcen_all[4][64];  // centre,
    __shared__ int2 lrng_all[4][64];     // first source and count,
    __shared__ int32_t lsel_all[4][64];  // those summed (index | guard << 8)
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    Src4T<T> *lds = lds_all[wid];
    int32_t *stk = stk_all[wid];
    double4 *lbox = lbox_all[wid];
    double2 *lcen = lcen_all[wid];
    int2 *lrng = lrng_all[wid];
    int32_t *lsel = lsel_all[wid];
    const uint64_t below = (1ull << lane) - 1;
    const int64_t grp = __builtin_amdgcn_readfirstlane((int)(xcd_block() * 4 + wid));
    if (grp >= t.ngroup) return;
    const TargetGroup G = target_group(t, grp, lane);
    const int64_t ti = G.ti;
    const bool ok = G.ok;
    const int32_t i = ok ? t.perm[ti] : 0;
    const double sdi = ok ? sd[ti] : 0.0;  // negative: listed for k_diff_far (d > 0 there)
    const double di = fabs(sdi);
    const bool mine = ok && !(sdi < 0.0);
    const double R = cut * wave_max(mine ? di : 0.0);
    const double R2 = R * R;
    const double cAx = G.cx, cAy = G.cy;
    const double4 bA = G.box;
    T ax, ay, kd, ri2;
    if constexpr (F32) {
        ax = ok ? (float)(t.sx[ti] - cAx) : 0.f;  // target relative to the group centre
        ay = ok ? (float)(t.sy[ti] - cAy) : 0.f;
        kd = (float)(-1.4426950408889634 / (ok ? di : 1.0));
        ri2 = mine ? (float)(cut * di * cut * di * (1.0 + 1e-4)) : -1.f;
    } else {
        ax = ok ? t.sx[ti] : cAx;
        ay = ok ? t.sy[ti] : cAy;
        kd = 1.0 / (ok ? di : 1.0);
        ri2 = mine ? cut * di * cut * di : -1.0;
    }
    const T thr2 = F32 ? (T)(thr * thr) : (T)thr;
    const T gap2 = (T)(4.0 * thr * thr);
    // sources farther than R from the group's box reach no target's radius
    const T Rs2 = F32 ? (T)(R2 * (1.0 + 1e-4)) : (T)R2;
    int64_t npairs = 0, nin = 0;
    double a2x = 0.0, a2y = 0.0, a1 = 0.0;
    // Traversal as in bs_wave: a stack in the wave's LDS slot popped up to 64
    // nodes at a time, one per lane, each tested against the group's box
    // widened by R.  The leaves found are listed with their box, centre and
    // source range; those that some lane's own radius reaches are summed one
    // at a time, each leaf's sources loaded while the previous one is summed.
    int sp = R > 0.0 ? 1 : 0;
    if (lane == 0) stk[0] = t.root;
    while (sp > 0) {
        const int k = sp + 64 + 64 <= kBsStack ? (sp < 64 ? sp : 64) : 1;
        __builtin_amdgcn_wave_barrier();
        const int node = lane < k ? stk[sp - 1 - lane] : -1;
        __builtin_amdgcn_wave_barrier();
        sp -= k;
        bool isleaf = false, isopen = false;
        NodeRec nd{};
        double4 b{};
        if (node >= 0) {
            b = t.nbb[node];
            const double ex = fmax(0.0, fmax(b.x - bA.y, bA.x - b.y));
            const double ey = fmax(0.0, fmax(b.z - bA.w, bA.z - b.w));
            if (ex * ex + ey * ey < R2) {
                nd = t.rec[node];
                isleaf = nd.a < 0;
                isopen = !isleaf;
            }
        }
        // children of the opened nodes, the first child on top
        const uint64_t om = __ballot(isopen);
        if (isopen) {
            const int pos = sp + 2 * __popcll(om & below);
            stk[pos] = nd.b;
            stk[pos + 1] = nd.a;
        }
        sp += 2 * __popcll(om);
        const uint64_t lm = __ballot(isleaf);
        if (isleaf) {
            const int q = __popcll(lm & below);
            lbox[q] = b;
            lcen[q] = make_double2(nd.cx, nd.cy);
            lrng[q] = make_int2(-1 - nd.a, nd.b);
        }
        const int nl = __popcll(lm);
        __builtin_amdgcn_wave_barrier();
        // the leaves some lane's own radius reaches, in list order
        int nr = 0;
        for (int li = 0; li < nl; ++li) {
            const double4 lb = lbox[li];
            T lx, ly;
            if constexpr (F32) {
                const float x0 = (float)(lb.x - cAx), x1 = (float)(lb.y - cAx);
                const float y0 = (float)(lb.z - cAy), y1 = (float)(lb.w - cAy);
                lx = fmaxf(0.f, fmaxf(x0 - ax, ax - x1));
                ly = fmaxf(0.f, fmaxf(y0 - ay, ay - y1));
            } else {
                lx = fmax(0.0, fmax(lb.x - ax, ax - lb.y));
                ly = fmax(0.0, fmax(lb.z - ay, ay - lb.w));
            }
            const T l2 = lx * lx + ly * ly;
            if (__ballot(l2 < ri2) == 0) continue;
            // the threshold test is needed only where a pair can come within
            // it: a leaf box nearer than twice the threshold to some target
            const int guard = __ballot(mine && !(l2 > gap2)) != 0;
            if (lane == 0) lsel[nr] = li | (guard << 8);
            ++nr;
        }
        __builtin_amdgcn_wave_barrier();
        Src4T<T> qn = pad_src<T>();
        // every lane loads (lanes past the count re-read the first source,
        // which the filter drops): with no branch around them the loads stay
        // in flight through the pair loop
        auto fetch = [&](int r) {
            const int2 rg = lrng[lsel[r] & 255];
            const int64_t j = rg.x + (lane < rg.y ? lane : 0);
            Src4T<T> q;
            if constexpr (F32) {
                q.x = t.fx[j];
                q.y = t.fy[j];
                q.g = t.fg[j];
            } else {
                q.x = t.sx[j];
                q.y = t.sy[j];
                q.g = t.sg[j];
            }
            q.pad = T(0);
            return q;
        };
        if (nr > 0) qn = fetch(0);
        for (int r = 0; r < nr; ++r) {
            const int sel = lsel[r];
            const int li = sel & 255;
            const int cnt_l = lrng[li].y;
            T ox, oy, bx0, bx1, by0, by1;
            if constexpr (F32) {
                const double2 lc = lcen[li];
                ox = ax + (float)(cAx - lc.x);
                oy = ay + (float)(cAy - lc.y);
                bx0 = (float)(bA.x - lc.x);
                bx1 = (float)(bA.y - lc.x);
                by0 = (float)(bA.z - lc.y);
                by1 = (float)(bA.w - lc.y);
            } else {
                ox = ax;
                oy = ay;
                bx0 = bA.x;
                bx1 = bA.y;
                by0 = bA.z;
                by1 = bA.w;
            }
            // sources within sqrt(Rs2) of the group's box, packed to the
            // front of the slot, padding after them
            bool keep = false;
            if (lane < cnt_l) {
                const T ex = qn.x < bx0 ? bx0 - qn.x : (qn.x > bx1 ? qn.x - bx1 : T(0));
                const T ey = qn.y < by0 ? by0 - qn.y : (qn.y > by1 ? qn.y - by1 : T(0));
                keep = ex * ex + ey * ey < Rs2;
            }
            if constexpr (F32) {
                if (G.ls.per == 1) {
                    // one lane per target: positive sources first, then
                    // negative ones, each with log2|g| (no circulation: no
                    // term), summed by the folded pair term
                    const bool kp = keep && qn.g > 0.f, kn = keep && qn.g < 0.f;
                    const uint64_t mp = __ballot(kp), mn = __ballot(kn);
                    const int np_ = (int)__popcll(mp), nn_ = (int)__popcll(mn);
                    Src4T<T> w = qn;
                    w.g = __builtin_amdgcn_logf(fabsf(qn.g));
                    __builtin_amdgcn_wave_barrier();
                    if (kp) lds[__popcll(mp & below)] = w;
                    if (kn) lds[np_ + __popcll(mn & below)] = w;
                    __builtin_amdgcn_wave_barrier();
                    if (r + 1 < nr) qn = fetch(r + 1);
                    const int ne = np_ + nn_;
                    npairs += ne;
                    if constexpr (CENSUS) {
                        if (mine)
                            for (int jj = 0; jj < ne; ++jj) {
                                const Src4T<T> q = lds[jj];
                                const T dx = ox - q.x, dy = oy - q.y;
                                nin += dx * dx + dy * dy < ri2 / (T)(1.0 + 1e-4) ? 1 : 0;
                            }
                    }
                    float s2x = 0.f, s2y = 0.f, s1 = 0.f;
                    if (sel >> 8) {
#pragma unroll 4
                        for (int jj = 0; jj < np_; ++jj) {
                            const Src4T<T> q = lds[jj];
                            diff_pair_log<true, false>(ox - q.x, oy - q.y, q.g, kd, thr2, s2x, s2y, s1);
                        }
#pragma unroll 4
                        for (int jj = np_; jj < ne; ++jj) {
                            const Src4T<T> q = lds[jj];
                            diff_pair_log<true, true>(ox - q.x, oy - q.y, q.g, kd, thr2, s2x, s2y, s1);
                        }
                    } else {
#pragma unroll 4
                        for (int jj = 0; jj < np_; ++jj) {
                            const Src4T<T> q = lds[jj];
                            diff_pair_log<false, false>(ox - q.x, oy - q.y, q.g, kd, thr2, s2x, s2y, s1);
                        }
#pragma unroll 4
                        for (int jj = np_; jj < ne; ++jj) {
                            const Src4T<T> q = lds[jj];
                            diff_pair_log<false, true>(ox - q.x, oy - q.y, q.g, kd, thr2, s2x, s2y, s1);
                        }
                    }
                    a2x += (double)s2x;
                    a2y += (double)s2y;
                    a1 += (double)s1;
                    continue;
                }
            }
            const uint64_t km = __ballot(keep);
            const int cnt_j = (int)__popcll(km);
            __builtin_amdgcn_wave_barrier();
            if (keep) lds[__popcll(km & below)] = qn;
            if (lane >= cnt_j) lds[lane] = pad_src<T>();
            __builtin_amdgcn_wave_barrier();
            if (r + 1 < nr) qn = fetch(r + 1);
            T s2x = 0, s2y = 0, s1 = 0;
            npairs += cnt_j;
            // this lane's share of the staged sources (padding adds nothing)
            const int h = share_len(G.ls, cnt_j);
            const Src4T<T> *sh = lds + G.ls.sub * h;
            if constexpr (CENSUS) {
                if (mine)
                    for (int jj = 0; jj < h; ++jj) {
                        const Src4T<T> q = sh[jj];
                        const T dx = ox - q.x, dy = oy - q.y;
                        nin += dx * dx + dy * dy < (F32 ? ri2 / (T)(1.0 + 1e-4) : ri2) ? 1 : 0;
                    }
            }
            if (sel >> 8) {
                for (int jj = 0; jj < h; ++jj) {
                    const Src4T<T> q = sh[jj];
                    DiffPair<T>::template add<true>(ox - q.x, oy - q.y, q.g, kd, thr2, s2x, s2y, s1);
                }
            } else {
                for (int jj = 0; jj < h; ++jj) {
                    const Src4T<T> q = sh[jj];
                    DiffPair<T>::template add<false>(ox - q.x, oy - q.y, q.g, kd, thr2, s2x, s2y, s1);
                }
            }
            a2x += (double)s2x;
            a2y += (double)s2y;
            a1 += (double)s1;
        }
        __builtin_amdgcn_wave_barrier();
    }
    const int nmine = __popcll(__ballot(mine && G.ls.sub == 0));
    cpairs = (unsigned long long)(npairs * nmine);
    if constexpr (CENSUS) {
        for (int o = 32; o > 0; o >>= 1) nin += __shfl_xor(nin, o, 64);
        cin = (unsigned long long)nin;
    }
    // the target's shares, added in the same order on each of its lanes
    for (int o = G.ls.tpw; o < 64; o <<= 1) {
        a2x += __shfl_xor(a2x, o, 64);
        a2y += __shfl_xor(a2y, o, 64);
        a1 += __shfl_xor(a1, o, 64);
    }
    if (!mine || G.ls.sub != 0) return;
    if (fabs(a1) > 1e-5) {
        u[i] += nu * a2x / (a1 * di);
        v[i] += nu * a2y / (a1 * di);
    }
}

// CENSUS also counts the pairs within each target's own radius (cnt[5]).
template <typename T, bool CENSUS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 6 : 1)))
void k_tree_diffusion(TreeDev t, double cut, double nu, double thr,
                                                        const double *__restrict__ sd,
                                                        double *__restrict__ u, double *__restrict__ v,
                                                        unsigned long long *__restrict__ cnt) {
    unsigned long long cp = 0, ci = 0;
    diffusion_wave<T, CENSUS>(t, cut, nu, thr, sd, u, v, cnt, cp, ci);
    block_count_add(cp, ci, cnt, CENSUS ? cnt + 5 : nullptr);
}

// The diffusion's d in sorted order, negated for the targets listed for
// k_diff_far: those with min(cut, kFarCut) d beyond far_radius(), the first
// t.fcap of them in sorted order (an ordered selection, so the listed set
// does not depend on scheduling); cnt[2] counts every such target, and
// targets past the capacity stay with their group's wave.
struct FarFlag {
    TreeDev t;
    double cut;
    const double *sd;  // d in sorted order (all positive before the selection)
    __device__ int operator()(int64_t ti) const { return fmin(cut, kFarCut) * sd[ti] > far_radius(t); }
};
struct FarWrite {
    TreeDev t;
    double *sd;
    int32_t *lists;
    __device__ void operator()(int64_t ti, int32_t pos, int fl) const {
        if (fl && pos < t.fcap) {
            lists[pos] = (int32_t)ti;
            sd[ti] = -sd[ti];
        }
    }
};

// sd[ti] = d[perm[ti]] (when the Biot-Savart pass's sorted copy is stale).
__global__ void k_sd_gather(const int32_t *__restrict__ perm, const double *__restrict__ d, int64_t m,
                            double *__restrict__ sd) {
    const int64_t ti = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ti < m) sd[ti] = d[perm[ti]];
}

// Targets whose radius exceeds far_radius(), listed by k_tree_diffusion.  The
// buckets are cut into chunks of 64 consecutive ones (Morton order) with a
// box each; k_far_pairs lists the (target, chunk) pairs whose boxes meet, and
// one wave per pair tests the chunk's 64 bucket boxes at once, then sums the
// buckets in range with one source per lane, eight buckets' loads in flight.
// Partial sums go to part[(chunk * 3 + q) * H + k] (zero for pairs not
// listed).
template <typename T>
__device__ __forceinline__ void far_load(const TreeDev &t, int64_t j0, int cnt, double bcx, double bcy, int lane,
                                         double ax, double ay, T &dx, T &dy, T &g) {
    if (lane < cnt) {
        if constexpr (sizeof(T) == 4) {
            dx = (float)(ax - bcx) - t.fx[j0 + lane];
            dy = (float)(ay - bcy) - t.fy[j0 + lane];
            g = t.fg[j0 + lane];
        } else {
            dx = ax - t.sx[j0 + lane];
            dy = ay - t.sy[j0 + lane];
            g = t.sg[j0 + lane];
        }
    } else {
        dx = T(1);  // outside the bucket: zero circulation
        dy = T(1);
        g = T(0);
    }
}

__device__ __forceinline__ double readlane_d(double a, int l) {
    const int2 v = *reinterpret_cast<int2 *>(&a);
    int2 r;
    r.x = __builtin_amdgcn_readlane(v.x, l);
    r.y = __builtin_amdgcn_readlane(v.y, l);
    return *reinterpret_cast<double *>(&r);
}

// One wave per chunk: the box of its 64 buckets.
__global__ __launch_bounds__(256) void k_chunk_box(TreeDev t, int64_t nchunk, double4 *__restrict__ cbox) {
    const int lane = threadIdx.x & 63;
    const int64_t ch = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (ch >= nchunk) return;
    const int64_t b = ch * 64 + lane;
    const bool ok = b < t.nleaf;
    const double4 q = ok ? t.nbb[b] : make_double4(1e300, -1e300, 1e300, -1e300);
    const double4 r = make_double4(wave_min(q.x), wave_max(q.y), wave_min(q.z), wave_max(q.w));
    if (lane == 0) cbox[ch] = r;
}

// One wave per listed target: the chunks whose box lies within its radius,
// appended as (target, chunk) pairs, one atomic per wave and 64 chunks.
__device__ void far_pairs_of(const TreeDev &t, const int32_t *__restrict__ list, int64_t k, int64_t nchunk,
                             const double4 *__restrict__ cbox, double cut, const double *__restrict__ dd,
                             int2 *__restrict__ pairs, unsigned long long *__restrict__ npairs,
                             double4 *__restrict__ ftgt, int lane) {
    const int64_t ti = list[k];
    const double di = dd[t.perm[ti]];
    const double R2 = (cut * di) * (cut * di);
    const double ax = t.sx[ti], ay = t.sy[ti];
    if (lane == 0) ftgt[k] = make_double4(ax, ay, di, 0.0);  // for the pair kernel
    for (int64_t c0 = 0; c0 < nchunk; c0 += 64) {
        const int64_t ch = c0 + lane;
        const bool need = ch < nchunk && box_d2(cbox[ch], ax, ay) < R2;
        const uint64_t m = __ballot(need);
        if (m == 0) continue;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(npairs, (unsigned long long)__popcll(m));
        base = __shfl(base, 0, 64);
        if (need) pairs[base + __popcll(m & ((1ull << lane) - 1))] = make_int2((int)k, (int)ch);
    }
}

__global__ __launch_bounds__(256) void k_far_pairs(TreeDev t, const int32_t *__restrict__ list,
                                                   const unsigned long long *__restrict__ cnt, int64_t nchunk,
                                                   const double4 *__restrict__ cbox, double cut,
                                                   const double *__restrict__ dd, int2 *__restrict__ pairs,
                                                   unsigned long long *__restrict__ npairs,
                                                   double4 *__restrict__ ftgt) {
    const int lane = threadIdx.x & 63;
    const int64_t H = far_count(cnt, t.fcap);
    for (int64_t k = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); k < H; k += (int64_t)gridDim.x * 4)
        far_pairs_of(t, list, k, nchunk, cbox, cut, dd, pairs, npairs, ftgt, lane);
}

// Zero the counters a[q] whose bit q is set in mask.
__global__ void k_zero_u64(unsigned long long *__restrict__ a, unsigned mask) {
    if (threadIdx.x < 32 && ((mask >> threadIdx.x) & 1u)) a[threadIdx.x] = 0ull;
}

// Zero the partials of the listed targets (grid-stride; H on the device).
__global__ void k_far_zero(double *__restrict__ a, int64_t nchunk, const unsigned long long *__restrict__ cnt,
                           int64_t cap) {
    const int64_t count = 3 * nchunk * far_count(cnt, cap);
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += (int64_t)gridDim.x * blockDim.x)
        a[t] = 0.0;
}

template <typename T>
__device__ void diff_far_pair(const TreeDev &t, const double4 *__restrict__ ftgt, int64_t H, int64_t k, int64_t ch,
                              double cut, double thr, double *__restrict__ part, unsigned long long &cpairs) {
    constexpr bool F32 = sizeof(T) == 4;
    const int lane = threadIdx.x & 63;
    const double4 tg = ftgt[k];
    const double ax = tg.x, ay = tg.y, di = tg.z;
    const double R2 = (cut * di) * (cut * di);
    const T kd = F32 ? (T)(-1.4426950408889634 / di) : (T)(1.0 / di);
    const T thr2 = F32 ? (T)(thr * thr) : (T)thr;
    // the chunk's buckets, one per lane: box test and the record in one load
    const int64_t b = ch * 64 + lane;
    const bool inb = b < t.nleaf;
    const NodeRec rb = inb ? t.rec[b] : NodeRec{0.0, 0.0, 0.0, -1, 0};
    const bool need = inb && box_d2(t.nbb[b], ax, ay) < R2;
    uint64_t mask = __ballot(need);
    double a2x = 0.0, a2y = 0.0, a1 = 0.0;
    int64_t npairs = 0;
    // the buckets in range, eight at a time: all their loads first, then the
    // sums in bucket order
    constexpr int kB = 8;
    while (mask) {
        T dx[kB], dy[kB], g[kB];
        int nbk = 0;
#pragma unroll
        for (int q = 0; q < kB; ++q) {
            dx[q] = dy[q] = T(1);
            g[q] = T(0);
            if (mask) {
                const int bl = __builtin_ctzll(mask);
                mask &= mask - 1;
                const int64_t j0 = -1 - __builtin_amdgcn_readlane(rb.a, bl);
                const int cnt = __builtin_amdgcn_readlane(rb.b, bl);
                far_load<T>(t, j0, cnt, readlane_d(rb.cx, bl), readlane_d(rb.cy, bl), lane, ax, ay, dx[q], dy[q],
                            g[q]);
                npairs += cnt;
                ++nbk;
            }
        }
#pragma unroll
        for (int q = 0; q < kB; ++q) {
            if (q < nbk) {
                T s2x = 0, s2y = 0, s1 = 0;
                DiffPair<T>::template add<true>(dx[q], dy[q], g[q], kd, thr2, s2x, s2y, s1);
                a2x += (double)s2x;
                a2y += (double)s2y;
                a1 += (double)s1;
            }
        }
    }
    a2x = wave_sum(a2x);
    a2y = wave_sum(a2y);
    a1 = wave_sum(a1);
    cpairs += (unsigned long long)npairs;
    if (lane != 0) return;
    part[(ch * 3 + 0) * H + k] = a2x;
    part[(ch * 3 + 1) * H + k] = a2y;
    part[(ch * 3 + 2) * H + k] = a1;
}

// Grid-stride over the listed pairs, one wave each; the count stays on the device.
template <typename T>
__global__ __launch_bounds__(256) void k_diff_far(TreeDev t, const double4 *__restrict__ ftgt,
                                                  const int2 *__restrict__ pairs,
                                                  const unsigned long long *__restrict__ npair_list, double cut,
                                                  double thr, double *__restrict__ part,
                                                  unsigned long long *__restrict__ cnt) {
    const int64_t H = far_count(cnt, t.fcap);
    const int64_t total = (int64_t)*npair_list;
    const int64_t nw = (int64_t)gridDim.x * 4;
    unsigned long long cp = 0;
    for (int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); w < total; w += nw)
        diff_far_pair<T>(t, ftgt, H, pairs[w].x, pairs[w].y, cut, thr, part, cp);
    block_count_add(cp, 0ull, cnt, nullptr);
}

// One work-group per listed target: the chunk partials summed in a fixed
// order (strided per thread, then a tree in LDS), then the velocity update.
__device__ void far_final_of(const TreeDev &t, const int32_t *__restrict__ list, int64_t H, int64_t k,
                             int64_t nchunk, const double *__restrict__ part, const double *__restrict__ dd,
                             double nu, double *__restrict__ u, double *__restrict__ v) {
    __shared__ double r[3][256];
    double a2x = 0.0, a2y = 0.0, a1 = 0.0;
    for (int64_t ch = threadIdx.x; ch < nchunk; ch += 256) {
        a2x += part[(ch * 3 + 0) * H + k];
        a2y += part[(ch * 3 + 1) * H + k];
        a1 += part[(ch * 3 + 2) * H + k];
    }
    r[0][threadIdx.x] = a2x;
    r[1][threadIdx.x] = a2y;
    r[2][threadIdx.x] = a1;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int q = 0; q < 3; ++q) r[q][threadIdx.x] += r[q][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const int32_t i = t.perm[list[k]];
    if (fabs(r[2][0]) > 1e-5) {
        const double di = dd[i];
        u[i] += nu * r[0][0] / (r[2][0] * di);
        v[i] += nu * r[1][0] / (r[2][0] * di);
    }
}

__global__ __launch_bounds__(256) void k_diff_far_final(TreeDev t, const int32_t *__restrict__ list,
                                                        const unsigned long long *__restrict__ cnt,
                                                        int64_t nchunk, const double *__restrict__ part,
                                                        const double *__restrict__ dd, double nu,
                                                        double *__restrict__ u, double *__restrict__ v) {
    const int64_t H = far_count(cnt, t.fcap);
    for (int64_t k = blockIdx.x; k < H; k += gridDim.x) far_final_of(t, list, H, k, nchunk, part, dd, nu, u, v);
}

// Slots left out of the tree (dead and far from the box) see no neighbour.
__global__ void k_tree_outside(const int32_t *__restrict__ perm, int64_t m, int64_t s, double vinf_x,
                               double vinf_y, double *__restrict__ u, double *__restrict__ v,
                               double *__restrict__ d) {
    const int64_t t = m + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= s) return;
    const int32_t i = perm[t];
    u[i] = vinf_x;
    v[i] = vinf_y;
    d[i] = sqrt(1e5);
}

// ---- host side ----------------------------------------------------------------
void tree_d_changed(vx_case *c) {
    if (c->tree) c->tree->sd_fresh = false;
}

static TreeBufs &tree_of(vx_case *c) {
    if (!c->tree) c->tree = new TreeBufs();
    return *c->tree;
}

void tree_release(vx_case *c) {
    if (!c->tree) return;
    TreeBufs &t = *c->tree;
    t.keys.release(); t.keys_alt.release(); t.idx.release(); t.perm.release(); t.incl.release();
    t.child0.release(); t.child1.release(); t.first.release(); t.count.release(); t.parent.release();
    t.fb.release(); t.rlo.release(); t.rhi.release(); t.rleft.release(); t.rright.release();
    t.bstart.release(); t.bscan.release(); t.upper.release(); t.uscan.release(); t.toplist.release();
    t.sx.release(); t.sy.release(); t.sg.release(); t.ncx.release(); t.ncy.release(); t.nr.release();
    t.fx.release(); t.fy.release(); t.fg.release(); t.nbb.release(); t.part.release(); t.rec.release();
    t.cbox.release(); t.fpairs.release(); t.ftgt.release();
    t.gcs.release(); t.gcs2.release(); t.gflag.release(); t.gscan.release(); t.gfirst.release(); t.pcount.release(); t.arr.release(); t.sizes.release();
    t.cd.release(); t.cf.release(); t.cnt.release(); t.tmp.release(); t.tmp2.release(); t.fbr2.release(); t.hpart.release();
    t.sd.release();
    if (t.h_far) (void)hipHostFree(t.h_far);
    if (t.ev) (void)hipEventDestroy(t.ev);
    delete c->tree;
    c->tree = nullptr;
}

// Coefficients stored per node (the upward pass works on groups of 16 or 32
// lanes) and terms used by the far field: 16, 20 or 32.
static int tree_order(const vx_case *c) { return c->prm.order > 16 ? 32 : 16; }
static int tree_terms(const vx_case *c) { return c->prm.order >= 32 ? 32 : (c->prm.order >= 20 ? 20 : 16); }

// Builds the tree on stream st in two parts: A (inclusion, box, keys, sort;
// no host wait) and B (the rest, after the included count is back).  With
// B.P > 0 the newborns' positions come from B and the gather of the sorted
// particles first waits for born_ev.
static int tree_build_a(vx_case *c, int64_t s, hipStream_t st, Births B) {
    TreeBufs &t = tree_of(c);
    const int order = tree_order(c);
    t.order = order;
    t.terms = tree_terms(c);
    t.s = s;
    const int64_t nbs = (s + 255) / 256;
    if (t.keys.ensure(s, st) || t.keys_alt.ensure(s, st) || t.idx.ensure(s, st) || t.perm.ensure(s, st) ||
        t.incl.ensure(s, st) || t.part.ensure(nbs + 1, st) || t.pcount.ensure(nbs, st) || t.cnt.ensure(16, st))
        return 1;
    VX_CHECK(prim::fill(t.cnt.p, 16, 0ull, st));
    const double4 dbox = make_double4(c->prm.x_min, c->prm.x_max, c->prm.y_min, c->prm.y_max);
    hipLaunchKernelGGL(k_incl_box, dim3((unsigned)nbs), dim3(256), 0, st, c->x.p, c->y.p, c->g.p, s, dbox, t.incl.p,
                       t.part.p, t.pcount.p, B);
    hipLaunchKernelGGL(k_box_final, dim3(1), dim3(256), 0, st, t.part.p, t.pcount.p, nbs, t.part.p + nbs, t.cnt.p,
                       c->d_dev);
    // the included count reaches the host (c->h_dev[0]) while the keys are sorted
    if (!t.ev) VX_CHECK(hipEventCreateWithFlags(&t.ev, hipEventDisableTiming));
    VX_CHECK(hipEventRecord(t.ev, st));
    hipLaunchKernelGGL(k_tree_keys, dim3((unsigned)nbs), dim3(256), 0, st, c->x.p, c->y.p, t.incl.p, s,
                       t.part.p + nbs, t.keys.p, t.idx.p, B);
    // stable LSD radix sort of (key, index); the sorted pairs end in keys and
    // idx, which then take the names keys_alt and perm
    if (t.tmp.ensure(sizeof(int32_t) * (size_t)prim::radix_tmp_size(s), st)) return 1;
    VX_CHECK(prim::radix_sort_pairs(t.keys.p, t.keys_alt.p, t.idx.p, t.perm.p, s, (int32_t *)t.tmp.p,
                                    (int32_t *)(t.cnt.p + 15), st));
    std::swap(t.keys, t.keys_alt);
    std::swap(t.idx, t.perm);
    return 0;
}

static int tree_build_b(vx_case *c, hipStream_t st, Births B) {
    TreeBufs &t = *c->tree;
    const int order = t.order;
    VX_CHECK(hipEventSynchronize(t.ev));
    const int64_t m = c->h_dev[0];
    t.m = m;
    t.nleaf = 0;
    t.ngroup = 0;
    t.nnode = 0;
    if (m == 0) return 0;
    if (t.sx.ensure(m, st) || t.sy.ensure(m, st) || t.sg.ensure(m, st) || t.fx.ensure(m, st) || t.fy.ensure(m, st) ||
        t.fg.ensure(m, st) || t.fb.ensure(m, st) || t.fbr2.ensure(m, st) || t.rlo.ensure(m, st) || t.rhi.ensure(m, st) ||
        t.rleft.ensure(m, st) || t.rright.ensure(m, st) || t.bstart.ensure(m, st) || t.bscan.ensure(m, st) ||
        t.upper.ensure(m, st) || t.uscan.ensure(m, st) || t.gcs.ensure(m, st) || t.gcs2.ensure(m, st) ||
        t.gflag.ensure(m, st) || t.gscan.ensure(m, st) || t.gfirst.ensure(m, st))
        return 1;
    const unsigned gm = (unsigned)((m + 255) / 256);
    int64_t nb = 1, nu = 0;
    // Key prefix depth of a bucket: at most one quadtree level coarser than a
    // cell holding kLeaf particles at the mean density.  Keys normalise each
    // axis of the bounding box separately, so a level-L cell holds m / 4^L
    // particles on average whatever the box's aspect ratio.
    const double lv = floor(0.5 * log2((double)m / kLeaf)) - (double)(c->tune.leaf_off >= 0 ? c->tune.leaf_off : 1);
    const int dmin = 2 * (int)fmin(fmax(lv, 0.0), 15.0);
    // Three independent chains follow the sort: the target groups (aux[0]),
    // the gather of the sorted particles (aux[1]) and the bucket and upper
    // node marks (st).  st waits for the groups before the counts and for
    // the gather before the buckets' multipoles.
    if (t.tmp.ensure(sizeof(int32_t) * (size_t)(prim::tiles_of(m) + 1), st) ||
        t.tmp2.ensure(sizeof(int32_t) * (size_t)(prim::tiles_of(m) + 1), st) || aux_streams(c))
        return 1;
    VX_CHECK(hipEventRecord(c->aux_fork, st));
    {
        hipStream_t sg = c->aux[0];
        VX_CHECK(hipStreamWaitEvent(sg, c->aux_fork, 0));
        hipLaunchKernelGGL(k_group_cells, dim3(gm), dim3(256), 0, sg, t.keys_alt.p, m, 32 - dmin, t.gcs.p);
        int32_t *tt = (int32_t *)t.tmp2.p;
        VX_CHECK(prim::scan_values<true>(prim::ArrayValue{t.gcs.p}, m, prim::OpMax(), prim::WriteScan{t.gcs2.p}, tt,
                                         (int32_t *)nullptr, sg));
        hipLaunchKernelGGL(k_group_flags, dim3(gm), dim3(256), 0, sg, t.gcs2.p, m, t.gflag.p);
        VX_CHECK(prim::select_flags(prim::ArrayFlag{t.gflag.p}, m, prim::WriteScan{t.gscan.p}, tt, (int32_t *)nullptr,
                                    sg));
        hipLaunchKernelGGL(k_group_first, dim3(gm), dim3(256), 0, sg, t.gflag.p, t.gscan.p, m, t.gfirst.p);
        VX_CHECK(hipEventRecord(c->aux_done[0], sg));
    }
    {
        // the sorted particles, newborns included
        hipStream_t sa = c->aux[1];
        VX_CHECK(hipStreamWaitEvent(sa, c->aux_fork, 0));
        if (B.P > 0) VX_CHECK(hipStreamWaitEvent(sa, c->born_ev, 0));
        hipLaunchKernelGGL(k_tree_gather, dim3(gm), dim3(256), 0, sa, c->x.p, c->y.p, c->g.p, t.perm.p, m, t.sx.p,
                           t.sy.p, t.sg.p);
        VX_CHECK(hipEventRecord(c->aux_done[1], sa));
    }
    if (m > kLeaf) {
        VX_CHECK(prim::fill(t.bstart.p, m, 0, st));
        hipLaunchKernelGGL(k_radix, dim3(gm), dim3(256), 0, st, t.keys_alt.p, m, t.rlo.p, t.rhi.p, t.rleft.p,
                           t.rright.p);
        hipLaunchKernelGGL(k_bucket_mark, dim3(gm), dim3(256), 0, st, t.keys_alt.p, t.rlo.p, t.rhi.p, t.rleft.p,
                           t.rright.p, m, dmin, t.bstart.p, t.upper.p);
        int32_t *tt = (int32_t *)t.tmp.p;
        VX_CHECK(prim::select_flags(prim::ArrayFlag{t.bstart.p}, m, prim::WriteScan{t.bscan.p}, tt,
                                    (int32_t *)nullptr, st));
        VX_CHECK(prim::select_flags(prim::ArrayFlag{t.upper.p}, m - 1, prim::WriteScan{t.uscan.p}, tt,
                                    (int32_t *)nullptr, st));
    }
    // the group, bucket and upper-node counts in one read-back; the node
    // links and the cut list, which need only the device counts, are queued
    // before the host waits for it, their arrays sized for the bound
    // nn <= 2 m - 1
    const int64_t nnb = 2 * m;
    if (t.sizes.ensure(4, st) || t.first.ensure(nnb, st) || t.count.ensure(nnb, st) || t.child0.ensure(nnb, st) ||
        t.child1.ensure(nnb, st) || t.parent.ensure(nnb, st) || t.arr.ensure((size_t)nnb, st))
        return 1;
    VX_CHECK(hipStreamWaitEvent(st, c->aux_done[0], 0));
    hipLaunchKernelGGL(k_tree_counts, dim3(1), dim3(1), 0, st, t.gscan.p, t.gflag.p, t.bscan.p, t.bstart.p, t.uscan.p,
                       t.upper.p, m, (int)(m > kLeaf), (const int32_t *)(t.cnt.p + 15), t.sizes.p,
                       c->d_dev + 4);
    VX_CHECK(hipEventRecord(t.ev, st));
    hipLaunchKernelGGL(k_parent_init, dim3((unsigned)std::min<int64_t>(1024, (nnb + 255) / 256)), dim3(256), 0, st,
                       (const int64_t *)t.sizes.p, t.parent.p);
    if (m > kLeaf) {
        hipLaunchKernelGGL(k_bucket_first, dim3(gm), dim3(256), 0, st, t.bstart.p, t.bscan.p, m, t.first.p);
        hipLaunchKernelGGL(k_bucket_count, dim3(gm), dim3(256), 0, st, t.first.p, (const int64_t *)t.sizes.p, m,
                           t.count.p);
        hipLaunchKernelGGL(k_upper_nodes, dim3(gm), dim3(256), 0, st, t.rlo.p, t.rhi.p, t.rleft.p, t.rright.p,
                           t.upper.p, t.uscan.p, t.bscan.p, m, (const int64_t *)t.sizes.p, t.first.p, t.count.p,
                           t.child0.p, t.child1.p, t.parent.p);
    } else {
        const int32_t one[2] = {0, (int32_t)m};
        VX_CHECK(hipMemcpyAsync(t.first.p, &one[0], sizeof(int32_t), hipMemcpyHostToDevice, st));
        VX_CHECK(hipMemcpyAsync(t.count.p, &one[1], sizeof(int32_t), hipMemcpyHostToDevice, st));
    }
    // cut nodes, in any order (count in cnt[10], cleared with the counters)
    int32_t *cut = t.arr.p;
    unsigned long long *ncut = t.cnt.p + 10;
    hipLaunchKernelGGL(k_cut_list, dim3((unsigned)std::min<int64_t>(1024, (nnb + 255) / 256)), dim3(256), 0, st,
                       CutNode{t.first.p, t.count.p, t.bscan.p, t.parent.p, (const int64_t *)t.sizes.p, m},
                       (const int64_t *)t.sizes.p, cut, ncut);
    VX_CHECK(hipEventSynchronize(t.ev));
    const int64_t *hc = c->h_dev + 4;
    if (hc[3]) {
        set_error("tree key sort: a tile's look-back did not finish");
        return 1;
    }
    t.ngroup = hc[0];
    nb = hc[1];
    nu = hc[2];
    t.nleaf = nb;
    t.nnode = nb + nu;
    t.root = (int32_t)(nu > 0 ? nb : 0);
    const int64_t nn = t.nnode;
    if (t.ncx.ensure(nn, st) || t.ncy.ensure(nn, st) || t.nr.ensure(nn, st) || t.nbb.ensure(nn, st) ||
        t.toplist.ensure(nn, st) || t.rec.ensure(nn, st) || t.cd.ensure(nn * order, st) ||
        t.cf.ensure(nn * order, st))
        return 1;
    const unsigned lb = (unsigned)((nb + 3) / 4);
    const unsigned gc = (unsigned)std::min<int64_t>(nb, 2048);
    VX_CHECK(hipStreamWaitEvent(st, c->aux_done[1], 0));  // the gathered particles
    if (order == 32) {
        hipLaunchKernelGGL(k_bucket_build<32>, dim3(lb), dim3(256), 0, st, t.sx.p, t.sy.p, t.sg.p, nb, t.first.p,
                           t.count.p, t.fx.p, t.fy.p, t.fg.p, t.ncx.p, t.ncy.p, t.nr.p, t.nbb.p, t.child0.p,
                           t.child1.p, t.cd.p, t.cf.p);
        hipLaunchKernelGGL(k_upper_local<32>, dim3(gc), dim3(256), 0, st, cut, ncut, nb, t.child0.p, t.child1.p,
                           t.ncx.p, t.ncy.p, t.nr.p, t.nbb.p, t.cd.p, t.cf.p);
        hipLaunchKernelGGL(k_upper_top<32>, dim3(1), dim3(1024), 0, st, t.root, nb, m, t.first.p, t.count.p,
                           t.bscan.p, t.toplist.p, t.child0.p, t.child1.p, t.ncx.p, t.ncy.p, t.nr.p, t.nbb.p, t.cd.p,
                           t.cf.p);
    } else {
        hipLaunchKernelGGL(k_bucket_build<16>, dim3(lb), dim3(256), 0, st, t.sx.p, t.sy.p, t.sg.p, nb, t.first.p,
                           t.count.p, t.fx.p, t.fy.p, t.fg.p, t.ncx.p, t.ncy.p, t.nr.p, t.nbb.p, t.child0.p,
                           t.child1.p, t.cd.p, t.cf.p);
        hipLaunchKernelGGL(k_upper_local<16>, dim3(gc), dim3(256), 0, st, cut, ncut, nb, t.child0.p, t.child1.p,
                           t.ncx.p, t.ncy.p, t.nr.p, t.nbb.p, t.cd.p, t.cf.p);
        hipLaunchKernelGGL(k_upper_top<16>, dim3(1), dim3(1024), 0, st, t.root, nb, m, t.first.p, t.count.p,
                           t.bscan.p, t.toplist.p, t.child0.p, t.child1.p, t.ncx.p, t.ncy.p, t.nr.p, t.nbb.p, t.cd.p,
                           t.cf.p);
    }
    hipLaunchKernelGGL(k_pack_nodes, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, nn, t.ncx.p, t.ncy.p,
                       t.nr.p, t.child0.p, t.child1.p, t.first.p, t.count.p, t.rec.p);
    VX_CHECK(hipGetLastError());
    return 0;
}

// Part A of the next tree build on the side stream, queued before the birth
// stage's work: it waits for the state before the birth (pre_ev) and takes
// the newborns' positions at [at, at + P) from the case's birth points.
int tree_build_early(vx_case *c, int64_t s, int64_t at) {
    if (side_stream(c)) return 1;
    VX_CHECK(hipStreamWaitEvent(c->side, c->pre_ev, 0));
    const int rc = tree_build_a(c, s, c->side, Births{at, c->P, c->bpos});
    if (rc) {
        (void)side_join(c);
        return rc;
    }
    c->born_at = at;
    c->tree_early_s = s;
    return 0;
}

// With part A queued by tree_build_early, part B follows on the side stream
// and the main stream continues after it; else the whole build runs on the
// main stream.
int tree_build(vx_case *c, int64_t s) {
    if (c->born_at < 0) {
        const Births none{-1, 0, nullptr};
        return tree_build_a(c, s, c->stream, none) || tree_build_b(c, c->stream, none);
    }
    const Births B{c->born_at, c->P, c->bpos};
    c->born_at = -1;
    VX_ARG(s == c->tree_early_s, "tree size differs from the one started with the birth stage");
    const int rc = tree_build_b(c, c->side, B);
    const int rj = side_join(c);
    return rc ? rc : rj;
}

static double theta_of(const vx_case *c) { return c->prm.theta > 0 ? c->prm.theta : 0.5; }

// Velocities at probe points over the tree of the preceding
// launch_tree_biot_savart (A19's sums, src/kernel.cu:1434-1475: every source,
// max(r^2, eps^2), no nearest-3).  One wave per 8 consecutive probes, 8 lanes
// per probe: the traversal of k_tree_bs with the 8 probes' box as the target
// circle; each probe's far nodes (M2P at the probe, as the FP64 Biot-Savart
// path evaluates them) and each near bucket's sources are dealt round-robin
// to its 8 lanes, and the lanes' sums are added in a fixed order at the end.
// A few hundred probes give a hundred waves of short traversals, where 64
// probes per wave left a dozen long ones.  FP32: positions relative to the
// source bucket's centre or the node's, differenced in FP64; sums in FP64 per
// bucket and node.
template <typename T, int ORDER, int STRIDE>
__global__ __launch_bounds__(256) void k_tree_probe(TreeDev t, const double *__restrict__ px,
                                                    const double *__restrict__ py, int64_t m, double theta,
                                                    double eps, double vinf_x, double vinf_y,
                                                    double *__restrict__ pu, double *__restrict__ pv) {
    using PM = PairMath<T>;
    constexpr bool F32 = sizeof(T) == 4;
    constexpr int kSub = 8, kPer = 64 / kSub;
    __shared__ Src4T<T> lds_all[4][kLeaf];
    __shared__ int32_t stk_all[4][kBsStack];
    __shared__ int32_t lq_all[4][64];
    __shared__ int32_t fq_all[4][64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, sub = lane & (kSub - 1);
    Src4T<T> *lds = lds_all[wid];
    int32_t *stk = stk_all[wid];
    int32_t *lq = lq_all[wid];
    int32_t *fq = fq_all[wid];
    const int64_t base = ((int64_t)blockIdx.x * 4 + wid) * kPer;
    if (base >= m) return;
    const int64_t i = base + lane / kSub;
    const bool ok = i < m;
    const double qx = px[ok ? i : base], qy = py[ok ? i : base];
    const double x0 = wave_min(qx), x1 = wave_max(qx), y0 = wave_min(qy), y1 = wave_max(qy);
    const double cAx = 0.5 * (x0 + x1), cAy = 0.5 * (y0 + y1);
    const double rA = 0.5 * sqrt((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0));
    const T eps2 = (T)(eps * eps);
    const double th2 = theta * theta;
    const uint64_t below = (1ull << lane) - 1;
    double su = 0.0, sv = 0.0;
    int sp = 1;
    if (lane == 0) stk[0] = t.root;
    while (sp > 0) {
        const int k = sp + 64 + 64 <= kBsStack ? (sp < 64 ? sp : 64) : 1;
        __builtin_amdgcn_wave_barrier();
        const int node = lane < k ? stk[sp - 1 - lane] : -1;
        __builtin_amdgcn_wave_barrier();
        sp -= k;
        bool isfar = false, isleaf = false, isopen = false;
        int ca = 0, cb = 0;
        if (node >= 0) {
            const NodeRec nd = t.rec[node];
            const double ddx = nd.cx - cAx, ddy = nd.cy - cAy;
            const double d2 = ddx * ddx + ddy * ddy;
            const double rs = nd.r + rA;
            isfar = d2 * th2 > rs * rs && d2 > (rs + eps) * (rs + eps);
            isleaf = !isfar && nd.a < 0;
            isopen = !isfar && !isleaf;
            ca = nd.a;
            cb = nd.b;
        }
        const uint64_t om = __ballot(isopen);
        if (isopen) {
            const int pos = sp + 2 * __popcll(om & below);
            stk[pos] = cb;
            stk[pos + 1] = ca;
        }
        sp += 2 * __popcll(om);
        const uint64_t lm = __ballot(isleaf);
        if (isleaf) lq[__popcll(lm & below)] = node;
        const int nl = __popcll(lm);
        const uint64_t fm = __ballot(isfar);
        if (isfar) fq[__popcll(fm & below)] = node;
        const int nf = __popcll(fm);
        __builtin_amdgcn_wave_barrier();
        // M2P: S = (1/D) sum_k b_k (R/D)^k, D = z - c_node; u += Im S, v += Re S
        for (int fi = sub; fi < nf; fi += kSub) {
            const int fnode = fq[fi];
            const NodeRec nd = t.rec[fnode];
            const T Dx = (T)(qx - nd.cx), Dy = (T)(qy - nd.cy);
            const T den = Dx * Dx + Dy * Dy;
            const T inv = PM::div(T(1), den);
            const T qre = Dx * inv, qim = -Dy * inv;
            const T R = (T)fmax(nd.r, kTiny);
            const T tx = R * qre, ty = R * qim;
            T sr, si;
            if constexpr (F32) {
                const float2 *bm = t.cf + (int64_t)fnode * STRIDE;
                sr = bm[ORDER - 1].x;
                si = bm[ORDER - 1].y;
#pragma unroll
                for (int q = ORDER - 2; q >= 0; --q) {
                    const T nr_ = sr * tx - si * ty + bm[q].x;
                    si = sr * ty + si * tx + bm[q].y;
                    sr = nr_;
                }
            } else {
                const double2 *bm = t.cd + (int64_t)fnode * STRIDE;
                sr = bm[ORDER - 1].x;
                si = bm[ORDER - 1].y;
#pragma unroll
                for (int q = ORDER - 2; q >= 0; --q) {
                    const T nr_ = sr * tx - si * ty + bm[q].x;
                    si = sr * ty + si * tx + bm[q].y;
                    sr = nr_;
                }
            }
            su += (double)(sr * qim + si * qre);
            sv += (double)(sr * qre - si * qim);
        }
        for (int li = 0; li < nl; ++li) {
            const int lnode = __builtin_amdgcn_readfirstlane(lq[li]);
            const NodeRec nd = t.rec[lnode];
            const int cnt = nd.b;
            stage_leaf<T>(lds, lane, -1 - nd.a, cnt, t);
            // FP32 sources are staged relative to their bucket's centre
            const T ox = F32 ? (T)(qx - nd.cx) : (T)qx, oy = F32 ? (T)(qy - nd.cy) : (T)qy;
            T pu_ = 0, pv_ = 0;
            for (int jj = sub; jj < cnt; jj += kSub) {
                const Src4T<T> q = lds[jj];
                const T dx = ox - q.x, dy = oy - q.y;
                const T r2 = dx * dx + dy * dy;
                const T w = PM::div(q.g, r2 > eps2 ? r2 : eps2);
                pv_ += w * dx;
                pu_ -= w * dy;
            }
            su += (double)pu_;
            sv += (double)pv_;
        }
        __builtin_amdgcn_wave_barrier();
    }
#pragma unroll
    for (int o = 1; o < kSub; o <<= 1) {
        su += __shfl_xor(su, o);
        sv += __shfl_xor(sv, o);
    }
    if (!ok || sub != 0) return;
    pu[i] = su / (2 * kPi) + vinf_x;
    pv[i] = sv / (2 * kPi) + vinf_y;
}

int launch_tree_probes(vx_case *c, const double *px, const double *py, int64_t m, double vinf_x, double vinf_y,
                       double *pu, double *pv, hipStream_t st) {
    if (m <= 0) return 0;
    if (!c->tree || c->tree->m == 0) {
        set_error("internal: tree probes without a tree");
        return 1;
    }
    TreeBufs &t = *c->tree;
    const TreeDev tv = view(t);
    const unsigned nb = (unsigned)((m + 31) / 32);  // 4 waves of 8 probes
    const double th = theta_of(c), eps = c->prm.ve_size;
    const bool mixed = c->prm.precision == VX_PREC_MIXED;
#define VX_TREE_PROBE(T_, P_, S_)                                                                               \
    hipLaunchKernelGGL((k_tree_probe<T_, P_, S_>), dim3(nb), dim3(256), 0, st, tv, px, py, m, th, eps, vinf_x, \
                       vinf_y, pu, pv)
    if (mixed) {
        if (t.terms == 32)
            VX_TREE_PROBE(float, 32, 32);
        else if (t.terms == 20)
            VX_TREE_PROBE(float, 20, 32);
        else
            VX_TREE_PROBE(float, 16, 16);
    } else {
        if (t.terms == 32)
            VX_TREE_PROBE(double, 32, 32);
        else if (t.terms == 20)
            VX_TREE_PROBE(double, 20, 32);
        else
            VX_TREE_PROBE(double, 16, 16);
    }
#undef VX_TREE_PROBE
    VX_CHECK(hipGetLastError());
    return 0;
}

int launch_tree_biot_savart(vx_case *c, int64_t s, double vinf_x, double vinf_y) {
    {
        Timer tb(c, &c->tm.tree_build_ms, 5);
        if (tree_build(c, s)) return 1;
    }
    TreeBufs &t = *c->tree;
    hipStream_t st = c->stream;
    if (s > t.m)
        hipLaunchKernelGGL(k_tree_outside, dim3((unsigned)((s - t.m + 255) / 256)), dim3(256), 0, st, t.perm.p, t.m, s,
                           vinf_x, vinf_y, c->u.p, c->v.p, c->d.p);
    if (t.m == 0) return 0;
    if (t.sd.ensure((size_t)t.m, st)) return 1;
    const TreeDev tv = view(t);
    const unsigned lb = (unsigned)((t.nleaf + 3) / 4 + 7) / 8 * 8;  // padded for xcd_block()
    const double th = theta_of(c), eps = c->prm.ve_size;
    unsigned long long *nfb = t.cnt.p + 1;
    const bool mixed = c->prm.precision == VX_PREC_MIXED;
    const int exact_all = (c->prm.flags & VX_FLAG_EXACT_NEAREST) ? 1 : 0;
    Timer tk(c, &c->tm.bs_kernel_ms, 6);
#define VX_TREE_BS(T_, P_, S_)                                                                                       \
    hipLaunchKernelGGL((k_tree_bs<T_, P_, S_>), dim3(lb), dim3(256), 0, st, tv, th, eps, vinf_x, vinf_y, c->u.p,     \
                       c->v.p, c->d.p, t.sd.p, t.fb.p, t.fbr2.p, nfb, exact_all)
    if (mixed) {
        if (t.terms == 32)
            VX_TREE_BS(float, 32, 32);
        else if (t.terms == 20)
            VX_TREE_BS(float, 20, 32);
        else
            VX_TREE_BS(float, 16, 16);
    } else {
        if (t.terms == 32)
            VX_TREE_BS(double, 32, 32);
        else if (t.terms == 20)
            VX_TREE_BS(double, 20, 32);
        else
            VX_TREE_BS(double, 16, 16);
    }
#undef VX_TREE_BS
    tk.stop();
    hipLaunchKernelGGL(k_tree_nearest_fallback, dim3(512), dim3(1024), 0, st, tv, t.arr.p, t.cnt.p + 10, t.fb.p,
                       t.fbr2.p, nfb, c->d.p, t.sd.p);
    VX_CHECK(hipGetLastError());
    t.sd_fresh = true;
    return 0;
}

// The diffusion over the tree of the preceding launch_tree_biot_savart (same
// positions) in three parts: tree_diffusion_pre (main stream: counters, far
// list, sorted d), tree_diffusion_far (the listed targets; any stream ordered
// after the pre part) and tree_diffusion_main (the groups' waves; main
// stream).  The far and main parts write disjoint targets.
static double diff_cut_of(const vx_case *c) {
    const bool mixed = c->prm.precision == VX_PREC_MIXED;
    // omitted weights total at most (1 + c) e^{-c} of the near sum's scale:
    // 3e-7 at c = 18 (the FP32 rounding level of the sums), 3e-12 at c = 30
    return c->prm.diff_cut > 0 ? c->prm.diff_cut : (mixed ? 18.0 : 30.0);
}

int tree_diffusion_pre(vx_case *c) {
    if (!c->tree || c->tree->m == 0) return 0;
    TreeBufs &t = *c->tree;
    hipStream_t st = c->stream;
    const int64_t nchunk = (t.nleaf + 63) / 64;
    // The last diffusion's count comes from pinned memory, copied at the end
    // of its far part; a caller that has not synchronised since reads an
    // older count, which costs speed only.  The growth stays within a 4 GiB
    // budget for the chunk partials.
    if (!t.h_far) {
        VX_CHECK(hipHostMalloc((void **)&t.h_far, sizeof(unsigned long long), hipHostMallocDefault));
        *t.h_far = 0;
    }
    const int64_t last = (int64_t)*(volatile unsigned long long *)t.h_far;
    const int64_t budget = (int64_t(1) << 32) / (3 * (int64_t)sizeof(double) * std::max<int64_t>(nchunk, 1));
    const int64_t grow = std::min<int64_t>({last + last / 4, t.m, budget});
    t.fcap = std::max<int64_t>({kFarCap, t.m / 16, grow});
    if (t.hpart.ensure((size_t)(3 * nchunk * t.fcap), st) || t.cbox.ensure((size_t)nchunk, st) ||
        t.fpairs.ensure((size_t)(nchunk * t.fcap), st) || t.ftgt.ensure((size_t)t.fcap, st) ||
        t.sd.ensure((size_t)t.m, st))
        return 1;
    // pairs, overflow, far targets, far (target, chunk) pairs; [4] is the
    // BS's M2L count, [5] the census of pairs within the targets' radii
    unsigned long long *cnt = t.cnt.p + 4;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, cnt, 0x2du);  // not the overflow flag nor M2L count
    if (c->ptmp.ensure((size_t)prim::tiles_of(t.m) + 1, st)) return 1;
    const TreeDev tv = view(t);
    // d in sorted order: the Biot-Savart pass's copy unless d changed since
    // or an earlier selection negated entries of it
    if (!t.sd_fresh)
        hipLaunchKernelGGL(k_sd_gather, dim3((unsigned)((t.m + 255) / 256)), dim3(256), 0, st, t.perm.p, c->d.p, t.m,
                           t.sd.p);
    t.sd_fresh = false;
    VX_CHECK(prim::select_flags(FarFlag{tv, diff_cut_of(c), t.sd.p}, t.m, FarWrite{tv, t.sd.p, t.fb.p}, c->ptmp.p,
                                cnt + 2, st));
    return 0;
}

int tree_diffusion_far(vx_case *c, hipStream_t st) {
    if (!c->tree || c->tree->m == 0) return 0;
    TreeBufs &t = *c->tree;
    const TreeDev tv = view(t);
    const bool mixed = c->prm.precision == VX_PREC_MIXED;
    const double cut = diff_cut_of(c), thr = 0.001 * c->prm.ve_size;
    unsigned long long *cnt = t.cnt.p + 4;
    // all counts stay on the device (grid-stride kernels sized for the list's capacity)
    const int64_t nchunk = (t.nleaf + 63) / 64;
    hipLaunchKernelGGL(k_far_zero, dim3(512), dim3(256), 0, st, t.hpart.p, nchunk, cnt, t.fcap);
    VX_CHECK(hipMemcpyAsync(t.h_far, cnt + 2, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    hipLaunchKernelGGL(k_chunk_box, dim3((unsigned)((nchunk + 3) / 4)), dim3(256), 0, st, tv, nchunk, t.cbox.p);
    hipLaunchKernelGGL(k_far_pairs, dim3(256), dim3(256), 0, st, tv, t.fb.p, cnt, nchunk, t.cbox.p, cut, c->d.p,
                       t.fpairs.p, cnt + 3, t.ftgt.p);
    if (mixed)
        hipLaunchKernelGGL(k_diff_far<float>, dim3(4096), dim3(256), 0, st, tv, t.ftgt.p, t.fpairs.p, cnt + 3, cut,
                           thr, t.hpart.p, cnt);
    else
        hipLaunchKernelGGL(k_diff_far<double>, dim3(4096), dim3(256), 0, st, tv, t.ftgt.p, t.fpairs.p, cnt + 3, cut,
                           thr, t.hpart.p, cnt);
    hipLaunchKernelGGL(k_diff_far_final, dim3(1024), dim3(256), 0, st, tv, t.fb.p, cnt, nchunk, t.hpart.p, c->d.p,
                       c->prm.nu, c->u.p, c->v.p);
    VX_CHECK(hipGetLastError());
    return 0;
}

int tree_diffusion_main(vx_case *c) {
    if (!c->tree || c->tree->m == 0) return 0;
    TreeBufs &t = *c->tree;
    hipStream_t st = c->stream;
    const unsigned lb8 = (unsigned)((t.ngroup + 3) / 4 + 7) / 8 * 8;  // target groups, padded for xcd_block()
    const bool mixed = c->prm.precision == VX_PREC_MIXED;
    const double cut = diff_cut_of(c), thr = 0.001 * c->prm.ve_size;
    unsigned long long *cnt = t.cnt.p + 4;
    Timer tk(c, &c->tm.diff_kernel_ms, 8);
    if (c->census) {
        if (mixed)
            hipLaunchKernelGGL((k_tree_diffusion<float, true>), dim3(lb8), dim3(256), 0, st, view(t), cut, c->prm.nu,
                               thr, t.sd.p, c->u.p, c->v.p, cnt);
        else
            hipLaunchKernelGGL((k_tree_diffusion<double, true>), dim3(lb8), dim3(256), 0, st, view(t), cut,
                               c->prm.nu, thr, t.sd.p, c->u.p, c->v.p, cnt);
    } else if (mixed) {
        hipLaunchKernelGGL((k_tree_diffusion<float, false>), dim3(lb8), dim3(256), 0, st, view(t), cut, c->prm.nu,
                           thr, t.sd.p, c->u.p, c->v.p, cnt);
    } else {
        hipLaunchKernelGGL((k_tree_diffusion<double, false>), dim3(lb8), dim3(256), 0, st, view(t), cut, c->prm.nu,
                           thr, t.sd.p, c->u.p, c->v.p, cnt);
    }
    VX_CHECK(hipGetLastError());
    return 0;
}

int launch_tree_diffusion(vx_case *c) {
    return tree_diffusion_pre(c) || tree_diffusion_far(c, c->stream) || tree_diffusion_main(c);
}

// The domain-scale target count the next diffusion sizes its list from (the
// last diffusion's, read after the caller's synchronisation); part of an
// exact checkpoint, since the list's capacity decides which targets the far
// kernel sums.
int64_t tree_far_hint(vx_case *c) {
    return (c->tree && c->tree->h_far) ? (int64_t) * (volatile unsigned long long *)c->tree->h_far : 0;
}

int tree_set_far_hint(vx_case *c, int64_t h) {
    TreeBufs &t = tree_of(c);
    if (!t.h_far) VX_CHECK(hipHostMalloc((void **)&t.h_far, sizeof(unsigned long long), hipHostMallocDefault));
    *t.h_far = (unsigned long long)(h > 0 ? h : 0);
    return 0;
}

int tree_copy_order(vx_case *c, uint32_t *keys, int32_t *perm, int64_t count) {
    if (!c->tree || count > c->tree->s) {
        set_error("bad argument: count beyond the slots of the last tree build");
        return 2;
    }
    if (count <= 0) return 0;
    TreeBufs &t = *c->tree;
    if (keys) VX_CHECK(hipMemcpyAsync(keys, t.keys_alt.p, sizeof(uint32_t) * count, hipMemcpyDeviceToHost, c->stream));
    if (perm) VX_CHECK(hipMemcpyAsync(perm, t.perm.p, sizeof(int32_t) * count, hipMemcpyDeviceToHost, c->stream));
    VX_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int tree_stats(vx_case *c, vx_tree_info *info) {
    *info = vx_tree_info{};
    if (!c->tree) return 0;
    unsigned long long h[16];
    VX_CHECK(hipMemcpyAsync(h, c->tree->cnt.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    VX_CHECK(hipStreamSynchronize(c->stream));
    info->m = c->tree->m;
    info->leaves = c->tree->nleaf;
    info->nodes = c->tree->nnode;
    info->fallback = (int64_t)h[1];
    info->bs_pairs = (int64_t)h[2];
    info->m2p_evals = (int64_t)h[3];
    info->diff_pairs = (int64_t)h[4];
    if (h[5]) {
        set_error("tree walk stack overflow");
        return 1;
    }
    info->order = c->tree->terms;
    const unsigned long long cap = (unsigned long long)c->tree->fcap;
    info->heavy = (int32_t)(h[6] < cap ? h[6] : cap);
    info->m2l = (int64_t)h[8];
    info->diff_in_radius = c->census ? (int64_t)h[9] : -1;
    return 0;
}

}  // namespace vx
