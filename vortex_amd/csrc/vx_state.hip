// Capacity tail and precision packing of the vortex state.
#include "vx_internal.h"

namespace vx {

// Unused capacity slots hold g = 0 far away, as zero_Kernel does
// (src/kernel.cu:767-777: r = 2e5 + 2e5 * rand()).  The factor is a 31-bit
// splitmix64 hash of the slot index so that the tail is reproducible.
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

constexpr uint64_t kTailSeed = 12345;

__global__ void k_fill_tail(double *__restrict__ x, double *__restrict__ y, double *__restrict__ g,
                            int64_t from, int64_t to) {
    const int64_t i = from + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= to) return;
    const double h = (double)(splitmix64((uint64_t)i ^ kTailSeed) >> 33);
    const double r = 2.0e5 * h + 2.0e5;
    x[i] = r;
    y[i] = r;
    g[i] = 0.0;
}

int side_stream(vx_case *c) {
    if (!c->side) {
        VX_CHECK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        VX_CHECK(hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming));
        VX_CHECK(hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming));
        VX_CHECK(hipEventCreateWithFlags(&c->pre_ev, hipEventDisableTiming));
        VX_CHECK(hipEventCreateWithFlags(&c->born_ev, hipEventDisableTiming));
    }
    return 0;
}

int aux_streams(vx_case *c) {
    if (!c->aux[0]) {
        for (int k = 0; k < 2; ++k) {
            VX_CHECK(hipStreamCreateWithFlags(&c->aux[k], hipStreamNonBlocking));
            VX_CHECK(hipEventCreateWithFlags(&c->aux_done[k], hipEventDisableTiming));
        }
        VX_CHECK(hipEventCreateWithFlags(&c->aux_fork, hipEventDisableTiming));
    }
    return 0;
}

int side_fork(vx_case *c) {
    if (side_stream(c)) return 1;
    VX_CHECK(hipEventRecord(c->fork_ev, c->stream));
    VX_CHECK(hipStreamWaitEvent(c->side, c->fork_ev, 0));
    return 0;
}

int side_join(vx_case *c) {
    VX_CHECK(hipEventRecord(c->join_ev, c->side));
    VX_CHECK(hipStreamWaitEvent(c->stream, c->join_ev, 0));
    return 0;
}

int launch_fill_tail(vx_case *c, int64_t from, int64_t to) {
    if (to <= from) return 0;
    const int64_t cnt = to - from;
    hipLaunchKernelGGL(k_fill_tail, dim3((unsigned)((cnt + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       c->stream, c->x.p, c->y.p, c->g.p, from, to);
    VX_CHECK(hipGetLastError());
    return 0;
}

__global__ void k_pack_f32(const double *__restrict__ x, const double *__restrict__ y,
                           const double *__restrict__ g, float *__restrict__ xf,
                           float *__restrict__ yf, float *__restrict__ gf, Cnt n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n.get()) return;
    xf[i] = (float)x[i];
    yf[i] = (float)y[i];
    gf[i] = (float)g[i];
}

int launch_pack_f32(vx_case *c, Cnt count) {
    if (count.ub <= 0) return 0;
    if (c->xf.ensure((size_t)c->cap, c->stream) || c->yf.ensure((size_t)c->cap, c->stream) ||
        c->gf.ensure((size_t)c->cap, c->stream))
        return 1;
    hipLaunchKernelGGL(k_pack_f32, dim3((unsigned)((count.ub + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       c->stream, c->x.p, c->y.p, c->g.p, c->xf.p, c->yf.p, c->gf.p, count);
    VX_CHECK(hipGetLastError());
    return 0;
}

int launch_pack_f32(vx_case *c, int64_t count) { return launch_pack_f32(c, host_cnt(count)); }

// Device-count helpers for captured steps: *a += v; the step's count back to
// slot 1 and the step's forces and count into the ring (slot k = st[0]++).
__global__ void k_count_add(int64_t *__restrict__ a, int64_t v) { *a += v; }

__global__ void k_step_record(int64_t *__restrict__ dcount, int slot, const double *__restrict__ red,
                              double *__restrict__ ring, int64_t *__restrict__ nring, int64_t *__restrict__ k) {
    const int64_t n = dcount[slot];
    dcount[1] = n;
    const int64_t s = *k;
    for (int q = 0; q < 4; ++q) ring[4 * s + q] = red[q];
    nring[s] = n;
    *k = s + 1;
}

int launch_count_add(vx_case *c, int64_t *a, int64_t v) {
    hipLaunchKernelGGL(k_count_add, dim3(1), dim3(1), 0, c->stream, a, v);
    VX_CHECK(hipGetLastError());
    return 0;
}

int launch_step_record(vx_case *c) {
    hipLaunchKernelGGL(k_step_record, dim3(1), dim3(1), 0, c->stream, c->dcount.p, c->dn_slot, c->red.p,
                       c->gr_ring.p, (int64_t *)(c->gr_ring.p + 4 * c->gr_cap), c->dcount.p + 9);
    VX_CHECK(hipGetLastError());
    c->dn_slot = 1;
    return 0;
}

}  // namespace vx
