// Device-wide primitives for wave64 (gfx950): exclusive/inclusive scans,
// order-preserving selection (stream compaction) and an LSD radix sort of
// 32-bit keys with 32-bit values.
//
// Every primitive is reduce-then-scan over tiles of kTile items: one launch
// reduces each tile, one work-group scans the tile totals, one launch scans
// each tile again from its offset.  No work-group ever waits on another
// (no decoupled look-back), so a primitive on a second stream beside a long
// kernel is only slowed by sharing the CUs, never held up by predecessors
// that cannot be scheduled.
//
// Items of a tile are read striped (round r, thread t: item r * kT + t), so
// loads coalesce; the scan order is still the item order.  Flags (0/1) are
// scanned with one ballot and popcount per wave and round; general values
// with shuffles.
//
// Functor interfaces (all __device__, called only for i < n of the launch):
//   flag:   int operator()(int64_t i) const            -> 0 or 1
//   value:  int32_t operator()(int64_t i) const
//   writer: void operator()(int64_t i, int32_t pos, int flag) const  (select / flag scan)
//           void operator()(int64_t i, int32_t v) const              (value scan)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vx {
namespace prim {

constexpr int kT = 256;             // threads per work-group
constexpr int kIPT = 16;            // items per thread
constexpr int kTile = kT * kIPT;    // items per tile
constexpr int kW = kT / 64;         // waves per work-group

inline int64_t tiles_of(int64_t n) { return n > 0 ? (n + kTile - 1) / kTile : 0; }

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }
__device__ __forceinline__ uint64_t lanes_below() { return (1ull << lane_id()) - 1ull; }

struct OpSum {
    static constexpr int32_t id = 0;
    __device__ __forceinline__ int32_t operator()(int32_t a, int32_t b) const { return a + b; }
};
struct OpMax {
    static constexpr int32_t id = INT32_MIN;
    __device__ __forceinline__ int32_t operator()(int32_t a, int32_t b) const { return a > b ? a : b; }
};

template <class Op>
__device__ __forceinline__ int32_t wave_reduce(int32_t a, Op op) {
    for (int o = 32; o > 0; o >>= 1) a = op(a, __shfl_xor(a, o, 64));
    return a;
}

// Inclusive scan over the 64 lanes of a wave.
template <class Op>
__device__ __forceinline__ int32_t wave_inclusive(int32_t a, Op op) {
    const int lane = lane_id();
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t b = __shfl_up(a, o, 64);
        if (lane >= o) a = op(a, b);
    }
    return a;
}

// ---- flags: count, scan, select ------------------------------------------------
// Phase 1: flags per tile.
template <class F>
__global__ __launch_bounds__(kT) void k_flag_count(F f, int64_t n, int32_t *__restrict__ tile_tot) {
    __shared__ int32_t wc[kW];
    const int64_t base = (int64_t)blockIdx.x * kTile;
    int32_t c = 0;
#pragma unroll 4
    for (int r = 0; r < kIPT; ++r) {
        const int64_t i = base + r * kT + threadIdx.x;
        const int fl = i < n ? f(i) : 0;
        c += __popcll(__ballot(fl));
    }
    if (lane_id() == 0) wc[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t s = 0;
        for (int w = 0; w < kW; ++w) s += wc[w];
        tile_tot[blockIdx.x] = s;
    }
}

// Phase 2: exclusive scan of the tile totals in place, in one work-group of
// 1024 threads (each takes a contiguous run of tiles); the grand total goes
// to *total (converted to TT).
template <class TT, class Op = OpSum>
__global__ __launch_bounds__(1024) void k_tile_scan(int32_t *__restrict__ tot, int64_t ntiles, TT *__restrict__ total,
                                                    Op op = Op()) {
    __shared__ int32_t ws[16];
    const int64_t per = (ntiles + 1023) / 1024;
    const int64_t b = (int64_t)threadIdx.x * per, e = b + per < ntiles ? b + per : ntiles;
    int32_t s = Op::id;
    for (int64_t k = b; k < e; ++k) s = op(s, tot[k]);
    // exclusive scan of the per-thread sums over the work-group
    const int32_t inc = wave_inclusive(s, op);
    if (lane_id() == 63) ws[threadIdx.x >> 6] = inc;
    __syncthreads();
    if (threadIdx.x < 64) {
        const int32_t v = threadIdx.x < 16 ? ws[threadIdx.x] : Op::id;
        const int32_t vi = wave_inclusive(v, op);
        if (threadIdx.x < 16) ws[threadIdx.x] = vi;  // inclusive over waves
    }
    __syncthreads();
    const int w = threadIdx.x >> 6;
    int32_t pre = w > 0 ? ws[w - 1] : Op::id;
    const int32_t excl_lane = __shfl_up(inc, 1, 64);
    if (lane_id() > 0) pre = op(pre, excl_lane);
    for (int64_t k = b; k < e; ++k) {
        const int32_t v = tot[k];
        tot[k] = pre;
        pre = op(pre, v);
    }
    if (threadIdx.x == 1023 && total) *total = (TT)(ws[15]);
}

// The first value of a tile's run: with SUMS, tile holds the tiles' totals
// and the work-group reduces those before it (few tiles: each reads at most
// kFuseTiles of them), and the last work-group stores the grand total;
// else tile holds the tiles' exclusive offsets (k_tile_scan).  Integer sums
// and maxima do not depend on the order of reduction.
constexpr int64_t kFuseTiles = 2048;

// With GRAND, *grand also gets the reduction of every tile: with SUMS from
// the tiles' totals, else from *total (which k_tile_scan wrote).
template <bool SUMS, bool GRAND = false, class Op, class TT>
__device__ __forceinline__ int32_t tile_start(const int32_t *__restrict__ tile, int64_t nt, Op op, TT *total,
                                              int32_t *grand = nullptr) {
    if constexpr (!SUMS) {
        if constexpr (GRAND) *grand = (int32_t)*total;
        return tile[blockIdx.x];
    } else {
        __shared__ int32_t ws[2][kW];
        int32_t a = Op::id, b = Op::id;
        const int64_t lim = GRAND ? nt : (int64_t)blockIdx.x;
        for (int64_t k = threadIdx.x; k < lim; k += kT) {
            const int32_t v = tile[k];
            if (k < (int64_t)blockIdx.x) a = op(a, v);
            if constexpr (GRAND) b = op(b, v);
        }
        a = wave_reduce(a, op);
        if constexpr (GRAND) b = wave_reduce(b, op);
        if (lane_id() == 0) {
            ws[0][threadIdx.x >> 6] = a;
            ws[1][threadIdx.x >> 6] = b;
        }
        __syncthreads();
        int32_t pre = Op::id, all = Op::id;
#pragma unroll
        for (int q = 0; q < kW; ++q) {
            pre = op(pre, ws[0][q]);
            all = op(all, ws[1][q]);
        }
        if constexpr (GRAND) *grand = all;
        if (total && threadIdx.x == 0 && (int64_t)blockIdx.x == nt - 1) *total = (TT)op(pre, tile[blockIdx.x]);
        return pre;
    }
}

// Phase 3: each flagged item's rank among the flagged items (its output
// position); the writer sees every item i < n with its flag, and with GRAND
// also the count of all flagged items.
template <class F, class Wr, bool SUMS = false, class TT = int32_t, bool GRAND = false>
__global__ __launch_bounds__(kT) void k_flag_apply(F f, int64_t n, const int32_t *__restrict__ tile_off, Wr wr,
                                                   TT *total = nullptr) {
    __shared__ int32_t wc[2][kW];
    const int64_t base = (int64_t)blockIdx.x * kTile;
    const int w = threadIdx.x >> 6;
    int32_t grand = 0;
    int32_t run = tile_start<SUMS, GRAND>(tile_off, (n + kTile - 1) / kTile, OpSum(), total, &grand);
    for (int r = 0; r < kIPT; ++r) {
        const int64_t i = base + r * kT + threadIdx.x;
        const int fl = i < n ? f(i) : 0;
        const uint64_t bal = __ballot(fl);
        if (lane_id() == 0) wc[r & 1][w] = __popcll(bal);
        __syncthreads();
        int32_t off = run, all = 0;
#pragma unroll
        for (int q = 0; q < kW; ++q) {
            const int32_t v = wc[r & 1][q];
            off += q < w ? v : 0;
            all += v;
        }
        if (i < n) {
            if constexpr (GRAND)
                wr(i, off + (int32_t)__popcll(bal & lanes_below()), fl, grand);
            else
                wr(i, off + (int32_t)__popcll(bal & lanes_below()), fl);
        }
        run += all;
    }
}

// ---- general values: scan ---------------------------------------------------------
template <class V, class Op>
__global__ __launch_bounds__(kT) void k_value_reduce(V val, int64_t n, int32_t *__restrict__ tile_tot, Op op) {
    __shared__ int32_t wc[kW];
    const int64_t base = (int64_t)blockIdx.x * kTile;
    int32_t a = Op::id;
#pragma unroll 4
    for (int r = 0; r < kIPT; ++r) {
        const int64_t i = base + r * kT + threadIdx.x;
        if (i < n) a = op(a, val(i));
    }
    a = wave_reduce(a, op);
    if (lane_id() == 0) wc[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t s = Op::id;
        for (int w = 0; w < kW; ++w) s = op(s, wc[w]);
        tile_tot[blockIdx.x] = s;
    }
}

// INCL: the writer gets the inclusive scan, else the exclusive one.
template <class V, class Op, bool INCL, class Wr, bool SUMS = false, class TT = int32_t>
__global__ __launch_bounds__(kT) void k_value_apply(V val, int64_t n, const int32_t *__restrict__ tile_off, Op op,
                                                    Wr wr, TT *total = nullptr) {
    __shared__ int32_t wc[2][kW];
    const int64_t base = (int64_t)blockIdx.x * kTile;
    const int w = threadIdx.x >> 6;
    int32_t run = tile_start<SUMS>(tile_off, (n + kTile - 1) / kTile, op, total);
    for (int r = 0; r < kIPT; ++r) {
        const int64_t i = base + r * kT + threadIdx.x;
        const int32_t v = i < n ? val(i) : Op::id;
        const int32_t inc = wave_inclusive(v, op);
        if (lane_id() == 63) wc[r & 1][w] = inc;
        __syncthreads();
        int32_t pre = run, all = run;
#pragma unroll
        for (int q = 0; q < kW; ++q) {
            const int32_t t = wc[r & 1][q];
            if (q < w) pre = op(pre, t);
            all = op(all, t);
        }
        const int32_t ex = __shfl_up(inc, 1, 64);  // outside the branch: every lane takes part
        if (i < n) wr(i, INCL ? op(pre, inc) : (lane_id() > 0 ? op(pre, ex) : pre));
        run = all;
    }
}

// ---- host wrappers ---------------------------------------------------------------------
// tmp must hold tiles_of(n) int32 entries.  total (device) receives the count
// of flagged items (or the value scan's total), converted to TT.

// Up to kFuseTiles tiles the apply launch reduces the tile totals itself
// (two launches); beyond, k_tile_scan turns them into offsets first.
template <class F, class Wr, class TT>
inline hipError_t select_flags(F f, int64_t n, Wr wr, int32_t *tmp, TT *total, hipStream_t st) {
    const int64_t nt = tiles_of(n);
    if (nt > 0) hipLaunchKernelGGL(k_flag_count<F>, dim3((unsigned)nt), dim3(kT), 0, st, f, n, tmp);
    if (nt > 0 && nt <= kFuseTiles) {
        hipLaunchKernelGGL((k_flag_apply<F, Wr, true, TT>), dim3((unsigned)nt), dim3(kT), 0, st, f, n,
                           (const int32_t *)tmp, wr, total);
        return hipGetLastError();
    }
    // with no tiles this only writes a zero total
    hipLaunchKernelGGL((k_tile_scan<TT, OpSum>), dim3(1), dim3(1024), 0, st, tmp, nt, total, OpSum());
    if (nt > 0)
        hipLaunchKernelGGL((k_flag_apply<F, Wr>), dim3((unsigned)nt), dim3(kT), 0, st, f, n, (const int32_t *)tmp, wr,
                           (int32_t *)nullptr);
    return hipGetLastError();
}

// select_flags whose writer also takes the count of all flagged items,
// wr(i, pos, flag, count); total (device, not null) receives that count.
template <class F, class Wr>
inline hipError_t select_flags_counted(F f, int64_t n, Wr wr, int32_t *tmp, int32_t *total, hipStream_t st) {
    const int64_t nt = tiles_of(n);
    if (nt > 0) hipLaunchKernelGGL(k_flag_count<F>, dim3((unsigned)nt), dim3(kT), 0, st, f, n, tmp);
    if (nt > 0 && nt <= kFuseTiles) {
        hipLaunchKernelGGL((k_flag_apply<F, Wr, true, int32_t, true>), dim3((unsigned)nt), dim3(kT), 0, st, f, n,
                           (const int32_t *)tmp, wr, total);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((k_tile_scan<int32_t, OpSum>), dim3(1), dim3(1024), 0, st, tmp, nt, total, OpSum());
    if (nt > 0)
        hipLaunchKernelGGL((k_flag_apply<F, Wr, false, int32_t, true>), dim3((unsigned)nt), dim3(kT), 0, st, f, n,
                           (const int32_t *)tmp, wr, total);
    return hipGetLastError();
}

template <bool INCL, class V, class Op, class Wr, class TT>
inline hipError_t scan_values(V val, int64_t n, Op op, Wr wr, int32_t *tmp, TT *total, hipStream_t st) {
    const int64_t nt = tiles_of(n);
    if (nt == 0) return hipSuccess;
    hipLaunchKernelGGL((k_value_reduce<V, Op>), dim3((unsigned)nt), dim3(kT), 0, st, val, n, tmp, op);
    if (nt <= kFuseTiles) {
        hipLaunchKernelGGL((k_value_apply<V, Op, INCL, Wr, true, TT>), dim3((unsigned)nt), dim3(kT), 0, st, val, n,
                           (const int32_t *)tmp, op, wr, total);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((k_tile_scan<TT, Op>), dim3(1), dim3(1024), 0, st, tmp, nt, total, op);
    hipLaunchKernelGGL((k_value_apply<V, Op, INCL, Wr>), dim3((unsigned)nt), dim3(kT), 0, st, val, n,
                       (const int32_t *)tmp, op, wr, (int32_t *)nullptr);
    return hipGetLastError();
}

// ---- fill (instead of memsets in per-step code) --------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_fill(T *__restrict__ a, int64_t n, T v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = v;
}
template <typename T>
inline hipError_t fill(T *a, int64_t n, T v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int64_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)(nb < 4096 ? nb : 4096)), dim3(256), 0, st, a, n, v);
    return hipGetLastError();
}

// ---- common functors -----------------------------------------------------------------
template <class P>
struct PredFlag {  // a boolean predicate of an int32 index as a flag
    P p;
    __device__ int operator()(int64_t i) const { return p((int32_t)i) ? 1 : 0; }
};
template <class P>
inline PredFlag<P> pred_flag(P p) { return PredFlag<P>{p}; }
template <class P, class N>
struct PredFlagN {  // the predicate below a count read on the device (N::get())
    P p;
    N n;
    __device__ int operator()(int64_t i) const { return (i < n.get() && p((int32_t)i)) ? 1 : 0; }
};
template <class P, class N>
inline PredFlagN<P, N> pred_flag(P p, N n) { return PredFlagN<P, N>{p, n}; }

struct ArrayFlag {  // a[i] != 0
    const int32_t *a;
    __device__ int operator()(int64_t i) const { return a[i] != 0; }
};
struct ArrayValue {
    const int32_t *a;
    __device__ int32_t operator()(int64_t i) const { return a[i]; }
};
struct WriteScan {  // out[i] = exclusive count of flags before i (flag scan) / the scan value (value scan)
    int32_t *out;
    __device__ void operator()(int64_t i, int32_t pos, int) const { out[i] = pos; }
    __device__ void operator()(int64_t i, int32_t v) const { out[i] = v; }
};
struct WriteIndex {  // out[pos] = i for flagged items (ordered selection of indices)
    int32_t *out;
    __device__ void operator()(int64_t i, int32_t pos, int fl) const {
        if (fl) out[pos] = (int32_t)i;
    }
};

// ---- LSD radix sort: 32-bit keys, 32-bit values, 8-bit digits ---------------------------
// Four stable passes of one launch each, after one launch that counts every
// pass's digits over all the keys.  A pass takes tiles of kSortTile keys; a
// tile counts its digits (one LDS add per run of equal digits in a wave,
// found with 8 ballots), publishes the counts in its status words, ranks its
// keys in LDS by digit (the wave's earlier lanes with the same digit, plus
// the tile's earlier waves and rounds) and writes them out digit run by digit
// run.  A run's output offset is the global count of the smaller digits plus
// the digit's count in the earlier tiles, which the tile reads back from
// their status words: a word holds a tile's own count (kStAgg) or its count
// including every earlier tile (kStInc); the look-back adds own counts until
// it meets an inclusive one, then publishes its own inclusive count.
// Progress: a work-group takes its tile by ticket, so every tile it reads
// belongs to a work-group that took its ticket earlier and is resident, and
// a tile publishes its own counts before it reads any other tile's, so each
// read waits at most until an earlier tile has counted.  A read still
// unpublished after kSpinMax polls (which this argument excludes) sets *err,
// and the tile publishes anyway, so a failure ends as a reported error.
// The status words are read and written with relaxed atomics at agent scope:
// each carries its count itself, and no other data is passed through them.
// Stable: equal keys keep their input order.
constexpr int kRadixBits = 8;
constexpr int kBins = 1 << kRadixBits;
constexpr int kSortPasses = 32 / kRadixBits;
constexpr int kSortIPT = 8;                 // keys per thread per tile
constexpr int kSortTile = kT * kSortIPT;
constexpr int64_t kSpinMax = 1 << 20;
constexpr uint64_t kStAgg = 1ull << 62, kStInc = 2ull << 62, kStMask = 3ull << 62, kStCount = 0xffffffffull;

inline int64_t sort_tiles_of(int64_t n) { return n > 0 ? (n + kSortTile - 1) / kSortTile : 0; }
// int32 entries of scratch radix_sort_pairs needs for n keys: every pass's
// status words (64-bit, [pass][tile][digit]), digit counts ([pass][digit])
// and ticket
inline int64_t radix_tmp_size(int64_t n) {
    return 2 * (int64_t)kSortPasses * kBins * sort_tiles_of(n) + kSortPasses * kBins + kSortPasses;
}

// Lanes of this wave whose digit equals this lane's (among lanes with ok).
__device__ __forceinline__ uint64_t digit_peers(int d, bool ok) {
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < kRadixBits; ++b) {
        const uint64_t bb = __ballot((d >> b) & 1);
        peers &= ((d >> b) & 1) ? bb : ~bb;
    }
    return peers;
}

// Every pass's digit counts over all n keys, added into ghist (zeroed).
template <int DUMMY = 0>
__global__ __launch_bounds__(kT) void k_radix_counts(const uint32_t *__restrict__ key, int64_t n,
                                                     int32_t *__restrict__ ghist) {
    __shared__ int32_t h[kSortPasses][kBins];
    for (int p = 0; p < kSortPasses; ++p) h[p][threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    for (int r = 0; r < kSortIPT; ++r) {
        const int64_t i = base + r * kT + threadIdx.x;
        const bool ok = i < n;
        const uint32_t k = ok ? key[i] : 0u;
#pragma unroll
        for (int p = 0; p < kSortPasses; ++p) {
            const int d = (int)((k >> (p * kRadixBits)) & (kBins - 1));
            const uint64_t peers = digit_peers(d, ok);
            if (ok && (peers & lanes_below()) == 0) atomicAdd(&h[p][d], (int32_t)__popcll(peers));
        }
    }
    __syncthreads();
    for (int p = 0; p < kSortPasses; ++p) {
        const int32_t v = h[p][threadIdx.x];
        if (v) atomicAdd(&ghist[p * kBins + threadIdx.x], v);
    }
}

// One pass: kin/vin by the digit at shift into kout/vout.  ghist: this
// pass's digit counts; ticket: zeroed; status: ntiles x kBins zeroed words.
template <int DUMMY = 0>
__global__ __launch_bounds__(kT) void k_radix_onesweep(const uint32_t *__restrict__ kin,
                                                       const int32_t *__restrict__ vin, int64_t n, int shift,
                                                       const int32_t *__restrict__ ghist, int32_t *__restrict__ ticket,
                                                       uint64_t *__restrict__ status, uint32_t *__restrict__ kout,
                                                       int32_t *__restrict__ vout, int32_t *__restrict__ err) {
    __shared__ uint32_t sk[kSortTile];   // the tile sorted by digit
    __shared__ int32_t sv[kSortTile];
    __shared__ int32_t cnt[kBins];       // the tile's count of each digit
    __shared__ int32_t loc[kBins];       // first slot of each digit in sk
    __shared__ int32_t run[kBins];       // next free slot of each digit in sk
    __shared__ int32_t gbase[kBins];     // first output position of each digit's run
    __shared__ int32_t wcnt[kW][kBins];
    __shared__ int32_t ws[2][kW];
    __shared__ int32_t tile_s;
    const int w = threadIdx.x >> 6, dd = threadIdx.x;  // thread dd handles digit dd in the per-digit steps
    if (threadIdx.x == 0) tile_s = atomicAdd(ticket, 1);
    cnt[dd] = 0;
    __syncthreads();
    const int64_t tile = tile_s;
    const int64_t base = tile * kSortTile;
    uint32_t kr[kSortIPT];
    int32_t vr[kSortIPT];
#pragma unroll
    for (int r = 0; r < kSortIPT; ++r) {
        const int64_t i = base + r * kT + threadIdx.x;
        kr[r] = i < n ? kin[i] : 0u;
        vr[r] = i < n ? vin[i] : 0;
    }
#pragma unroll
    for (int r = 0; r < kSortIPT; ++r) {
        const bool ok = base + r * kT + threadIdx.x < n;
        const int d = (int)((kr[r] >> shift) & (kBins - 1));
        const uint64_t peers = digit_peers(d, ok);
        if (ok && (peers & lanes_below()) == 0) atomicAdd(&cnt[d], (int32_t)__popcll(peers));
    }
    __syncthreads();
    const int32_t c = cnt[dd];
    uint64_t *own = status + tile * kBins + dd;
    __hip_atomic_store(own, (tile == 0 ? kStInc : kStAgg) | (uint64_t)(uint32_t)c, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    // the digit's first slot in sk, and the global count of the smaller digits
    const int32_t g = ghist[dd];
    const int32_t inc = wave_inclusive(c, OpSum()), ginc = wave_inclusive(g, OpSum());
    if (lane_id() == 63) {
        ws[0][w] = inc;
        ws[1][w] = ginc;
    }
    __syncthreads();
    int32_t pre = inc - c, gpre = ginc - g;
    for (int k = 0; k < w; ++k) {
        pre += ws[0][k];
        gpre += ws[1][k];
    }
    loc[dd] = pre;
    run[dd] = pre;
    for (int r = 0; r < kSortIPT; ++r) {
        for (int k = 0; k < kW; ++k) wcnt[k][threadIdx.x] = 0;
        const bool ok = base + r * kT + threadIdx.x < n;
        const uint32_t k = kr[r];
        const int d = (int)((k >> shift) & (kBins - 1));
        const uint64_t peers = digit_peers(d, ok);
        const int32_t below = (int32_t)__popcll(peers & lanes_below());
        __syncthreads();  // wcnt cleared, run up to date
        if (ok && below == 0) wcnt[w][d] = (int32_t)__popcll(peers);  // the digit's first lane in this wave
        __syncthreads();
        if (ok) {
            int32_t pos = run[d] + below;
            for (int q = 0; q < w; ++q) pos += wcnt[q][d];
            sk[pos] = k;
            sv[pos] = vr[r];
        }
        __syncthreads();
        int32_t add = 0;
#pragma unroll
        for (int q = 0; q < kW; ++q) add += wcnt[q][threadIdx.x];
        run[threadIdx.x] += add;
    }
    // the digit's count in the earlier tiles
    int64_t excl = 0;
    if (tile > 0) {
        int64_t t = tile - 1, polls = 0;
        for (;;) {
            const uint64_t sw = __hip_atomic_load(status + t * kBins + dd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint64_t st = sw & kStMask;
            if (st == 0) {
                if (++polls > kSpinMax) {
                    if (err) atomicOr(err, 1);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            excl += (int64_t)(sw & kStCount);
            if (st == kStInc || t == 0) break;
            --t;
        }
        __hip_atomic_store(own, kStInc | (uint64_t)(uint32_t)(excl + c), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    gbase[dd] = (int32_t)(gpre + excl);
    __syncthreads();
    const int64_t m = n - base < kSortTile ? n - base : kSortTile;
    for (int j = threadIdx.x; j < m; j += kT) {
        const uint32_t k = sk[j];
        const int d = (int)((k >> shift) & (kBins - 1));
        const int64_t pos = (int64_t)gbase[d] + (j - loc[d]);
        kout[pos] = k;
        vout[pos] = sv[j];
    }
}

// Sorts (k0, v0) of n items; k1, v1 are the same size.  The result ends in
// k0, v0 (four passes).  tmp must hold radix_tmp_size(n) int32 entries (64-bit
// aligned); *err (if given) gets bit 1 when a look-back gave up.
inline hipError_t radix_sort_pairs(uint32_t *k0, uint32_t *k1, int32_t *v0, int32_t *v1, int64_t n, int32_t *tmp,
                                   int32_t *err, hipStream_t st) {
    const int64_t nt = sort_tiles_of(n);
    if (nt == 0) return hipSuccess;
    uint64_t *status = reinterpret_cast<uint64_t *>(tmp);
    int32_t *ghist = tmp + 2 * (int64_t)kSortPasses * kBins * nt;
    int32_t *ticket = ghist + kSortPasses * kBins;
    const hipError_t e = fill(tmp, radix_tmp_size(n), 0, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_radix_counts<0>, dim3((unsigned)nt), dim3(kT), 0, st, (const uint32_t *)k0, n, ghist);
    uint32_t *ka = k0, *kb = k1;
    int32_t *va = v0, *vb = v1;
    for (int pass = 0; pass < kSortPasses; ++pass) {
        hipLaunchKernelGGL(k_radix_onesweep<0>, dim3((unsigned)nt), dim3(kT), 0, st, (const uint32_t *)ka,
                           (const int32_t *)va, n, pass * kRadixBits, (const int32_t *)(ghist + pass * kBins),
                           ticket + pass, status + (int64_t)pass * kBins * nt, kb, vb, err);
        uint32_t *tk = ka; ka = kb; kb = tk;
        int32_t *tv = va; va = vb; vb = tv;
    }
    return hipGetLastError();
}

}  // namespace prim
}  // namespace vx
