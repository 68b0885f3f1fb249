// Internal declarations shared by the HIP translation units of libvortex_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "vortex.h"

namespace vx {

constexpr double kPi = 3.14159265358979323846;
constexpr double kDelt = 1e-12;        // include/definitions.h:36
constexpr int kBlock = 256;

void set_error(const std::string &msg);

#define VX_CHECK(call)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            ::vx::set_error(std::string(hipGetErrorString(e_)) + " (" __FILE__ ":" +     \
                            std::to_string(__LINE__) + ")");                             \
            return 1;                                                                    \
        }                                                                                \
    } while (0)

#define VX_ARG(cond, msg)                                                                \
    do {                                                                                 \
        if (!(cond)) {                                                                   \
            ::vx::set_error(std::string("bad argument: ") + (msg));                      \
            return 2;                                                                    \
        }                                                                                \
    } while (0)

struct TreeBufs;
struct LuBufs;

// A vortex count for a launch.  Host form (dev == nullptr): the value ub.
// Device form: *dev read by the kernel, rounded up to a multiple of pad and
// clamped to ub, the host's upper bound that sized the grid; this is what a
// captured graph uses, since its launches cannot take per-step host values.
struct Cnt {
    const int64_t *dev;
    int64_t ub;
    int32_t pad;
    __host__ __device__ __forceinline__ int64_t get() const {
#if defined(__HIP_DEVICE_COMPILE__)
        if (!dev) return ub;
        const int64_t v = *dev;
        const int64_t r = pad > 1 ? (v + pad - 1) / pad * pad : v;
        return r < ub ? r : ub;
#else
        return ub;
#endif
    }
};
inline Cnt host_cnt(int64_t n) { return Cnt{nullptr, n, 1}; }

// Panels on the device as structure-of-arrays (one allocation).
struct PanelsSoA {
    double *lx, *ly, *rx, *ry, *cx, *cy, *bx, *by, *nx, *ny, *tx, *ty, *len;
    int32_t *rp;
    int64_t P;
};

// A run of consecutive panels for the birth right-hand side: centre and
// radius of a circle holding its panels (the box centre, the farthest
// endpoint), first panel and count.
struct RhsSeg {
    double cx, cy, r;
    int32_t first, count;
};

// Growable device buffer.  The first allocation takes an eighth more than
// asked and growth is geometric (at least 1.25x), so sizes that creep up
// from step to step (a vortex count near its start) reallocate O(log) times,
// and rarely early in a run.
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    int ensure(size_t want, hipStream_t s, bool keep = false) {
        if (want <= n) return 0;
        size_t cap = want < 1024 ? 1024 : want;
        if (!n) cap += cap / 8;
        if (n && cap < n + n / 4) cap = n + n / 4;
        T *q = nullptr;
        hipError_t e = hipMalloc((void **)&q, cap * sizeof(T));
        if (e != hipSuccess) {
            set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
            return 1;
        }
        if (keep && p && n) {
            e = hipMemcpyAsync(q, p, n * sizeof(T), hipMemcpyDeviceToDevice, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) {
                set_error(std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
                (void)hipFree(q);
                return 1;
            }
        }
        if (p) {
            (void)hipStreamSynchronize(s);
            (void)hipFree(p);
        }
        p = q;
        n = cap;
        return 0;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

}  // namespace vx

struct vx_case {
    vx_params prm;
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream (default priority) for work that overlaps the diffusion:
    // fork_ev orders it after the main stream, join_ev the main stream after it
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    // two more streams (default priority) for the independent chains of a
    // tree build: aux_fork orders them after the build's stream, aux_done[k]
    // the build's stream after stream k's chain
    hipStream_t aux[2] = {nullptr, nullptr};
    hipEvent_t aux_fork = nullptr, aux_done[2] = {nullptr, nullptr};
    bool wall_pending = false;                        // a wall pass forked and not yet joined
    // Within vx_time_step, the tree build of the speed stage starts on the
    // side stream while the birth stage runs on the main stream: pre_ev marks
    // the state before the birth, born_ev the newborns written, born_at their
    // first slot (-1: no such overlap pending).
    bool overlap_births = false;
    hipEvent_t pre_ev = nullptr, born_ev = nullptr;
    int64_t born_at = -1;
    int64_t tree_early_s = 0;                         // the size tree_build_early started
    vx::PanelsSoA pan{};
    void *pan_block = nullptr;
    vx::RhsSeg *segs = nullptr;                       // birth RHS segments (device)
    double *bpos = nullptr;                           // newborn positions: x [0, P), y [P, 2P) (device)
    bool births_near = false;                         // every newborn position within the tree's margin
    int32_t *seg_of = nullptr;                        // segment of each panel (device)
    int64_t nseg = 0;
    int64_t P = 0;
    double4 wall_box{};             // birth-point box grown by 5 max(length)
    // panel system
    vx::DBuf<double> L, M;          // (P+1)^2
    bool have_matrix = false;       // the births can be solved (inverse or LU)
    bool have_inverse = false;      // M holds the inverse
    bool census = false;            // tree diffusion also counts in-radius pairs
    // vortex state, FP64 SoA
    int64_t n = 0, cap = 0;
    double d_g = 0.0;               // circulation absorbed by the last Step (host copy)
    vx::DBuf<double> x, y, g, u, v, d;
    // FP32 copies of the sources for the mixed precision pairwise stages
    vx::DBuf<float> xf, yf, gf;
    // scratch
    vx::DBuf<double> rp, gsheet, partial, red;       // RHS, sheet, split-K and reductions
    vx::DBuf<double> rhs_s;                           // RHS far-field expansion sums
    vx::DBuf<double> tx, ty, tg;                      // compaction/collapse staging
    vx::DBuf<int32_t> flags, scan, lists, col, misc;
    vx::DBuf<uint64_t> keys, keys_alt;
    vx::DBuf<int32_t> vals, vals_alt;
    vx::DBuf<int32_t> cnode;                           // step collapse grid: node links, then pool slots
    vx::DBuf<int32_t> ptmp;                            // tile totals of the scans and selections (main stream)
    vx::DBuf<unsigned char> gridp;                     // collapse grid parameters (device)
    vx::DBuf<int64_t> dcount;                         // device counters
    vx::DBuf<int32_t> colseg;                         // collapse candidate counts per list segment
    vx::DBuf<int32_t> wlist;                          // wall targets (side stream)
    vx::DBuf<int32_t> wtmp;                           // their selection's tile totals
    vx::DBuf<double> wstage;                          // their velocity terms, per list slot
    int dn_slot = 1;                                  // dcount slot of the step's current count (1 or 8)
    int64_t *h_count = nullptr;                       // pinned read-back slots (16)
    // pinned, mapped and coherent slots that kernels store into themselves
    // (system-scope stores), read by the host after an event: h_dev on the
    // host, d_dev the same memory on the device (16 slots)
    int64_t *h_dev = nullptr, *d_dev = nullptr;
    double *h_red = nullptr;                          // pinned reduction read-back
    // probes
    vx::DBuf<double> px, py, pu, pv, pg;
    // the per-step probes of vx_case_set_probes (velocity_control,
    // src/main.c:744): points, last velocities, their split-K partials
    vx::DBuf<double> prb_x, prb_y, prb_u, prb_v, prb_part;
    int64_t prb_m = 0;
    bool prb_in_speed = false;  // vx_time_step: the step's probes go over the tree in vx_speed
    // Morton tree of the last velocity stage
    vx::TreeBufs *tree = nullptr;
    // FP32 LU of the panel system with FP64 refinement (VX_FLAG_LU_REFINE)
    vx::LuBufs *lu = nullptr;
    // Device-count steps (vx_run_steps_graph): every launch takes the vortex
    // count from dcount[1] (Cnt device form) with grids sized for n_ub, the
    // absorbed circulation from red[0], and each step ends by recording its
    // forces and count in gr_ring (slot dcount[9]); the step is captured once
    // as a graph and replayed.
    bool devcount = false;
    int64_t n_ub = 0;
    // launch shapes (vx_case_autotune / vx_case_set_tuning)
    vx_tuning tune{0, 0, -1, 0, 0.0, 0.0, 0};
    vx::DBuf<double> gr_ring;                         // 4 forces + 1 count per step
    int64_t gr_cap = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int64_t graph_cap = -1;                           // capacity the graph was captured for
    double graph_vinf[2] = {0.0, 0.0};
    // timing
    bool timing = false;
    hipEvent_t ev[18] = {};
    vx_timings tm{};
};

namespace vx {

// Event pair on the case stream accumulating elapsed milliseconds (slot k
// uses events 2k, 2k+1); active only when timing is enabled on the case.
struct Timer {
    vx_case *c;
    double *acc;
    bool on;
    hipStream_t s;
    Timer(vx_case *cc, double *a, int slot, hipStream_t st = nullptr)
        : c(cc), acc(a), on(cc->timing), s(st ? st : cc->stream) {
        if (on) (void)hipEventRecord(c->ev[slot * 2], s);
        this->slot = slot;
    }
    int slot;
    void stop() {
        if (!on) return;
        (void)hipEventRecord(c->ev[slot * 2 + 1], s);
        (void)hipEventSynchronize(c->ev[slot * 2 + 1]);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, c->ev[slot * 2], c->ev[slot * 2 + 1]);
        *acc += ms;
        on = false;
    }
    ~Timer() { stop(); }
};


// ---- launchers (kernels in the .hip units) --------------------------------
// vx_state.hip
int launch_fill_tail(vx_case *c, int64_t from, int64_t to);
int side_stream(vx_case *c);  // creates the side stream and its events once
int side_fork(vx_case *c);  // the side stream waits for the work queued so far on the main stream
int side_join(vx_case *c);  // the main stream waits for the work queued so far on the side stream
int aux_streams(vx_case *c);  // creates aux[] and their events once
int launch_pack_f32(vx_case *c, int64_t count);
int launch_pack_f32(vx_case *c, Cnt count);
int launch_count_add(vx_case *c, int64_t *a, int64_t v);  // *a += v (stream-ordered)
int launch_step_record(vx_case *c);  // count back to slot 1; forces and count into gr_ring
// The counts of the current stage: host values, or in device-count mode the
// count slot dcount[1] (n, or rounded up to a multiple of 32: s).
inline Cnt cnt_n(const vx_case *c) {
    return c->devcount ? Cnt{c->dcount.p + 1, c->n_ub, 1} : host_cnt(c->n);
}
inline Cnt cnt_s(const vx_case *c) {
    const int64_t s_ub = (c->devcount ? c->n_ub : c->n) + 31;
    return c->devcount ? Cnt{c->dcount.p + 1, std::min(s_ub / 32 * 32, c->cap), 32}
                       : host_cnt(std::min((c->n + 31) / 32 * 32, c->cap));
}
// vx_pairwise.hip
int launch_biot_savart(vx_case *c, int64_t n_tgt, int64_t n_src, double vinf_x, double vinf_y);
int launch_biot_savart(vx_case *c, Cnt n_tgt, Cnt n_src, double vinf_x, double vinf_y);
int launch_diffusion(vx_case *c, int64_t n_tgt, int64_t n_src);
int launch_diffusion(vx_case *c, Cnt n_tgt, Cnt n_src);
int launch_right_part(vx_case *c, int64_t n_src, double vinf_x, double vinf_y);  // -> c->rp
int launch_right_part(vx_case *c, Cnt n_src, double vinf_x, double vinf_y);
int launch_probe(vx_case *c, int64_t m, int64_t n_src, double vinf_x, double vinf_y); // c->px.. -> c->pu,pv
// The probe kernels on explicit buffers: part holds the split-K partials (at
// most max_part doubles, max_part <= 0: as many as the split wants, see
// probe_split).
void probe_split(int64_t m, int64_t n_src, int64_t max_part, int64_t *chunk, int64_t *nchunk);
int launch_probe_to(vx_case *c, const double *px, const double *py, int64_t m, int64_t n_src, double vinf_x,
                    double vinf_y, double *pu, double *pv, double *part, int64_t max_part);
int launch_probe_to(vx_case *c, const double *px, const double *py, int64_t m, Cnt n_src, double vinf_x,
                    double vinf_y, double *pu, double *pv, double *part, int64_t max_part);
int launch_step_probes(vx_case *c, double vinf_x, double vinf_y);  // c->prb_x.. -> c->prb_u, prb_v
// The probes over the tree of the last launch_tree_biot_savart, on stream st
void tree_d_changed(vx_case *c);  // d was set by the caller: the tree's sorted copy is stale
int launch_tree_probes(vx_case *c, const double *px, const double *py, int64_t m, double vinf_x,
                       double vinf_y, double *pu, double *pv, hipStream_t st);
// The probes on caller-owned device buffers with the partials in the
// library's file-scope device scratch (no allocation; calls that overlap on
// one GPU race for it).
int launch_probe_noalloc(vx_case *c, const double *px, const double *py, int64_t m, double vinf_x, double vinf_y,
                         double *pu, double *pv);
// vx_wall.hip
// launch_wall_side queues the wall term on stream st (the main stream, or the
// side stream after side_fork), into per-target staging; launch_wall_add
// (main stream, after side_join if st was the side stream) adds the terms to
// u, v.  launch_diffusion_wall is the two in a row on the main stream.
int launch_wall_side(vx_case *c, int64_t n_tgt, hipStream_t st);
int launch_wall_side(vx_case *c, Cnt n_tgt, hipStream_t st);
int launch_wall_add(vx_case *c);
int launch_diffusion_wall(vx_case *c, int64_t n_tgt);
int launch_diffusion_wall(vx_case *c, Cnt n_tgt);
// vx_panel.hip
int launch_panel_matrix(vx_case *c);                   // -> c->L
int launch_invert(vx_case *c);                         // c->L -> c->M (L destroyed copy)
int launch_birth(vx_case *c, double d_g);              // c->rp -> sheet, newborns at [n, n+P)
const double *birth_sheet(vx_case *c);                 // the last birth's sheet strengths (device)
int launch_birth_positions(vx_case *c);                // -> c->bpos (once per case)
// vx_lu.hip: FP32 LU of c->L and its triangular inverses (once per case),
// then per birth the refined solve of L x = [R; d_g] into lu_solution(c)
int launch_lu_setup(vx_case *c);
int launch_lu_solve(vx_case *c, const double *R, double d_g, const double *d_g_dev = nullptr);
const double *lu_solution(vx_case *c);
int lu_info(vx_case *c, int32_t *iters, int32_t *converged, double *resid);
void lu_release(vx_case *c);
// vx_tree.hip
int tree_build(vx_case *c, int64_t s);
int tree_build_early(vx_case *c, int64_t s, int64_t at);  // part A on the side stream (birth overlap)
int launch_tree_biot_savart(vx_case *c, int64_t s, double vinf_x, double vinf_y);
int launch_tree_diffusion(vx_case *c);  // reuses the tree of the last launch_tree_biot_savart
// launch_tree_diffusion in three parts: pre (main stream), far (any stream
// ordered after pre), main (main stream); far and main write disjoint targets
int tree_diffusion_pre(vx_case *c);
int tree_diffusion_far(vx_case *c, hipStream_t st);
int tree_diffusion_main(vx_case *c);
int tree_stats(vx_case *c, vx_tree_info *info);
int tree_copy_order(vx_case *c, uint32_t *keys, int32_t *perm, int64_t count);
int64_t tree_far_hint(vx_case *c);
int tree_set_far_hint(vx_case *c, int64_t h);
void tree_release(vx_case *c);
// vx_step.hip
// The step stages run on a device-resident vortex count *dn (dcount[1]):
// n_ub is an upper bound of it (n never grows inside Step), used for grids.
int64_t *step_dn(vx_case *c);                          // the device count
int launch_set_dn(vx_case *c, int64_t n);              // *dn = n (stream-ordered)
int launch_step_move(vx_case *c);                      // sums and live box -> c->red[0..7]
int launch_points_move(vx_case *c, int64_t m);         // c->px, py, pg moved by c->pu, pv
// A collapse grid kept over the passes of one step (launch_step_collapse):
// a kernel that gives a slot new data (a merge's survivor, a compaction's
// filled hole) pushes a pool node for the slot at its new cell.
struct GridPush {
    const void *grid;                      // the step's Grid on the device
    int32_t *head, *next, *pslot;          // cell heads, node links, pool node -> slot
    int32_t *pcount, *err;                 // pool nodes used; error flag (|= 2 when full)
    int64_t base, cap;                     // first pool node id, pool capacity
};
// [0, *dn) compacted, new count -> *dn; with merged given, the identity when
// *merged == 0 (the count of candidates the preceding collapse pass listed)
int launch_compact(vx_case *c, int64_t n_ub, const int64_t *merged = nullptr, const GridPush *gp = nullptr);
inline int64_t *col_merged(vx_case *c) { return c->dcount.p + 11; }  // written by launch_collapse
int launch_live_box(vx_case *c);                       // box of g != 0 in [0, *dn) -> c->red[4..7]
// launch_setka -> c->col (box from c->red[4..7]); with prepare it also fills
// the collapse counts and candidate list, and launch_collapse is then called
// with prepared.  With after_pass (a pass of the same kind and its compaction
// just ran) it searches nothing if that pass left the count unchanged.
// Collapse error flag in dcount[3].
int launch_setka(vx_case *c, int kind, int64_t n_ub, bool prepare, bool after_pass = false);
// default targets per thread of the direct kernels for n_tgt targets (vx_pairwise.hip)
int direct_tpt_default(bool f32, int64_t n_tgt);
// With reset, the collapse work arrays are left prepared for the next pass.
int launch_collapse(vx_case *c, int kind, int64_t n_ub, bool prepared, const GridPush *gp = nullptr,
                    bool reset = false);
// Every collapse pass of a step (both kinds, n_col each, with their
// compactions), from 2^17 slots on over a grid kept from pass to pass.
// Collapse error flag in dcount[3] (bit 2: the grid's node pool was full).
int launch_step_collapse(vx_case *c, int64_t n_ub);

}  // namespace vx
