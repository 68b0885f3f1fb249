/*
 * vortex.h -- C-ABI of the MI355X vortex-element time step (libvortex_amd.so).
 *
 * One vx_case holds one flow case: parameters, panels, the panel-system
 * inverse and the vortex state, all resident in HBM as structure-of-arrays,
 * plus its own HIP stream.  Every entry point returns 0 on success, 1 on a
 * HIP/runtime error and 2 on bad arguments (the reference's convention,
 * include/main.h:20-26 and src/unit_1.cu:459-461); vx_last_error() gives the
 * message of the last failure on the calling thread.
 *
 * Host arrays passed in or out are plain double/int pointers; panels are
 * arrays of the reference's 120-byte tPanel layout (vx_panel below).
 *
 * Entry points and the reference routine each one replaces
 * (grechkin-pogrebnyakov/vortex, path:line):
 *   vx_case_create            init_device_conf_values    src/unit_1.cu:1220-1252
 *                             + panel upload             src/main.c:541-552
 *   vx_case_build_matrix      matr_creation              src/unit_1.cu:256-402
 *   vx_case_set_matrix        load_matrix                src/unit_1.cu:404-456
 *   vx_case_reserve           incr_vort_quant / allocate_arrays / randomize_tail
 *                                                        src/unit_1.cu:458-558
 *   vx_vort_creation          vort_creation              src/unit_1.cu:560-595
 *   vx_speed                  Speed                      src/unit_1.cu:963-1075
 *   vx_step                   Step                       src/unit_1.cu:1077-1218
 *   vx_velocity_probe         velocity_control           src/unit_1.cu:1254-1269
 *   vx_time_step              one iteration of the loop  src/main.c:652-785
 * Stage-level entry points (vx_right_part, vx_birth, vx_biot_savart,
 * vx_diffusion, vx_diffusion_wall, vx_step_move, vx_compact, vx_setka,
 * vx_collapse) run one kernel group each, as the reference kernels of the
 * same names do (src/kernel.cu:779-1271), for parity testing and profiling.
 */
#ifndef VORTEX_AMD_H_
#define VORTEX_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VX_ABI_VERSION 2

/* Byte-identical to the reference tPanel (include/definitions.h:92-113). */
typedef struct vx_panel {
    unsigned n;
    double left[2];
    double right[2];
    double contr[2];
    double birth[2];
    double norm[2];
    double tang[2];
    double length;
    unsigned n_of_lpanel;
    unsigned n_of_rpanel;
} vx_panel;

/* Precision of the pairwise stages.  VX_PREC_F64: the reference method in
 * double precision throughout.  VX_PREC_MIXED: pairwise and far-field sums in
 * FP32 with per-tile FP64 accumulation; panel system, reductions, moves,
 * compaction and collapse stay FP64. */
enum { VX_PREC_F64 = 0, VX_PREC_MIXED = 1 };

/* The reference's device constants and conf fields (src/unit_1.cu:20-44,
 * include/definitions.h:115-142).  birth_offset_rel = 0 keeps the reference
 * birth point (panel midpoint + 1e-6 normal); > 0 places newborns at
 * midpoint + birth_offset_rel * length * normal. */
typedef struct vx_params {
    double dt, ve_size, nu, rho;
    double vinf_x, vinf_y;
    double x_min, x_max, y_min, y_max;          /* deletion box                  */
    double px_min, px_max, py_min, py_max;      /* profile box (left end points) */
    double r_col_diff_sign, r_col_same_sign;    /* collapse radii / ve_size      */
    double max_ve_g;
    double rc_x, rc_y;
    double birth_offset_rel;
    int64_t n_of_points;                        /* wall quadrature points        */
    int64_t n_col;                              /* collapse passes per kind      */
    int64_t inc_step;                           /* capacity growth quantum       */
    int32_t precision;                          /* VX_PREC_*                     */
    int32_t tree_min_n;                         /* Morton-tree velocities at N >= this (0: never) */
    double  theta;                              /* tree opening angle (0: 0.5)   */
    double  diff_cut;                           /* diffusion cut-off / d_i (0: 18 mixed, 30 FP64) */
    int32_t order;                              /* multipole terms, 16, 20 or 32 (0: 16) */
    int32_t flags;                              /* VX_FLAG_*                     */
} vx_params;

/* Recompute the nearest-3 distance of every target by the exact search
 * instead of only where the tree's near field cannot guarantee it. */
#define VX_FLAG_EXACT_NEAREST 1
/* Solve the panel system at every birth by FP32 LU with partial pivoting
 * (MFMA trailing updates, factors inverted once per case) plus FP64-residual
 * iterative refinement against the FP64 matrix, instead of multiplying by
 * the FP64 Gauss-Jordan inverse (src/unit_1.cu:115-254, src/kernel.cu:849-873).
 * vx_case_build_matrix and vx_case_set_matrix then factor the panel matrix
 * (vx_case_set_matrix keeps M only for vx_case_get_matrix). */
#define VX_FLAG_LU_REFINE 2

typedef struct vx_case vx_case;

int         vx_abi_version(void);
const char *vx_last_error(void);
int         vx_device_count(int *count);

/* ---- case lifetime and state ------------------------------------------- */
int vx_case_create(const vx_params *prm, const vx_panel *panels, int64_t P, vx_case **out);
int vx_case_destroy(vx_case *c);
int vx_case_build_matrix(vx_case *c);                        /* L on the GPU + Gauss-Jordan */
int vx_case_set_matrix(vx_case *c, const double *M, int64_t m); /* m = P+1, row-major      */
int vx_case_get_matrix(vx_case *c, double *M, int64_t m);
int vx_case_get_panel_matrix(vx_case *c, double *L, int64_t m);  /* L before inversion     */
/* With VX_FLAG_LU_REFINE: the last solve's refinement corrections, whether
 * it met the test ||r|| <= ||x|| ||L|| eps64 sqrt(P+1), and its
 * ||r|| / (||L|| ||x||) (infinity norms). */
int vx_case_solve_info(vx_case *c, int32_t *iters, int32_t *converged, double *resid_rel);
int vx_case_reserve(vx_case *c, int64_t capacity);
int vx_case_set_state(vx_case *c, const double *x, const double *y, const double *g, int64_t n);
/* Copies slots [0, count) (count may exceed n: the tail beyond n is included). */
int vx_case_get_state(vx_case *c, double *x, double *y, double *g, int64_t count);
int vx_case_get_velocity(vx_case *c, double *u, double *v, double *d, int64_t count);
int vx_case_set_velocity(vx_case *c, const double *u, const double *v, const double *d, int64_t count);
int64_t vx_case_n(const vx_case *c);
int64_t vx_case_capacity(const vx_case *c);
/* Sets the count only (raw): the slots from n on keep their contents, which
 * the stages read up to ceil32(n + P); vx_case_set_state refills them. */
int vx_case_set_n(vx_case *c, int64_t n);
double  vx_case_d_g(const vx_case *c);
int vx_case_set_d_g(vx_case *c, double d_g);
int vx_case_sync(vx_case *c);
/* Exact checkpoint (no reference counterpart: the reference restarts from a
 * Kadr text snapshot printed with %lf, src/main.c:190-203,558-605, which
 * rounds the state).  vx_case_save writes the case's n, capacity, d_g, every
 * slot of x, y, g (the tail beyond n takes part in the nearest-3 search), the
 * tree's domain-scale target hint and the caller's blob `user` (user_bytes,
 * may be 0) to `path` in native byte order.  vx_case_load restores them
 * into a case created with the same panels and parameters (P, precision,
 * dt, ve_size and nu are checked: error 2), so that its following steps
 * equal those of the saved case bit for bit; the blob goes to `user`
 * (capacity user_cap bytes, error 2 if smaller) and its size to
 * *user_bytes.  Velocities are not saved: every step recomputes them. */
int vx_case_save(vx_case *c, const char *path, const void *user, int64_t user_bytes);
int vx_case_load(vx_case *c, const char *path, void *user, int64_t user_cap, int64_t *user_bytes);

/* ---- per-step stages (device resident) ---------------------------------- */
int vx_vort_creation(vx_case *c, double vinf_x, double vinf_y);
int vx_speed(vx_case *c, double vinf_x, double vinf_y);
int vx_step(vx_case *c, double out4[4]);
int vx_time_step(vx_case *c, double vinf_x, double vinf_y, double forces4[4]);
int vx_run_steps(vx_case *c, int64_t steps, double vinf_x, double vinf_y, double *forces4_per_step);
/* vx_run_steps as replays of one HIP graph of the whole time step (birth,
 * probes, Biot-Savart, diffusion, wall, move, compaction, collapse) on the
 * direct path: every launch reads the vortex count on the device, the grids
 * are sized for the capacity (grown once up front for all the steps), and the
 * host waits once at the end instead of after every step as the reference
 * loop does (src/main.c:652-785).  The results equal vx_run_steps bit for bit.
 * The count must stay below tree_min_n for every step (else error 2).  The
 * graph is kept for the next call with the same capacity and free stream.
 * counts_per_step (optional) gets the vortex count after each step. */
int vx_run_steps_graph(vx_case *c, int64_t steps, double vinf_x, double vinf_y, double *forces4_per_step,
                       int64_t *counts_per_step);

/* ---- batches of independent cases ------------------------------------------
 * No counterpart in the reference, which runs one case per process (globals
 * at src/main.c:17-20, __constant__ parameters at src/unit_1.cu:20-44).  A
 * batch advances k cases of any panel and vortex counts together: each
 * case's graph step (as vx_run_steps_graph) is one branch of a single graph,
 * so one graph launch advances every case by one step and their kernels run
 * side by side.  Each case's results equal its own vx_run_steps bit for bit.
 * The caller keeps ownership of the cases and must not use them on other
 * calls while vx_batch_run_steps runs.  vinf holds 2 values per case;
 * forces4 (4 per step and case) and counts (one per step and case) are
 * case-major and optional. */
typedef struct vx_batch vx_batch;
int vx_batch_create(vx_case *const *cases, int64_t k, vx_batch **out);
int vx_batch_run_steps(vx_batch *b, int64_t steps, const double *vinf, double *forces4, int64_t *counts);
int vx_batch_destroy(vx_batch *b);
int vx_velocity_probe(vx_case *c, const double *px, const double *py, int64_t m,
                      double vinf_x, double vinf_y, double *u, double *v);
/* Probe points kept on the device (m = 0: none).  vx_time_step then computes
 * their velocities every step after the birth, from the same sources as
 * vx_velocity_probe, as the reference loop calls velocity_control
 * (src/main.c:744, src/unit_1.cu:1254-1269); vx_case_get_probes copies the
 * last step's m velocities. */
int vx_case_set_probes(vx_case *c, const double *px, const double *py, int64_t m);
int vx_case_get_probes(vx_case *c, double *u, double *v, int64_t m);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): vx_velocity_probe on caller-owned DEVICE arrays px, py
 * (inputs) and u, v (outputs) of m doubles, queued on the case stream with no
 * host wait (vx_case_sync waits).  Its split-K partials live in a file-scope
 * device array of 2^21 doubles (16 MiB per GPU, held until the process ends),
 * so m may be at most 2^20 (larger m: error 2) and no call allocates.  Every
 * call on one GPU shares that array: calls that overlap on one GPU race. */
int vx_velocity_probe_noalloc(vx_case *c, const double *px, const double *py, int64_t m,
                              double vinf_x, double vinf_y, double *u, double *v);

/* ---- second component (tracer points) ------------------------------------
 * The reference advects two copies of the points of second_points_file each
 * step (src/main.c:729-741): passive ones with the flow velocity
 * (first_speed, src/unit_1.cu:1271-1298) and inertial ones whose velocity
 * relaxes towards it with time constant rel_t (second_speed,
 * src/unit_1.cu:1300-1340).  The flow velocity at the points is
 * vx_velocity_probe.  Host arrays of m points.
 *   vx_points_move   step_Kernel on points (src/kernel.cu:1090-1143 with null
 *                    force pointers): x += u dt; absorbed by the profile -> g = 0;
 *                    out of the box -> (-2e5, -2e5), g = 0
 *   vx_points_relax  second_speed_Kernel (src/kernel.cu:1478-1489), host:
 *                    u = (u_env - u_prev) / rel_t * dt + u_prev, u_prev in FP32 */
int vx_points_move(vx_case *c, double *x, double *y, double *g, const double *u, const double *v, int64_t m);
int vx_points_relax(const double *g, const double *uenv, const double *venv, double *u, double *v, int64_t m,
                    double rel_t, double dt);

/* ---- stage-level entry points (parity tests, profiling) ------------------ */
int vx_right_part(vx_case *c, int64_t n_src, double vinf_x, double vinf_y, double *R_p);
int vx_birth(vx_case *c, const double *R_p, double d_g, double *gsheet);   /* writes [n, n+P) */
int vx_biot_savart(vx_case *c, int64_t n_tgt, int64_t n_src, double vinf_x, double vinf_y);
int vx_diffusion(vx_case *c, int64_t n_tgt, int64_t n_src);
int vx_diffusion_wall(vx_case *c, int64_t n_tgt);
/* Morton-tree forms of vx_biot_savart (targets = sources = [0, s)) and
 * vx_diffusion (on the tree of the preceding vx_tree_biot_savart); vx_speed
 * uses them when s >= tree_min_n.  vx_tree_stats reports the work of the
 * last tree launches. */
int vx_tree_biot_savart(vx_case *c, int64_t s, double vinf_x, double vinf_y);
int vx_tree_diffusion(vx_case *c);
typedef struct vx_tree_info {
    int64_t m, leaves, nodes;        /* particles in the tree, leaves, all nodes      */
    int64_t fallback;                /* targets whose nearest-3 used the exact search */
    int64_t bs_pairs, m2p_evals;     /* pairwise pairs and (target, far node)
                                        interactions, last BS; the FP64 path
                                        evaluates each by M2P                 */
    int64_t diff_pairs;              /* pairwise evaluations, last diffusion          */
    int32_t order;
    int32_t heavy;                   /* diffusion targets with a domain-scale radius  */
    int64_t m2l;                     /* far nodes translated into bucket local
                                        expansions (FP32 path; the FP64 path
                                        evaluates m2p_evals directly), last BS */
    int64_t diff_in_radius;          /* of the target-group kernel's diff_pairs,
                                        those within the target's cut-off radius
                                        (census on), else -1                  */
} vx_tree_info;
int vx_tree_stats(vx_case *c, vx_tree_info *info);
/* Census of the tree diffusion (no reference counterpart): with on != 0 the
 * following diffusions also count the evaluated pairs that lie within their
 * target's cut-off radius (vx_tree_info.diff_in_radius), at the cost of one
 * more pass over each staged leaf.  Results do not change. */
int vx_tree_census(vx_case *c, int on);
/* The last tree build's sorted Morton keys and slot permutation (its first
 * count entries, count <= the slots it keyed): keys[k] is the k-th smallest
 * key and perm[k] its slot; equal keys keep the slot order.  Excluded slots
 * (g = 0 farther than 400 from the deletion box) carry key 0xFFFFFFFF and
 * sort last.  No reference counterpart: the reference's tree
 * (src/kernel.cu:79-763) is a bisection kd-tree and is inactive. */
int vx_tree_order(vx_case *c, uint32_t *keys, int32_t *perm, int64_t count);
int vx_step_move(vx_case *c, double out4[4]);
int vx_compact(vx_case *c);
int vx_setka(vx_case *c, int kind, int32_t *COL);                           /* COL[n] out */
int vx_collapse(vx_case *c, int kind, const int32_t *COL);

/* ---- launch-shape tuning --------------------------------------------------
 * No counterpart in the reference (fixed 32-thread blocks, include/definitions.h:35).
 *   tpt_bs, tpt_diff  targets per thread of the direct Biot-Savart and
 *                     diffusion kernels: 1, 2, 4 or 8 (0: default, 4 in mixed
 *                     precision, 2 in FP64, halved while the launch would give
 *                     the GPU's 256 CUs fewer than two work-groups each).
 *                     Results do not depend on them.
 *   leaf_off          the tree's bucket and target-group depth: quadtree
 *                     levels above the cell that holds 64 vortices at the mean
 *                     density, 0..3 (-1: default 1).  It changes which pairs
 *                     are summed directly, within the tree's error bound.
 *   col_share_min     vortex count from which the collapse passes of a step
 *                     keep one cell grid from pass to pass, linking the
 *                     vortices a pass moves at their new cells (0: default
 *                     2^17; below it every pass builds its grid).  Results
 *                     do not depend on it.
 * vx_case_autotune times every choice on the case's current state (the
 * direct kernels one by one below tree_min_n, the whole velocity stage
 * above it), keeps the fastest and reports the default's and the chosen
 * configuration's time in milliseconds.  It overwrites u, v and d only. */
typedef struct vx_tuning {
    int32_t tpt_bs, tpt_diff;
    int32_t leaf_off;
    int32_t trials;                 /* out: configurations timed */
    double ms_default, ms_best;     /* out: timed stage, default and chosen */
    int64_t col_share_min;
} vx_tuning;
int vx_case_autotune(vx_case *c, double vinf_x, double vinf_y, vx_tuning *out);
int vx_case_set_tuning(vx_case *c, const vx_tuning *t);
int vx_case_get_tuning(vx_case *c, vx_tuning *t);

/* ---- timing of the last stages on the case stream (milliseconds) -------- */
typedef struct vx_timings {
    double creation_ms, speed_ms, step_ms;      /* cumulative, as the timings line */
    double bs_ms, diff_ms, wall_ms;             /* cumulative, per Speed kernel    */
    int64_t bs_launches;
    double tree_build_ms, bs_kernel_ms;         /* cumulative: tree build, BS kernel alone */
    double diff_kernel_ms;                      /* cumulative: tree diffusion kernel alone */
} vx_timings;
int vx_case_timings(vx_case *c, vx_timings *t);
int vx_case_enable_timing(vx_case *c, int on);

#ifdef __cplusplus
}
#endif
#endif /* VORTEX_AMD_H_ */
