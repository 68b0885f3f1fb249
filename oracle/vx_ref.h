/*
 * vx_ref.h -- FP64 CPU restatement of the reference vortex-element time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP library in
 * vortex_amd/csrc and the CPU baseline timed by bench.py.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path never calls it.
 *
 * Every function restates one reference routine in double precision, with the
 * reference's operation order and decision rules (compiled with
 * -ffp-contract=off so that comparisons and integer outcomes are reproducible
 * bit for bit).  Citations are path:line in grechkin-pogrebnyakov/vortex.
 *
 * Pinning (tests/test_oracle_*.py):
 *   vxr_panel_matrix / vxr_invert     circle sheet closed form, M.L = I,
 *                                     numpy.linalg.solve golden (P = 200),
 *                                     Gauss-Legendre quadrature of the
 *                                     sheet-averaged velocity
 *   vxr_right_part                    adaptive quadrature of the panel-averaged
 *                                     point-vortex velocity
 *   vxr_birth                         constraint row: sum of newborn g = d_g
 *   vxr_biot_savart                   antisymmetry invariants, vortex pairs,
 *                                     Thomson N-gon, brute-force nearest-3
 *   vxr_diffusion                     brute force on tiny sets (numpy)
 *   vxr_diffusion_wall                high-order quadrature of the same
 *                                     integrand; near-case closed form
 *   vxr_step_move                     integer-coordinate crossing cases,
 *                                     deletion box, impulse sums
 *   vxr_compact                       hole/donor closed form (bit-exact)
 *   vxr_setka / vxr_collapse          brute-force min-j selection + greedy,
 *                                     conservation of sum g and sum g r
 */
#ifndef VX_REF_H_
#define VX_REF_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Byte layout of the reference's tPanel (include/definitions.h:92-113). */
typedef struct vxr_panel {
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
} vxr_panel;

/* The reference's __constant__ block (src/unit_1.cu:20-44, set at :1220-1252)
 * plus the conf fields its host code reads.  birth_offset_rel = 0 keeps the
 * reference birth point (midpoint + 1e-6 normal, include/kernel.cuh:109-116);
 * > 0 places newborns at midpoint + birth_offset_rel * length * normal. */
typedef struct vxr_params {
    double dt, ve_size, nu, rho;
    double vinf_x, vinf_y;
    double x_min, x_max, y_min, y_max;          /* deletion box             */
    double px_min, px_max, py_min, py_max;      /* profile bbox (left ends) */
    double r_col_diff_sign, r_col_same_sign;    /* multiples of ve_size     */
    double max_ve_g;
    double rc_x, rc_y;
    double birth_offset_rel;
    int64_t n_of_points;
    int64_t n_col;
} vxr_params;

int  vxr_abi_version(void);
int  vxr_num_threads(void);
void vxr_set_num_threads(int n);
size_t vxr_sizeof_panel(void);

/* A1: panel influence matrix, (P+1)x(P+1) row-major (src/unit_1.cu:256-373). */
void vxr_panel_matrix(const vxr_panel *pan, int64_t P, double *L);
/* A2: Gauss-Jordan inverse with the reference's pivot order
 * (src/unit_1.cu:115-254).  Returns 0 ok, 1 singular (pivot < 1e-12).
 * A is destroyed.  Minv is (m x m) row-major. */
int  vxr_invert(double *A, int64_t m, double *Minv);
/* mode 0: vxr_invert; mode 1: standard partial pivoting (the fallback where
 * the reference rule fails on a regular matrix). */
int  vxr_invert_mode(double *A, int64_t m, double *Minv, int mode);

/* A3: right-hand side of the birth system (src/kernel.cu:779-847).
 * Sources are [0, n_src); R_p has P entries. */
void vxr_right_part(const double *x, const double *y, const double *g, int64_t n_src,
                    const vxr_panel *pan, int64_t P, double vinf_x, double vinf_y,
                    double *R_p);
/* A4: sheet strengths and newborns (src/kernel.cu:849-873).  M is (P+1)^2
 * with row stride P+1.  Writes P newborns into xo/yo/go and, if gsheet is
 * not NULL, the P sheet strengths g_i. */
void vxr_birth(const double *M, int64_t P, const double *R_p, double d_g,
               const vxr_panel *pan, double birth_offset_rel,
               double *xo, double *yo, double *go, double *gsheet);

/* A6: Biot-Savart + nearest-3 distance (src/kernel.cu:875-983).
 * Targets [0, n_tgt) are also sources; sources are [0, n_src). */
void vxr_biot_savart(const double *x, const double *y, const double *g,
                     int64_t n_tgt, int64_t n_src, double eps2,
                     double vinf_x, double vinf_y,
                     double *u, double *v, double *d);
/* A19: velocity at probe points, no self skip, no d (src/kernel.cu:1434-1475). */
void vxr_velocity_probe(const double *x, const double *y, const double *g, int64_t n_src,
                        const double *px, const double *py, int64_t m, double eps2,
                        double vinf_x, double vinf_y, double *u, double *v);
/* A7: vortex-vortex diffusive velocity, added to u,v (src/kernel.cu:985-1029). */
void vxr_diffusion(const double *x, const double *y, const double *g, const double *d,
                   int64_t n_tgt, int64_t n_src, double nu, double ve_size,
                   double *u, double *v);
/* A8: wall diffusive velocity, added to u,v (src/kernel.cu:1031-1088,1324-1346). */
void vxr_diffusion_wall(const double *x, const double *y, const double *d, int64_t n_tgt,
                        const vxr_panel *pan, int64_t P, double nu, int64_t n_of_points,
                        double *u, double *v);

/* A9 (+hitting): move, newborn impulse, absorption, deletion for i < n
 * (src/kernel.cu:1090-1143,1357-1431).  Per-vortex dg/F/M go to the
 * optional arrays; A10 sums them in index order (src/kernel.cu:1144-1157):
 * out3 = {d_g, Fx, Fy, M} scaled as the reference does. */
void vxr_step_move(double *x, double *y, double *g, const double *u, const double *v,
                   int64_t n, int64_t P, const vxr_panel *pan, const vxr_params *prm,
                   double *dg_i, double *fx_i, double *fy_i, double *m_i, double out4[4]);
/* hitting() on its own: returns 1 on crossing, *pan_out = nearest panel. */
int  vxr_hitting(const vxr_panel *pan, int64_t P, const vxr_params *prm,
                 double new_x, double new_y, double old_x, double old_y, int *pan_out);

/* Second component: step_Kernel on m tracer points with null force pointers
 * (src/unit_1.cu:1295,1334; src/kernel.cu:1090-1143): move, absorption
 * (g = 0), deletion; no impulse terms. */
void vxr_points_move(double *x, double *y, double *g, const double *u, const double *v, int64_t m,
                     int64_t P, const vxr_panel *pan, const vxr_params *prm);
/* second_speed_Kernel (src/kernel.cu:1478-1489): u = (u_env - u_prev)/rel_t*dt
 * + u_prev with u_prev read as float; 0 where g = 0. */
void vxr_points_relax(const double *g, const double *uenv, const double *venv, double *u, double *v,
                      int64_t m, double rel_t, double dt);

/* A11: swap-with-last removal of |g| < 1e-12 (src/kernel.cu:1159-1178). */
int64_t vxr_compact(double *x, double *y, double *g, int64_t n);

/* A12/A13 search: kind 1 = different signs (first_setka, src/kernel.cu:1208-1235),
 * kind 2 = same sign (second_setka, :1180-1206).  r2 is the squared radius. */
void vxr_setka(const double *x, const double *y, const double *g, int64_t n, int kind,
               double r2, double max_ve_g, int32_t *COL);
/* A12/A13 merge (src/kernel.cu:1237-1271): sequential in index order. */
void vxr_collapse(double *x, double *y, double *g, int32_t *COL, int64_t n, int kind);

/* A14: Step() orchestration (src/unit_1.cu:1077-1218).  n in/out.
 * out4 = {d_g, Fx, Fy, M}. */
void vxr_Step(double *x, double *y, double *g, const double *u, const double *v,
              int64_t *n, const vxr_panel *pan, int64_t P, const vxr_params *prm,
              double out4[4]);
/* A15: Speed() (src/unit_1.cu:963-1075): u,v,d over [0,s). */
void vxr_Speed(const double *x, const double *y, const double *g, int64_t s,
               const vxr_panel *pan, int64_t P, const vxr_params *prm,
               double vinf_x, double vinf_y, double *u, double *v, double *d);
/* A16: vort_creation() (src/unit_1.cu:560-595): RHS over [0, ceil32(n)),
 * births written at [n, n+P). */
void vxr_vort_creation(double *x, double *y, double *g, int64_t n, const double *M,
                       const vxr_panel *pan, int64_t P, const vxr_params *prm,
                       double vinf_x, double vinf_y, double d_g);

/* A18: one iteration of the time loop (src/main.c:652-785) from the state
 * after the previous Step.  Arrays must have capacity >= ceil32(n+P).
 * d_g is in/out (circulation absorbed by the previous Step); forces4 gets
 * {d_g, Fx, Fy, M} of this step's Step. */
void vxr_time_step(double *x, double *y, double *g, double *u, double *v, double *d,
                   int64_t *n, const double *M, const vxr_panel *pan, int64_t P,
                   const vxr_params *prm, double vinf_x, double vinf_y,
                   double *d_g, double forces4[4]);

/* Morton keys of the velocity tree (a build-new structure: the reference's
 * own tree, src/kernel.cu:79-763, is a bisection kd-tree, inactive in every
 * shipped conf).  A slot is included when its position is finite and g != 0
 * or it lies within 400 of the deletion box (box = {x_min, x_max, y_min,
 * y_max}); the included slots' bounding box (out_box, {x0, x1, y0, y1}; the
 * unit box when none is included) is mapped to [0, 65534] per axis and the
 * two 16-bit integers interleaved, x in the even bits.  Excluded slots get
 * 0xFFFFFFFF.  Returns the included count. */
uint32_t vxr_spread16(uint32_t v);
int64_t vxr_morton_keys(const double *x, const double *y, const double *g, int64_t s,
                        const double box[4], uint32_t *keys, double out_box[4]);
/* Stable order of the keys (ties keep the index order): perm[k] is the slot
 * of the k-th smallest key. */
void vxr_stable_order(const uint32_t *keys, int64_t s, int32_t *perm);

#ifdef __cplusplus
}
#endif
#endif /* VX_REF_H_ */
