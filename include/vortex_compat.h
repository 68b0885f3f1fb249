/*
 * vortex_compat.h -- the reference's host API, served by libvortex_amd.so.
 *
 * A caller written against grechkin-pogrebnyakov/vortex include/unit_1.cuh
 * (the L4 boundary between src/main.c and the device code) links against
 * libvortex_amd.so instead and keeps its call sequence.  Types and layouts
 * are the reference's (include/definitions.h:35-142, default build with the
 * tree field, so sizeof(Vortex) == 32): vortex arrays are caller-owned
 * device arrays of Vortex / PVortex, the panel-system inverse is the padded
 * (Pb+1)^2 row-major matrix.  The hidden inputs are the reference's globals:
 * `struct conf_t conf` is read from the caller (this library only provides a
 * weak default definition, which a caller's own definition replaces).
 *
 * Each entry point stages the caller's AoS arrays into the library's
 * structure-of-arrays state, runs the same kernels as include/vortex.h, and
 * writes the results back, with the reference's return convention (0 ok,
 * 1 runtime error, 2 bad arguments; NULL for the matrix routines).
 *
 * Entry point            replaces (path:line in the reference)
 *   matr_creation          include/unit_1.cuh:21, src/unit_1.cu:256-402
 *   load_matrix            include/unit_1.cuh:24, src/unit_1.cu:404-456
 *   allocate_arrays        include/unit_1.cuh:27, src/unit_1.cu:458-486
 *   randomize_tail         include/unit_1.cuh:29, src/unit_1.cu:488-509
 *   incr_vort_quant        include/unit_1.cuh:31, src/unit_1.cu:511-558
 *   vort_creation          include/unit_1.cuh:34, src/unit_1.cu:560-595
 *   start_timer            include/unit_1.cuh:38, src/unit_1.cu:597-603
 *   stop_timer             include/unit_1.cuh:40, src/unit_1.cu:605-613
 *   Speed                  include/unit_1.cuh:43, src/unit_1.cu:963-1075
 *   Step                   include/unit_1.cuh:46, src/unit_1.cu:1077-1218
 *   init_device_conf_values include/unit_1.cuh:48, src/unit_1.cu:1220-1252
 *   velocity_control       include/unit_1.cuh:51, src/unit_1.cu:1254-1269
 *   second_speed           include/unit_1.cuh:52, src/unit_1.cu:1300-1340
 *   first_speed            include/unit_1.cuh:53, src/unit_1.cu:1271-1298
 * The second-component calls write output/vels/VelNNNNNN_S.txt for the
 * caller's current_step as the reference's do (the directory must exist).
 */
#ifndef VORTEX_COMPAT_H_
#define VORTEX_COMPAT_H_

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double TVars;
typedef TVars TVctr[2];

/* The reference's default build carries the tree field (32 B); built with
 * NO_TREE it has none (24 B, include/definitions.h:61-69).  A NO_TREE caller
 * gets the 24-byte record here, and its init_device_conf_values() call
 * (which the reference's main makes before any other, src/main.c:446)
 * selects that layout for every later call that takes Vortex arrays. */
typedef struct Vortex {
    TVars r[2];
    TVars g;
#ifndef NO_TREE
    unsigned int tree_id;
#endif
} Vortex;

typedef struct PVortex {
    TVars v[2];
} PVortex;

typedef struct tPanel {
    unsigned int n;
    TVctr left, right, contr, birth, norm, tang;
    TVars length;
    unsigned int n_of_lpanel, n_of_rpanel;
} tPanel;

/* include/definitions.h:115-142 (tree fields present) */
struct conf_t {
    size_t steps, saving_step;
    TVars dt;
    size_t inc_step;
    TVars viscosity, rho;
    TVars v_inf_x;
    TVars v_inf_y;
    size_t n_of_points;
    TVars x_max, x_min, y_max, y_min;
    TVars profile_x_min;
    TVars profile_x_max;
    TVars profile_y_min;
    TVars profile_y_max;
    TVars ve_size;
    size_t n_col;
    TVars h_col_x, h_col_y;
    TVars rc_x;
    TVars rc_y;
    char pr_file[256];
    size_t birth_quant;
    TVars r_col_diff_sign, r_col_same_sign;
    uint8_t matrix_load;
    size_t v_inf_incr_steps;
    TVars max_ve_g;
    char log_file[256];
    char config_file[256];
    uint8_t log_level;
    char timings_file[256];
    char kadr_file[256];
    char second_points_file[256];
    TVars rel_t;
    uint8_t steady_flow;
    size_t tree_depth;
    TVars theta;
    uint8_t no_log_buf;
};

extern struct conf_t conf;
/* the reference's step counters (src/main.c:18-19), read by second_speed and
 * first_speed; the caller defines them as the reference's main.c does */
extern int current_step;
extern int first_step;

TVars *matr_creation(tPanel *panels, size_t s, size_t birth);
TVars *load_matrix(size_t *p);
int allocate_arrays(Vortex **p_host, Vortex **p_dev, PVortex **v_host, PVortex **v_dev, TVars **d_dev, size_t size);
int randomize_tail(Vortex **p_dev, size_t new_size, size_t increased);
int incr_vort_quant(Vortex **p_host, Vortex **p_dev, PVortex **v_host, PVortex **v_dev, TVars **d_dev, size_t *size);
int vort_creation(Vortex *pos, TVctr *V_infDev, size_t n_of_birth, size_t n_of_birth_BLOCK_S, size_t n, TVars *M_Dev,
                  TVars *d_g, tPanel *panels);
void start_timer(hipEvent_t *start, hipEvent_t *stop);
float stop_timer(hipEvent_t start, hipEvent_t stop);
int Speed(Vortex *pos, TVctr *v_inf, size_t s, PVortex *v, TVars *d, TVars nu, tPanel *panels);
int Step(Vortex *pos, PVortex *V, size_t *n, size_t s, TVars *d_g, PVortex *F_p, TVars *M, tPanel *panels);
int init_device_conf_values(void);
int velocity_control(Vortex *pos, TVctr *V_inf, size_t n, Vortex *Contr_points, PVortex *V, size_t n_contr);
int second_speed(Vortex *pos, TVctr *V_inf, size_t n, Vortex *second_pos, PVortex *v_second, PVortex *v_env,
                 size_t *n_second, tPanel *panels);
int first_speed(Vortex *pos, TVctr *V_inf, size_t n, Vortex *second_pos, PVortex *v_second, size_t *n_second,
                tPanel *panels);

/* Select the arithmetic of the compat entry points (VX_PREC_F64 by default). */
int vx_compat_set_precision(int precision);
/* init_device_conf_values for callers of the 24-byte NO_TREE Vortex (no
 * reference counterpart: the reference selects the layout at compile time);
 * a NO_TREE caller's init_device_conf_values() resolves to it. */
int vx_compat_init_no_tree(void);
#ifdef NO_TREE
#define init_device_conf_values() vx_compat_init_no_tree()
#endif

#ifdef __cplusplus
}
#endif
#endif /* VORTEX_COMPAT_H_ */
