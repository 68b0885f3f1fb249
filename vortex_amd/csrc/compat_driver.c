/* compat_driver PROFILE STEPS [mixed|f64] [POINTS]
 *
 * A run loop written only against the reference's host API
 * (include/vortex_compat.h), in the order of src/main.c:652-785: grow the
 * arrays when needed, vort_creation, Speed, Step; prints one line per step:
 * "step n Fx Fy M".  With a POINTS file (Kadr format) the second component
 * runs too (second_speed, first_speed after birth, rel_t = 0.1) and the final
 * points are printed as "first|second x y g".  Used to check that the compat
 * entry points reproduce the native API on the same case. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "vortex_compat.h"

struct conf_t conf; /* the caller owns the configuration, as src/main.c:17 does */

static int load(const char *fn, tPanel **pan, size_t *P, size_t *birth) {
    FILE *f = fopen(fn, "r");
    char buf[255];
    if (!f || !fgets(buf, sizeof buf, f) || !fgets(buf, sizeof buf, f) || fscanf(f, "%zu", P) != 1) return 1;
    *birth = (size_t)(32 * ceil((double)*P / 32));
    *pan = (tPanel *)calloc(*birth, sizeof(tPanel));
    double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
    for (size_t i = 0; i < *P; ++i) {
        tPanel *p = *pan + i;
        if (fscanf(f, " %u %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %u %u", &p->n, &p->left[0], &p->left[1],
                   &p->right[0], &p->right[1], &p->contr[0], &p->contr[1], &p->birth[0], &p->birth[1], &p->norm[0],
                   &p->norm[1], &p->tang[0], &p->tang[1], &p->length, &p->n_of_lpanel, &p->n_of_rpanel) != 16)
            return 1;
        x0 = fmin(x0, p->left[0]); x1 = fmax(x1, p->left[0]);
        y0 = fmin(y0, p->left[1]); y1 = fmax(y1, p->left[1]);
    }
    fclose(f);
    conf.profile_x_min = x0; conf.profile_x_max = x1;
    conf.profile_y_min = y0; conf.profile_y_max = y1;
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    tPanel *pan_h = NULL;
    size_t birth = 0;
    memset(&conf, 0, sizeof conf);
    conf.dt = 0.01; conf.inc_step = 8192; conf.viscosity = 1e-3; conf.rho = 1.0;
    conf.v_inf_x = 1.0; conf.v_inf_y = 0.0; conf.n_of_points = 20;
    conf.x_min = -2; conf.x_max = 10; conf.y_min = -10; conf.y_max = 10;
    conf.n_col = 2; conf.r_col_diff_sign = 0.3; conf.r_col_same_sign = 0.3;
    if (load(argv[1], &pan_h, &conf.birth_quant, &birth)) { fprintf(stderr, "profile\n"); return 1; }
    conf.ve_size = pan_h[0].length;
    conf.max_ve_g = 2 * conf.ve_size;
    const int steps = atoi(argv[2]);
    if (argc > 3 && !strcmp(argv[3], "mixed")) vx_compat_set_precision(1);
    /* second component: passive (first) and inertial (second) copies */
    size_t m = 0;
    Vortex *p1_h = NULL, *p1_d = NULL, *p2_d = NULL;
    PVortex *v1_d = NULL, *v2_d = NULL, *venv_d = NULL;
    if (argc > 4) {
        FILE *f = fopen(argv[4], "r");
        if (!f || fscanf(f, "%zu", &m) != 1) { fprintf(stderr, "points\n"); return 1; }
        p1_h = (Vortex *)calloc(m, sizeof(Vortex));
        for (size_t i = 0; i < m; ++i)
            if (fscanf(f, " %*u %*f %lf %lf %*f %*f %*f %lf", &p1_h[i].r[0], &p1_h[i].r[1], &p1_h[i].g) != 3) {
                fprintf(stderr, "points\n");
                return 1;
            }
        fclose(f);
        conf.rel_t = 0.1;
        hipMalloc((void **)&p1_d, m * sizeof(Vortex));
        hipMalloc((void **)&p2_d, m * sizeof(Vortex));
        hipMemcpy(p1_d, p1_h, m * sizeof(Vortex), hipMemcpyHostToDevice);
        hipMemcpy(p2_d, p1_h, m * sizeof(Vortex), hipMemcpyHostToDevice);
        hipMalloc((void **)&v1_d, m * sizeof(PVortex));
        hipMalloc((void **)&v2_d, m * sizeof(PVortex));
        hipMalloc((void **)&venv_d, m * sizeof(PVortex));
        hipMemset(v1_d, 0, m * sizeof(PVortex));
        hipMemset(v2_d, 0, m * sizeof(PVortex));
        hipMemset(venv_d, 0, m * sizeof(PVortex));
    }
    if (init_device_conf_values()) return 1;
    TVars *M = matr_creation(pan_h, conf.birth_quant, birth);
    if (!M) { fprintf(stderr, "matrix\n"); return 1; }
    tPanel *pan_d = NULL; TVars *M_d = NULL, *dg_d = NULL, *Mom_d = NULL, *d_d = NULL;
    TVctr *vinf_d = NULL; PVortex *F_d = NULL, *V_d = NULL, *V_h = NULL; Vortex *pos_d = NULL, *pos_h = NULL;
    TVctr vinf = {conf.v_inf_x, conf.v_inf_y};
    const double zero[2] = {0, 0};
    hipMalloc((void **)&pan_d, birth * sizeof(tPanel));
    hipMemcpy(pan_d, pan_h, birth * sizeof(tPanel), hipMemcpyHostToDevice);
    hipMalloc((void **)&M_d, (birth + 1) * (birth + 1) * sizeof(TVars));
    hipMemcpy(M_d, M, (birth + 1) * (birth + 1) * sizeof(TVars), hipMemcpyHostToDevice);
    hipMalloc((void **)&vinf_d, sizeof(TVctr));
    hipMemcpy(vinf_d, vinf, sizeof(TVctr), hipMemcpyHostToDevice);
    hipMalloc((void **)&dg_d, sizeof(TVars)); hipMemcpy(dg_d, zero, sizeof(TVars), hipMemcpyHostToDevice);
    hipMalloc((void **)&Mom_d, sizeof(TVars));
    hipMalloc((void **)&F_d, sizeof(PVortex));
    size_t size = 0, n = 0;
    if (incr_vort_quant(&pos_h, &pos_d, &V_h, &V_d, &d_d, &size)) return 1;
    for (int step = 0; step < steps; ++step) {
        const size_t s = (size_t)(32 * ceil((double)(n + conf.birth_quant) / 32));
        if (s > size && incr_vort_quant(&pos_h, &pos_d, &V_h, &V_d, &d_d, &size)) return 1;
        if (vort_creation(pos_d, vinf_d, conf.birth_quant, birth, n, M_d, dg_d, pan_d)) return 1;
        n += conf.birth_quant;
        if (m) {
            current_step = step;
            if (second_speed(pos_d, vinf_d, n, p2_d, v2_d, venv_d, &m, pan_d)) return 1;
            if (first_speed(pos_d, vinf_d, n, p1_d, v1_d, &m, pan_d)) return 1;
        }
        if (Speed(pos_d, vinf_d, s, V_d, d_d, conf.viscosity, pan_d)) return 1;
        hipMemcpy(F_d, zero, sizeof(PVortex), hipMemcpyHostToDevice);
        hipMemcpy(dg_d, zero, sizeof(TVars), hipMemcpyHostToDevice);
        hipMemcpy(Mom_d, zero, sizeof(TVars), hipMemcpyHostToDevice);
        if (Step(pos_d, V_d, &n, s, dg_d, F_d, Mom_d, pan_d)) return 1;
        PVortex F; TVars Mo;
        hipMemcpy(&F, F_d, sizeof F, hipMemcpyDeviceToHost);
        hipMemcpy(&Mo, Mom_d, sizeof Mo, hipMemcpyDeviceToHost);
        printf("%d %zu %.17g %.17g %.17g\n", step, n, F.v[0], F.v[1], Mo);
    }
    for (int which = 0; which < 2 && m; ++which) {
        hipMemcpy(p1_h, which ? p2_d : p1_d, m * sizeof(Vortex), hipMemcpyDeviceToHost);
        for (size_t i = 0; i < m; ++i)
            printf("%s %.17g %.17g %.17g\n", which ? "second" : "first", p1_h[i].r[0], p1_h[i].r[1], p1_h[i].g);
    }
    free(M); free(pan_h); free(pos_h); free(V_h); free(p1_h);
    return 0;
}
