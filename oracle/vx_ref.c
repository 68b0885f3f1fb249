/*
 * vx_ref.c -- FP64 CPU restatement of the reference per-step formulas.
 * TEST INFRASTRUCTURE ONLY (see vx_ref.h).  Built with -ffp-contract=off.
 *
 * Each function follows the reference routine cited above it, in double
 * precision, keeping its loop order, its decision rules and its expression
 * association so that integer outcomes (absorption, compaction order,
 * collapse pairs) are reproducible bit for bit on identical inputs.
 */
#include "vx_ref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define VXR_PI 3.14159265358979323846
#define VXR_DELT 1e-12 /* include/definitions.h:36 */

int vxr_abi_version(void) { return 3; }
size_t vxr_sizeof_panel(void) { return sizeof(vxr_panel); }

int vxr_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void vxr_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

static inline double dot2(double a0, double a1, double b0, double b1) { return a0 * b0 + a1 * b1; }
static inline double dist2(double a0, double a1, double b0, double b1) {
    return (a0 - b0) * (a0 - b0) + (a1 - b1) * (a1 - b1);
}

/* ---- A1: src/unit_1.cu:256-373 ------------------------------------------ */
void vxr_panel_matrix(const vxr_panel *pan, int64_t P, double *L) {
    const int64_t m = P + 1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < P; ++i) {
        const double t0 = pan[i].tang[0], t1 = pan[i].tang[1];
        for (int64_t j = 0; j < P; ++j) {
            if (j == i) { L[i * m + j] = -0.5; continue; }
            /* endpoints of panel i (a) and panel j (b); for the right-hand
             * neighbour both panels are traversed reversed (:303-313) */
            double a00 = pan[i].left[0], a01 = pan[i].left[1];
            double a10 = pan[i].right[0], a11 = pan[i].right[1];
            double b00 = pan[j].left[0], b01 = pan[j].left[1];
            double b10 = pan[j].right[0], b11 = pan[j].right[1];
            const int next = (j == i + 1) || (j == 0 && i == P - 1);
            const int prev = (j == i - 1) || (i == 0 && j == P - 1);
            if (next) {
                a10 = pan[i].left[0]; a11 = pan[i].left[1];
                a00 = pan[i].right[0]; a01 = pan[i].right[1];
                b10 = pan[j].left[0]; b11 = pan[j].left[1];
                b00 = pan[j].right[0]; b01 = pan[j].right[1];
            }
            const double dj0 = b10 - b00, dj1 = b11 - b01;
            const double di0 = a10 - a00, di1 = a11 - a01;
            const double p10 = a00 - b10, p11 = a01 - b11;
            const double p20 = a10 - b10, p21 = a11 - b11;
            const double s10 = a00 - b00, s11 = a01 - b01;
            const double s20 = a10 - b00, s21 = a11 - b01;
            const double z1 = p10 * p21 - p11 * p20;
            const double z2 = s10 * s21 - s11 * s20;
            const double z3 = s20 * p21 - s21 * p20;
            double q11, q21;
            if (next || prev) {
                q11 = 0.0; q21 = 0.0;
            } else {
                q11 = atan(dot2(di0, di1, p10, p11) / z1) - atan(dot2(di0, di1, p20, p21) / z1);
                q21 = 0.5 * log(dot2(p20, p21, p20, p21) / dot2(p10, p11, p10, p11));
            }
            const double q12 = atan(dot2(di0, di1, s20, s21) / z2) - atan(dot2(di0, di1, s10, s11) / z2);
            const double q13 = atan(dot2(dj0, dj1, p20, p21) / z3) - atan(dot2(dj0, dj1, s20, s21) / z3);
            const double q22 = 0.5 * log(dot2(s10, s11, s10, s11) / dot2(s20, s21, s20, s21));
            const double q23 = 0.5 * log(dot2(p20, p21, p20, p21) / dot2(s20, s21, s20, s21));
            const double djp1 = dot2(dj0, dj1, p10, p11);
            const double dis1 = dot2(di0, di1, s10, s11);
            const double didj = dot2(di0, di1, dj0, dj1);
            const double c10 = djp1 * di0 + dis1 * dj0 - didj * s10;
            const double c11 = djp1 * di1 + dis1 * dj1 - didj * s11;
            const double djj = dot2(dj0, dj1, dj0, dj1);
            const double dii = dot2(di0, di1, di0, di1);
            const double c20 = c10 + djj * di0, c21 = c11 + djj * di1;
            const double c30 = dii * dj0, c31 = dii * dj1;
            const double k = 1.0 / (2 * VXR_PI * sqrt(djj) * dii);
            const double v0 = k * (q11 * c10 + q12 * c20 + q13 * c30 + (q21 * c11 + q22 * c21 + q23 * c31));
            const double v1 = k * (q11 * c11 + q12 * c21 + q13 * c31 - (q21 * c10 + q22 * c20 + q23 * c30));
            L[i * m + j] = dot2(v0, v1, t0, t1);
        }
    }
    for (int64_t i = 0; i < P; ++i) {
        L[P * m + i] = pan[i].length;
        L[i * m + P] = 1.0;
    }
    L[P * m + P] = 0.0;
}

/* ---- A2: src/unit_1.cu:115-254 ------------------------------------------ */
static void swap_rows(double *A, int64_t m, int64_t r0, int64_t r1) {
    if (r0 == r1) return;
    double *a = A + r0 * m, *b = A + r1 * m;
    for (int64_t j = 0; j < m; ++j) { double t = a[j]; a[j] = b[j]; b[j] = t; }
}

/* Standard Gauss-Jordan with partial pivoting (pivot of column k searched
 * after columns < k are eliminated).  Used where the reference's look-ahead
 * pivot rule hits a pivot below 1e-12 on a regular matrix (the bordered
 * system of a closed body has a singular P x P block). */
static int invert_pp(double *A, int64_t m, double *Minv) {
    for (int64_t k = 0; k < m; ++k) {
        double mx = fabs(A[k * m + k]);
        int64_t line = k;
        for (int64_t i = k + 1; i < m; ++i)
            if (fabs(A[i * m + k]) > mx) { mx = fabs(A[i * m + k]); line = i; }
        if (mx < VXR_DELT) return 1;
        swap_rows(A, m, k, line);
        swap_rows(Minv, m, k, line);
        const double *Ak = A + k * m, *Ik = Minv + k * m;
        const double piv = Ak[k];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < m; ++i) {
            if (i == k) continue;
            double *Ai = A + i * m, *Ii = Minv + i * m;
            const double c = Ai[k] / piv;
            for (int64_t j = 0; j < m; ++j) {
                Ai[j] = Ai[j] - c * Ak[j];
                Ii[j] = Ii[j] - c * Ik[j];
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) {
        const double di = A[i * m + i];
        for (int64_t j = 0; j < m; ++j) Minv[i * m + j] = Minv[i * m + j] / di;
    }
    return 0;
}

int vxr_invert_mode(double *A, int64_t m, double *Minv, int mode) {
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < m; ++j) Minv[i * m + j] = (i == j) ? 1.0 : 0.0;
    if (mode == 1) return invert_pp(A, m, Minv);
    return vxr_invert(A, m, Minv);
}

int vxr_invert(double *A, int64_t m, double *Minv) {
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < m; ++j) Minv[i * m + j] = (i == j) ? 1.0 : 0.0;
    if (fabs(A[0]) < VXR_DELT) {
        double mx = fabs(A[0]);
        int64_t num = 0;
        for (int64_t i = 1; i < m; ++i)
            if (fabs(A[i * m]) > mx) { mx = fabs(A[i * m]); num = i; }
        if (num != 0) { swap_rows(A, m, 0, num); swap_rows(Minv, m, 0, num); }
    }
    for (int64_t k = 0; k < m - 1; ++k) {
        if (fabs(A[k * m + k]) < VXR_DELT) return 1;
        /* the pivot for column k+1 is chosen before column k is eliminated */
        double mx = fabs(A[(k + 1) * m + k + 1]);
        int64_t line = k + 1;
        for (int64_t i = k + 1; i < m; ++i)
            if (fabs(A[i * m + k + 1]) > mx) { mx = fabs(A[i * m + k + 1]); line = i; }
        if (mx < VXR_DELT) return 1;
        swap_rows(A, m, k + 1, line);
        swap_rows(Minv, m, k + 1, line);
        const double *Ak = A + k * m, *Ik = Minv + k * m;
        const double piv = Ak[k];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < m; ++i) {
            if (i == k) continue;
            double *Ai = A + i * m, *Ii = Minv + i * m;
            const double c = Ai[k] / piv;
            for (int64_t j = 0; j < m; ++j) {
                Ai[j] = Ai[j] - c * Ak[j];
                Ii[j] = Ii[j] - c * Ik[j];
            }
        }
    }
    const double last = A[(m - 1) * m + m - 1];
    if (fabs(last) < VXR_DELT) return 1;
    const double *Il = Minv + (m - 1) * m;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m - 1; ++i) {
        const double c = A[i * m + m - 1] / last;
        double *Ii = Minv + i * m;
        for (int64_t j = 0; j < m; ++j) Ii[j] = Ii[j] - c * Il[j];
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) {
        const double di = A[i * m + i];
        for (int64_t j = 0; j < m; ++j) Minv[i * m + j] = Minv[i * m + j] / di;
    }
    return 0;
}

/* ---- A3: src/kernel.cu:779-847 ------------------------------------------ */
void vxr_right_part(const double *x, const double *y, const double *g, int64_t n_src,
                    const vxr_panel *pan, int64_t P, double vinf_x, double vinf_y,
                    double *R_p) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < P; ++i) {
        const double al0 = pan[i].left[0], al1 = pan[i].left[1];
        const double ar0 = pan[i].right[0], ar1 = pan[i].right[1];
        const double t0 = pan[i].tang[0], t1 = pan[i].tang[1];
        const double d0 = ar0 - al0, d1 = ar1 - al1;
        const double kd0 = -d1, kd1 = d0;
        double y0 = 0.0, y1 = 0.0;
        for (int64_t j = 0; j < n_src; ++j) {
            const double s00 = al0 - x[j], s01 = al1 - y[j];
            const double s10 = ar0 - x[j], s11 = ar1 - y[j];
            const double z = d0 * s01 - d1 * s00;
            const double alpha = atan(dot2(s00, s01, d0, d1) / z) - atan(dot2(s10, s11, d0, d1) / z);
            const double beta = 0.5 * log(dot2(s10, s11, s10, s11) / dot2(s00, s01, s00, s01));
            y0 += g[j] * (alpha * d0 + beta * kd0);
            y1 += g[j] * (alpha * d1 + beta * kd1);
        }
        const double mnog = 1 / (2 * VXR_PI * dot2(d0, d1, d0, d1));
        R_p[i] = -((mnog * y0 + vinf_x) * t0 + (mnog * y1 + vinf_y) * t1);
    }
}

/* ---- A4: src/kernel.cu:849-873, include/kernel.cuh:109-116 --------------- */
void vxr_birth(const double *M, int64_t P, const double *R_p, double d_g,
               const vxr_panel *pan, double birth_offset_rel,
               double *xo, double *yo, double *go, double *gsheet) {
    const int64_t m = P + 1;
    double *gs = (double *)malloc(sizeof(double) * (size_t)P);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < P; ++i) {
        double acc = 0.0;
        for (int64_t j = 0; j < P; ++j) acc += M[i * m + j] * R_p[j];
        acc += M[i * m + P] * d_g;
        gs[i] = acc;
    }
    for (int64_t i = 0; i < P; ++i) {
        const int64_t nx = pan[i].n_of_rpanel;
        const double off = birth_offset_rel > 0.0 ? birth_offset_rel * pan[i].length : 1e-6;
        xo[i] = (pan[i].left[0] + pan[i].right[0]) / 2 + off * pan[i].norm[0];
        yo[i] = (pan[i].left[1] + pan[i].right[1]) / 2 + off * pan[i].norm[1];
        go[i] = 0.5 * (gs[i] * pan[i].length + gs[nx] * pan[nx].length);
        if (gsheet) gsheet[i] = gs[i];
    }
    free(gs);
}

/* ---- A6: src/kernel.cu:875-983 ------------------------------------------ */
void vxr_biot_savart(const double *x, const double *y, const double *g,
                     int64_t n_tgt, int64_t n_src, double eps2,
                     double vinf_x, double vinf_y,
                     double *u, double *v, double *d) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_tgt; ++i) {
        const double a0 = x[i], a1 = y[i];
        double y0 = 0.0, y1 = 0.0;
        double d1 = 1e5, d2 = 1e5, d3 = 1e5; /* sentinel of :922-924 */
        for (int64_t j = 0; j < n_src; ++j) {
            if (j == i) continue;
            double r2 = dist2(a0, a1, x[j], y[j]);
            if (d3 > r2) { /* keep the three smallest, sorted (:943-956) */
                d3 = r2;
                double t = fmin(d3, d2); d3 = fmax(d3, d2); d2 = t;
                t = fmin(d1, d2); d2 = fmax(d1, d2); d1 = t;
            }
            r2 = fmax(r2, eps2);
            const double w = g[j] / r2;
            y1 += w * (a0 - x[j]);
            y0 += -w * (a1 - y[j]);
        }
        if (d) d[i] = (sqrt(d1) + sqrt(d2) + sqrt(d3)) / 3;
        u[i] = y0 / (2 * VXR_PI) + vinf_x;
        v[i] = y1 / (2 * VXR_PI) + vinf_y;
    }
}

/* ---- A19: src/kernel.cu:1434-1475 --------------------------------------- */
void vxr_velocity_probe(const double *x, const double *y, const double *g, int64_t n_src,
                        const double *px, const double *py, int64_t m, double eps2,
                        double vinf_x, double vinf_y, double *u, double *v) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) {
        const double a0 = px[i], a1 = py[i];
        double y0 = 0.0, y1 = 0.0;
        for (int64_t j = 0; j < n_src; ++j) {
            double r2 = dist2(a0, a1, x[j], y[j]);
            r2 = fmax(r2, eps2);
            const double w = g[j] / r2;
            y1 += w * (a0 - x[j]);
            y0 += -w * (a1 - y[j]);
        }
        u[i] = y0 / (2 * VXR_PI) + vinf_x;
        v[i] = y1 / (2 * VXR_PI) + vinf_y;
    }
}

/* ---- A7: src/kernel.cu:985-1029 ----------------------------------------- */
void vxr_diffusion(const double *x, const double *y, const double *g, const double *d,
                   int64_t n_tgt, int64_t n_src, double nu, double ve_size,
                   double *u, double *v) {
    const double thr = 0.001 * ve_size;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_tgt; ++i) {
        const double a0 = x[i], a1 = y[i], dd = d[i];
        double I20 = 0.0, I21 = 0.0, I1 = 0.0;
        for (int64_t j = 0; j < n_src; ++j) {
            const double dist = sqrt(dist2(a0, a1, x[j], y[j]));
            if (dist > thr) {
                const double w = g[j] / dist;
                const double e = exp(-dist / dd);
                I20 += -w * (a0 - x[j]) * e;
                I21 += -w * (a1 - y[j]) * e;
                I1 += g[j] * e;
            }
        }
        if (fabs(I1) > 1e-5) {
            u[i] += -nu * I20 / (I1 * dd);
            v[i] += -nu * I21 / (I1 * dd);
        }
    }
}

/* ---- A8: src/kernel.cu:1031-1088 and I_0_I_3 at :1300-1346 -------------- */
static void i0_i3(double Ra0, double Ra1, double Rb0, double Rb1, double N0, double N1,
                  double Rj0, double Rj1, double dL, double dd, int64_t N,
                  double *r0, double *r30, double *r31) {
    double s0 = 0.0, s30 = 0.0, s31 = 0.0;
    const double dR0 = (Rb0 - Ra0) / N, dR1 = (Rb1 - Ra1) / N;
    const double delt = dL / N;
    for (int64_t k = 0; k < N; ++k) {
        const double Rk0 = (Ra0 + k * dR0 + Ra0 + (k + 1) * dR0) / 2;
        const double Rk1 = (Ra1 + k * dR1 + Ra1 + (k + 1) * dR1) / 2;
        const double e0 = (Rj0 - Rk0) / dd, e1 = (Rj1 - Rk1) / dd;
        const double me = sqrt(e0 * e0 + e1 * e1);
        const double ex = exp(-me);
        s0 += (e0 * N0 + e1 * N1) / (e0 * e0 + e1 * e1) * (me + 1) * ex * delt;
        s30 += N0 * ex * delt;
        s31 += N1 * ex * delt;
    }
    *r0 = s0; *r30 = s30; *r31 = s31;
}

void vxr_diffusion_wall(const double *x, const double *y, const double *d, int64_t n_tgt,
                        const vxr_panel *pan, int64_t P, double nu, int64_t n_of_points,
                        double *u, double *v) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n_tgt; ++i) {
        const double a0 = x[i], a1 = y[i], dd = d[i];
        double II0 = 2 * VXR_PI * dd * dd;
        double II30 = 0.0, II31 = 0.0;
        for (int64_t f = 0; f < P; ++f) {
            const double dL = pan[f].length;
            /* the file's birth column, not the midpoint (include/kernel.cuh:141-146) */
            const double r = sqrt(dist2(a0, a1, pan[f].birth[0], pan[f].birth[1]));
            if ((r < 5 * dL) && (r > 0.1 * dL)) {
                double R0, R30, R31;
                i0_i3(pan[f].left[0], pan[f].left[1], pan[f].right[0], pan[f].right[1],
                      -pan[f].norm[0], -pan[f].norm[1], a0, a1, dL, dd, n_of_points,
                      &R0, &R30, &R31);
                II0 += (-dd) * R0;
                II30 -= R30;
                II31 -= R31;
            } else if (r <= 0.1 * dL) {
                const double ex = 1 - exp(-dL / (2 * dd));
                II0 = VXR_PI * dd * dd;
                II30 = 2 * pan[f].norm[0] * dd * ex;
                II31 = 2 * pan[f].norm[1] * dd * ex;
                break;
            }
        }
        u[i] += nu * II30 / II0;
        v[i] += nu * II31 / II0;
    }
}

/* ---- hitting(): src/kernel.cu:1357-1431 --------------------------------- */
int vxr_hitting(const vxr_panel *pan, int64_t P, const vxr_params *prm,
                double x1, double y1, double x2, double y2, int *pan_out) {
    const double porog = 1e-12;
    int minN = -1;
    double minDist = 25.0;
    int hit = !(((x1 < prm->px_min) && (x2 < prm->px_min)) ||
                ((x1 > prm->px_max) && (x2 > prm->px_max)) ||
                ((y1 < prm->py_min) && (y2 < prm->py_min)) ||
                ((y1 > prm->py_max) && (y2 > prm->py_max)));
    if (hit) {
        hit = 0;
        const double A = y2 - y1, B = x1 - x2, D = y1 * x2 - x1 * y2;
        for (int64_t i = 0; i < P; ++i) {
            double r0 = A * pan[i].left[0] + B * pan[i].left[1] + D;
            double r1 = A * pan[i].right[0] + B * pan[i].right[1] + D;
            if (fabs(r0) < porog) r0 = 0.0;
            if (fabs(r1) < porog) r1 = 0.0;
            if (r0 * r1 <= 0) {
                const double A1 = pan[i].right[1] - pan[i].left[1];
                const double B1 = pan[i].left[0] - pan[i].right[0];
                const double D1 = pan[i].left[1] * pan[i].right[0] - pan[i].left[0] * pan[i].right[1];
                double r2 = A1 * x1 + B1 * y1 + D1;
                double r3 = A1 * x2 + B1 * y2 + D1;
                if (fabs(r2) < porog) r2 = 0.0;
                if (fabs(r3) < porog) r3 = 0.0;
                if (r2 * r3 <= 0) {
                    hit = 1;
                    const double den = A * B1 - B * A1;
                    const double X = (B * D1 - D * B1) / den;
                    const double Y = (A1 * D - D1 * A) / den;
                    const double dd2 = (x2 - X) * (x2 - X) + (y2 - Y) * (y2 - Y);
                    if (dd2 < minDist) { minDist = dd2; minN = (int)i; }
                }
            }
        }
    }
    *pan_out = minN;
    /* the reference reads panels[-1] when every crossing is farther than 5;
     * that case is treated as no crossing here */
    return hit && minN >= 0;
}

/* ---- A9 + A10: src/kernel.cu:1090-1157 ---------------------------------- */
void vxr_step_move(double *x, double *y, double *g, const double *u, const double *v,
                   int64_t n, int64_t P, const vxr_panel *pan, const vxr_params *prm,
                   double *dg_i, double *fx_i, double *fy_i, double *m_i, double out4[4]) {
    double *dgv = dg_i ? dg_i : (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double *fxv = fx_i ? fx_i : (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double *fyv = fy_i ? fy_i : (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double *mv = m_i ? m_i : (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    const double dt = prm->dt;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        double fx = 0.0, fy = 0.0, mm = 0.0, dg = 0.0;
        if (n >= P && i >= n - P) { /* unsigned arithmetic in the reference */
            fx = g[i] * (-y[i]);
            fy = g[i] * (x[i]);
            mm = g[i] * dist2(x[i], y[i], prm->rc_x, prm->rc_y);
        }
        const double xn = x[i] + u[i] * dt;
        const double yn = y[i] + v[i] * dt;
        int hp = -1;
        if ((g[i] != 0) && vxr_hitting(pan, P, prm, xn, yn, x[i], y[i], &hp)) {
            fx -= g[i] * (-pan[hp].contr[1]);
            fy -= g[i] * (pan[hp].contr[0]);
            mm -= g[i] * dist2(pan[hp].contr[0], pan[hp].contr[1], prm->rc_x, prm->rc_y);
            dg = g[i];
            g[i] = 0;
        }
        x[i] = xn;
        y[i] = yn;
        if ((g[i] != 0) && ((x[i] > prm->x_max) || (x[i] < prm->x_min) ||
                            (y[i] > prm->y_max) || (y[i] < prm->y_min))) {
            x[i] = -2.0e+5;
            y[i] = -2.0e+5;
            g[i] = 0;
        }
        dgv[i] = dg; fxv[i] = fx; fyv[i] = fy; mv[i] = mm;
    }
    double sdg = 0.0, sfx = 0.0, sfy = 0.0, sm = 0.0;
    for (int64_t k = 0; k < n; ++k) { sdg += dgv[k]; sfx += fxv[k]; sfy += fyv[k]; sm += mv[k]; }
    out4[0] = sdg;
    out4[1] = sfx * (prm->rho / dt);
    out4[2] = sfy * (prm->rho / dt);
    out4[3] = sm * (prm->rho / (2 * dt));
    if (!dg_i) free(dgv);
    if (!fx_i) free(fxv);
    if (!fy_i) free(fyv);
    if (!m_i) free(mv);
}

/* ---- second component: src/unit_1.cu:1271-1340, src/kernel.cu:1478-1489 -- */
void vxr_points_move(double *x, double *y, double *g, const double *u, const double *v, int64_t m,
                     int64_t P, const vxr_panel *pan, const vxr_params *prm) {
    const double dt = prm->dt;
    for (int64_t i = 0; i < m; ++i) {
        const double xn = x[i] + u[i] * dt;
        const double yn = y[i] + v[i] * dt;
        int hp = -1;
        if ((g[i] != 0) && vxr_hitting(pan, P, prm, xn, yn, x[i], y[i], &hp)) g[i] = 0;
        x[i] = xn;
        y[i] = yn;
        if ((g[i] != 0) && ((x[i] > prm->x_max) || (x[i] < prm->x_min) ||
                            (y[i] > prm->y_max) || (y[i] < prm->y_min))) {
            x[i] = -2.0e+5;
            y[i] = -2.0e+5;
            g[i] = 0;
        }
    }
}

void vxr_points_relax(const double *g, const double *uenv, const double *venv, double *u, double *v,
                      int64_t m, double rel_t, double dt) {
    for (int64_t i = 0; i < m; ++i) {
        if (!g[i]) {
            u[i] = 0.0;
            v[i] = 0.0;
            continue;
        }
        const float pu = (float)u[i], pv = (float)v[i];
        u[i] = (uenv[i] - pu) / rel_t * dt + pu;
        v[i] = (venv[i] - pv) / rel_t * dt + pv;
    }
}

/* ---- A11: src/kernel.cu:1159-1178 --------------------------------------- */
int64_t vxr_compact(double *x, double *y, double *g, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        if (fabs(g[i]) < VXR_DELT) {
            const double r0 = x[i], r1 = y[i];
            g[i] = g[n - 1]; x[i] = x[n - 1]; y[i] = y[n - 1];
            g[n - 1] = 0; x[n - 1] = r0; y[n - 1] = r1;
            n--;
            i--;
        }
    }
    return n;
}

/* ---- A12/A13 search: src/kernel.cu:1180-1235 ---------------------------- */
void vxr_setka(const double *x, const double *y, const double *g, int64_t n, int kind,
               double r2, double max_ve_g, int32_t *COL) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        const double r0 = x[i], r1 = y[i], gi = g[i];
        int32_t c = -1;
        for (int64_t j = i + 1; j < n; ++j) {
            if (kind == 1) {
                if ((gi * g[j] < 0) && (dist2(r0, r1, x[j], y[j]) < r2)) { c = (int32_t)j; break; }
            } else {
                if ((gi * g[j] > 0) && (dist2(r0, r1, x[j], y[j]) < r2) && (j != i) &&
                    (fabs(gi + g[j]) < max_ve_g)) { c = (int32_t)j; break; }
            }
        }
        COL[i] = c;
    }
}

/* ---- A12/A13 merge: src/kernel.cu:1237-1271 ----------------------------- */
void vxr_collapse(double *x, double *y, double *g, int32_t *COL, int64_t n, int kind) {
    for (int64_t i = 0; i < n; ++i) {
        if (COL[i] > -1) {
            const int64_t j = COL[i];
            if (kind == 1) {
                const double mn = fabs(g[i]) + fabs(g[j]);
                x[i] = (x[i] * fabs(g[i]) + x[j] * fabs(g[j])) / fabs(mn);
                y[i] = (y[i] * fabs(g[i]) + y[j] * fabs(g[j])) / fabs(mn);
                g[i] = g[i] + g[j];
            } else {
                const double ng = g[i] + g[j];
                x[i] = (x[i] * g[i] + x[j] * g[j]) / ng;
                y[i] = (y[i] * g[i] + y[j] * g[j]) / ng;
                g[i] = ng;
            }
            g[j] = 0;
            x[j] = 1e+10;
            y[j] = 1e+10;
            COL[j] = -1;
        }
    }
}

/* ---- A14: src/unit_1.cu:1077-1218 --------------------------------------- */
void vxr_Step(double *x, double *y, double *g, const double *u, const double *v,
              int64_t *n, const vxr_panel *pan, int64_t P, const vxr_params *prm,
              double out4[4]) {
    vxr_step_move(x, y, g, u, v, *n, P, pan, prm, NULL, NULL, NULL, NULL, out4);
    *n = vxr_compact(x, y, g, *n);
    const double ve2 = prm->ve_size * prm->ve_size;
    const double r2d = prm->r_col_diff_sign * prm->r_col_diff_sign * ve2;
    const double r2s = prm->r_col_same_sign * prm->r_col_same_sign * ve2;
    int32_t *COL = (int32_t *)malloc(sizeof(int32_t) * (size_t)(*n > 0 ? *n : 1));
    for (int kind = 1; kind <= 2; ++kind) {
        for (int64_t cc = 0; cc < prm->n_col; ++cc) {
            vxr_setka(x, y, g, *n, kind, kind == 1 ? r2d : r2s, prm->max_ve_g, COL);
            vxr_collapse(x, y, g, COL, *n, kind);
            const int64_t nt = vxr_compact(x, y, g, *n);
            if (nt == *n) break;
            *n = nt;
        }
    }
    free(COL);
}

/* ---- A15: src/unit_1.cu:963-1075 ---------------------------------------- */
void vxr_Speed(const double *x, const double *y, const double *g, int64_t s,
               const vxr_panel *pan, int64_t P, const vxr_params *prm,
               double vinf_x, double vinf_y, double *u, double *v, double *d) {
    vxr_biot_savart(x, y, g, s, s, prm->ve_size * prm->ve_size, vinf_x, vinf_y, u, v, d);
    vxr_diffusion(x, y, g, d, s, s, prm->nu, prm->ve_size, u, v);
    vxr_diffusion_wall(x, y, d, s, pan, P, prm->nu, prm->n_of_points, u, v);
}

/* ---- A16: src/unit_1.cu:560-595 ----------------------------------------- */
void vxr_vort_creation(double *x, double *y, double *g, int64_t n, const double *M,
                       const vxr_panel *pan, int64_t P, const vxr_params *prm,
                       double vinf_x, double vinf_y, double d_g) {
    double *R_p = (double *)malloc(sizeof(double) * (size_t)P);
    const int64_t n_src = (n + 31) / 32 * 32; /* the 32-wide tile loop reads to ceil32(n) */
    vxr_right_part(x, y, g, n_src, pan, P, vinf_x, vinf_y, R_p);
    vxr_birth(M, P, R_p, d_g, pan, prm->birth_offset_rel, x + n, y + n, g + n, NULL);
    free(R_p);
}

/* ---- A18: src/main.c:652-785 -------------------------------------------- */
void vxr_time_step(double *x, double *y, double *g, double *u, double *v, double *d,
                   int64_t *n, const double *M, const vxr_panel *pan, int64_t P,
                   const vxr_params *prm, double vinf_x, double vinf_y,
                   double *d_g, double forces4[4]) {
    vxr_vort_creation(x, y, g, *n, M, pan, P, prm, vinf_x, vinf_y, *d_g);
    *n += P;
    const int64_t s = (*n + 31) / 32 * 32;
    vxr_Speed(x, y, g, s, pan, P, prm, vinf_x, vinf_y, u, v, d);
    vxr_Step(x, y, g, u, v, n, pan, P, prm, forces4);
    *d_g = forces4[0];
}

/* ---- Morton keys of the velocity tree (build-new, no reference line) ----- */
uint32_t vxr_spread16(uint32_t v) {
    /* bit b of v moves to bit 2b, one bit at a time */
    uint32_t r = 0;
    for (int b = 0; b < 16; ++b) r |= ((v >> b) & 1u) << (2 * b);
    return r;
}

int64_t vxr_morton_keys(const double *x, const double *y, const double *g, int64_t s,
                        const double box[4], uint32_t *keys, double out_box[4]) {
    double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
    int64_t m = 0;
    char *in = (char *)malloc((size_t)(s > 0 ? s : 1));
    for (int64_t i = 0; i < s; ++i) {
        int inc = 0;
        if (isfinite(x[i]) && isfinite(y[i])) {
            const double ex = fmax(0.0, fmax(box[0] - x[i], x[i] - box[1]));
            const double ey = fmax(0.0, fmax(box[2] - y[i], y[i] - box[3]));
            inc = (g[i] != 0.0) || (ex * ex + ey * ey <= 400.0 * 400.0);
        }
        in[i] = (char)inc;
        if (inc) {
            ++m;
            x0 = fmin(x0, x[i]); x1 = fmax(x1, x[i]);
            y0 = fmin(y0, y[i]); y1 = fmax(y1, y[i]);
        }
    }
    if (!(x0 <= x1)) { x0 = 0.0; x1 = 1.0; y0 = 0.0; y1 = 1.0; }
    const double wx = x1 - x0, wy = y1 - y0;
    for (int64_t i = 0; i < s; ++i) {
        if (!in[i]) { keys[i] = 0xFFFFFFFFu; continue; }
        double fx = wx > 0 ? (x[i] - x0) / wx * 65534.0 : 0.0;
        double fy = wy > 0 ? (y[i] - y0) / wy * 65534.0 : 0.0;
        fx = fmin(fmax(fx, 0.0), 65534.0);
        fy = fmin(fmax(fy, 0.0), 65534.0);
        keys[i] = vxr_spread16((uint32_t)fx) | (vxr_spread16((uint32_t)fy) << 1);
    }
    free(in);
    if (out_box) { out_box[0] = x0; out_box[1] = x1; out_box[2] = y0; out_box[3] = y1; }
    return m;
}

/* Stable counting sort, one key byte at a time from the least significant. */
void vxr_stable_order(const uint32_t *keys, int64_t s, int32_t *perm) {
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(s > 0 ? s : 1));
    for (int64_t i = 0; i < s; ++i) perm[i] = (int32_t)i;
    for (int shift = 0; shift < 32; shift += 8) {
        int64_t cnt[257] = {0};
        for (int64_t i = 0; i < s; ++i) cnt[((keys[perm[i]] >> shift) & 255u) + 1]++;
        for (int b = 0; b < 256; ++b) cnt[b + 1] += cnt[b];
        for (int64_t i = 0; i < s; ++i) tmp[cnt[(keys[perm[i]] >> shift) & 255u]++] = perm[i];
        memcpy(perm, tmp, sizeof(int32_t) * (size_t)s);
    }
    free(tmp);
}
