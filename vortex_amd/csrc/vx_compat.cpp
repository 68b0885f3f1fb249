// The reference's host API (include/vortex_compat.h) over the library's
// structure-of-arrays case: each call stages the caller's device AoS arrays
// in, runs the same kernels as include/vortex.h and writes the results back.
// Calls are synchronous, as the reference's are.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "vortex_compat.h"
#include "vx_internal.h"

// The caller (the run driver) owns the configuration and the step counters;
// these weak definitions only keep the library loadable on its own.
__attribute__((weak)) struct conf_t conf;
__attribute__((weak)) int current_step = 0;
__attribute__((weak)) int first_step = 0;

namespace vx {
namespace {

static_assert(sizeof(Vortex) == 32, "Vortex layout of the reference's default build (NO_TREE: 24 B, at run time)");
static_assert(sizeof(PVortex) == 16, "PVortex layout");
static_assert(sizeof(tPanel) == sizeof(vx_panel), "tPanel layout");

struct Compat {
    vx_case *c = nullptr;
    int64_t vb = sizeof(Vortex);  // the caller's Vortex: 32 B, or 24 B built with NO_TREE
    const tPanel *panels_dev = nullptr;
    size_t P = 0;
    int precision = VX_PREC_F64;
    const TVars *M_src = nullptr;
    bool dirty = true;
};
Compat g_compat;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Vortex records of vb bytes: r[0], r[1], g at offsets 0, 8, 16, and with
// vb = 32 the tree_id at 24 (include/definitions.h:61-69).
__global__ void k_aos_in(const char *__restrict__ pos, int64_t vb, int64_t cnt, double *__restrict__ x,
                         double *__restrict__ y, double *__restrict__ g) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    const double *p = reinterpret_cast<const double *>(pos + i * vb);
    x[i] = p[0];
    y[i] = p[1];
    g[i] = p[2];
}

__global__ void k_aos_out(const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ g,
                          int64_t from, int64_t to, char *__restrict__ pos, int64_t vb) {
    const int64_t i = from + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= to) return;
    double *p = reinterpret_cast<double *>(pos + i * vb);
    p[0] = x[i];
    p[1] = y[i];
    p[2] = g[i];
}

__global__ void k_pv_in(const PVortex *__restrict__ v, int64_t cnt, double *__restrict__ u, double *__restrict__ w) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    u[i] = v[i].v[0];
    w[i] = v[i].v[1];
}

__global__ void k_pv_out(const double *__restrict__ u, const double *__restrict__ w, int64_t cnt,
                         PVortex *__restrict__ v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    v[i].v[0] = u[i];
    v[i].v[1] = w[i];
}

// zero_Kernel (src/kernel.cu:767-777): g = 0 far away; the factor is a hash
// of the slot index instead of rand().
__global__ void k_aos_tail(char *__restrict__ pos, int64_t vb, int64_t from, int64_t to) {
    const int64_t i = from + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= to) return;
    const double h = (double)(mix64((uint64_t)i ^ 12345ull) >> 33);
    double *p = reinterpret_cast<double *>(pos + i * vb);
    p[0] = 2.0e5 * h + 2.0e5;
    p[1] = 2.0e5 * h + 2.0e5;
    p[2] = 0.0;
    if (vb >= 28) *reinterpret_cast<unsigned int *>(pos + i * vb + 24) = 0u;
}

// Host copies of Vortex records of the caller's layout.
struct HostPts {
    std::vector<char> raw;
    int64_t vb = 32;
    size_t m = 0;
    double *rec(size_t i) { return reinterpret_cast<double *>(raw.data() + i * vb); }
    const double *rec(size_t i) const { return reinterpret_cast<const double *>(raw.data() + i * vb); }
};
int read_pts(HostPts &h, const void *dev, size_t m) {
    h.vb = g_compat.vb;
    h.m = m;
    h.raw.assign(m * (size_t)h.vb, 0);
    if (m) VX_CHECK(hipMemcpy(h.raw.data(), dev, h.raw.size(), hipMemcpyDeviceToHost));
    return 0;
}

unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }
int64_t ceil32(int64_t n) { return (n + 31) / 32 * 32; }

vx_params params_from_conf(int precision) {
    vx_params p{};
    p.dt = conf.dt;
    p.ve_size = conf.ve_size;
    p.nu = conf.viscosity;
    p.rho = conf.rho;
    p.vinf_x = conf.v_inf_x;
    p.vinf_y = conf.v_inf_y;
    p.x_min = conf.x_min;
    p.x_max = conf.x_max;
    p.y_min = conf.y_min;
    p.y_max = conf.y_max;
    p.px_min = conf.profile_x_min;
    p.px_max = conf.profile_x_max;
    p.py_min = conf.profile_y_min;
    p.py_max = conf.profile_y_max;
    p.r_col_diff_sign = conf.r_col_diff_sign;
    p.r_col_same_sign = conf.r_col_same_sign;
    p.max_ve_g = conf.max_ve_g;
    p.rc_x = conf.rc_x;
    p.rc_y = conf.rc_y;
    p.n_of_points = (int64_t)conf.n_of_points;
    p.n_col = (int64_t)conf.n_col;
    p.inc_step = conf.inc_step ? (int64_t)conf.inc_step : 8192;
    p.precision = precision;
    p.tree_min_n = conf.tree_depth > 1 ? 1 : 0;
    return p;
}

// The library case behind the compat calls, rebuilt when the panels or the
// configuration change.
vx_case *compat_case(const tPanel *panels_dev) {
    Compat &G = g_compat;
    const size_t P = conf.birth_quant;
    if (G.c && !G.dirty && G.panels_dev == panels_dev && G.P == P) return G.c;
    if (!panels_dev || P < 2) {
        set_error("bad argument: panels / conf.birth_quant");
        return nullptr;
    }
    std::vector<vx_panel> host(P);
    if (hipMemcpy(host.data(), panels_dev, sizeof(vx_panel) * P, hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("panel read-back failed");
        return nullptr;
    }
    if (G.c) vx_case_destroy(G.c);
    G.c = nullptr;
    const vx_params prm = params_from_conf(G.precision);
    if (vx_case_create(&prm, host.data(), (int64_t)P, &G.c)) return nullptr;
    G.panels_dev = panels_dev;
    G.P = P;
    G.M_src = nullptr;
    G.dirty = false;
    return G.c;
}

int stage_in(vx_case *c, const Vortex *pos, int64_t cnt) {
    if (cnt <= 0) return 0;
    hipLaunchKernelGGL(k_aos_in, dim3(blocks(cnt)), dim3(256), 0, c->stream, (const char *)pos, g_compat.vb, cnt,
                       c->x.p, c->y.p, c->g.p);
    VX_CHECK(hipGetLastError());
    return 0;
}

int stage_out(vx_case *c, Vortex *pos, int64_t from, int64_t to) {
    if (to <= from) return 0;
    hipLaunchKernelGGL(k_aos_out, dim3(blocks(to - from)), dim3(256), 0, c->stream, c->x.p, c->y.p, c->g.p, from, to,
                       (char *)pos, g_compat.vb);
    VX_CHECK(hipGetLastError());
    return 0;
}

template <typename T>
int read_dev(hipStream_t s, T *host, const void *dev, size_t bytes) {
    VX_CHECK(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s));
    VX_CHECK(hipStreamSynchronize(s));
    return 0;
}

TVars *padded_copy(const double *M, size_t P, size_t birth) {
    const size_t m = P + 1, b = birth + 1;
    TVars *MM = (TVars *)calloc(b * b, sizeof(TVars));
    if (!MM) return nullptr;
    for (size_t i = 0; i < m; ++i) memcpy(MM + i * b, M + i * m, sizeof(TVars) * m);
    return MM;
}

void save_text(const char *name, const double *M, size_t m) {
    FILE *f = fopen(name, "w");
    if (!f) return;
    fprintf(f, "%zu\n", m);
    for (size_t i = 0; i < m; ++i) {
        for (size_t j = 0; j < m; ++j) fprintf(f, "%lf\t", M[i * m + j]);
        fprintf(f, "\n");
    }
    fclose(f);
}

}  // namespace
}  // namespace vx

using namespace vx;

extern "C" {

int vx_compat_set_precision(int precision) {
    VX_ARG(precision == VX_PREC_F64 || precision == VX_PREC_MIXED, "precision");
    g_compat.precision = precision;
    g_compat.dirty = true;
    return 0;
}

int init_device_conf_values(void) {
    g_compat.dirty = true;
    g_compat.vb = sizeof(Vortex);
    return 0;
}

int vx_compat_init_no_tree(void) {
    g_compat.dirty = true;
    g_compat.vb = 24;
    return 0;
}

TVars *matr_creation(tPanel *panels, size_t s, size_t birth) {
    if (!panels || s < 2 || birth < s) return nullptr;
    vx_params prm = params_from_conf(VX_PREC_F64);
    if (!(prm.dt > 0)) prm.dt = 1;
    if (!(prm.ve_size > 0)) prm.ve_size = 1;
    vx_case *c = nullptr;
    if (vx_case_create(&prm, (const vx_panel *)panels, (int64_t)s, &c)) return nullptr;
    std::vector<double> L((s + 1) * (s + 1)), M((s + 1) * (s + 1));
    TVars *out = nullptr;
    if (vx_case_get_panel_matrix(c, L.data(), (int64_t)s + 1) == 0) {
        save_text("L.txt", L.data(), s + 1);
        if (vx_case_build_matrix(c) == 0 && vx_case_get_matrix(c, M.data(), (int64_t)s + 1) == 0) {
            save_text("M.txt", M.data(), s + 1);
            out = padded_copy(M.data(), s, birth);
        }
    }
    vx_case_destroy(c);
    return out;
}

TVars *load_matrix(size_t *p) {
    if (!p) return nullptr;
    FILE *f = fopen("M.txt", "r");
    if (!f) {
        set_error("unable to open matrix file M.txt");
        return nullptr;
    }
    size_t m = 0;
    if (fscanf(f, "%zu", &m) != 1 || m < 2 || m > 100000) {
        fclose(f);
        return nullptr;
    }
    std::vector<double> M(m * m);
    for (size_t k = 0; k < m * m; ++k)
        if (fscanf(f, "%lf", &M[k]) != 1) {
            fclose(f);
            return nullptr;
        }
    fclose(f);
    *p = m - 1;
    const size_t birth = (size_t)(32 * ceil((double)(m - 1) / 32));
    return padded_copy(M.data(), m - 1, birth);
}

int allocate_arrays(Vortex **p_host, Vortex **p_dev, PVortex **v_host, PVortex **v_dev, TVars **d_dev, size_t size) {
    VX_ARG(p_host && p_dev && v_host && v_dev && d_dev && size, "wrong parameters");
    const size_t vb = (size_t)g_compat.vb;
    *p_host = (Vortex *)calloc(size, vb);
    *v_host = (PVortex *)calloc(size, sizeof(PVortex));
    VX_ARG(*p_host && *v_host, "host allocation");
    VX_CHECK(hipMalloc((void **)p_dev, size * vb));
    VX_CHECK(hipMemset(*p_dev, 0, size * vb));
    VX_CHECK(hipMalloc((void **)d_dev, size * sizeof(TVars)));
    VX_CHECK(hipMemset(*d_dev, 0, size * sizeof(TVars)));
    VX_CHECK(hipMalloc((void **)v_dev, size * sizeof(PVortex)));
    VX_CHECK(hipMemset(*v_dev, 0, size * sizeof(PVortex)));
    return 0;
}

int randomize_tail(Vortex **p_dev, size_t new_size, size_t increased) {
    VX_ARG(p_dev && *p_dev && increased <= new_size, "wrong parameters");
    if (!increased) return 0;
    hipLaunchKernelGGL(k_aos_tail, dim3(blocks((int64_t)increased)), dim3(256), 0, 0, (char *)*p_dev, g_compat.vb,
                       (int64_t)(new_size - increased), (int64_t)new_size);
    VX_CHECK(hipGetLastError());
    VX_CHECK(hipDeviceSynchronize());
    return 0;
}

int incr_vort_quant(Vortex **p_host, Vortex **p_dev, PVortex **v_host, PVortex **v_dev, TVars **d_dev, size_t *size) {
    VX_ARG(p_host && p_dev && v_host && v_dev && d_dev && size, "wrong parameters");
    const size_t inc = conf.inc_step ? conf.inc_step : 8192;
    if (*p_host && *p_dev && *v_host && *v_dev && *d_dev) {
        Vortex *nd = nullptr;
        const size_t ns = *size + inc, vb = (size_t)g_compat.vb;
        VX_CHECK(hipMalloc((void **)&nd, ns * vb));
        VX_CHECK(hipMemcpy(nd, *p_dev, *size * vb, hipMemcpyDeviceToDevice));
        free(*p_host);
        free(*v_host);
        *p_host = (Vortex *)calloc(ns, vb);
        *v_host = (PVortex *)calloc(ns, sizeof(PVortex));
        VX_CHECK(hipFree(*p_dev));
        VX_CHECK(hipFree(*d_dev));
        VX_CHECK(hipFree(*v_dev));
        VX_CHECK(hipMalloc((void **)d_dev, ns * sizeof(TVars)));
        VX_CHECK(hipMalloc((void **)v_dev, ns * sizeof(PVortex)));
        *p_dev = nd;
        *size = ns;
    } else if (!*p_host && !*p_dev && !*v_host && !*v_dev && !*d_dev) {
        *size = inc;
        if (int rc = allocate_arrays(p_host, p_dev, v_host, v_dev, d_dev, *size)) return rc;
    } else {
        VX_ARG(false, "wrong parameters");
    }
    return randomize_tail(p_dev, *size, inc);
}

int vort_creation(Vortex *pos, TVctr *V_infDev, size_t n_of_birth, size_t n_of_birth_BLOCK_S, size_t n, TVars *M_Dev,
                  TVars *d_g, tPanel *panels) {
    VX_ARG(pos && V_infDev && M_Dev && d_g && panels, "null pointer");
    VX_ARG(n_of_birth == conf.birth_quant && n_of_birth_BLOCK_S >= n_of_birth, "birth counts");
    vx_case *c = compat_case(panels);
    if (!c) return 1;
    const int64_t P = (int64_t)n_of_birth, m = P + 1, b = (int64_t)n_of_birth_BLOCK_S + 1;
    if (vx_case_reserve(c, ceil32((int64_t)n + P))) return 1;
    if (g_compat.M_src != M_Dev) {
        if (c->M.ensure((size_t)(m * m), c->stream)) return 1;
        VX_CHECK(hipMemcpy2DAsync(c->M.p, sizeof(double) * m, M_Dev, sizeof(double) * b, sizeof(double) * m, m,
                                  hipMemcpyDeviceToDevice, c->stream));
        c->have_matrix = true;
        c->have_inverse = true;
        g_compat.M_src = M_Dev;
    }
    double vinf[2], dg = 0.0;
    if (read_dev(c->stream, vinf, V_infDev, sizeof(vinf)) || read_dev(c->stream, &dg, d_g, sizeof(double))) return 1;
    const int64_t src = ceil32((int64_t)n);  // the 32-wide tile loop reads to ceil32(n)
    if (stage_in(c, pos, src)) return 1;
    c->n = (int64_t)n;
    if (launch_right_part(c, src, vinf[0], vinf[1]) || launch_birth(c, dg)) return 1;
    if (stage_out(c, pos, (int64_t)n, (int64_t)n + P)) return 1;
    VX_CHECK(hipStreamSynchronize(c->stream));
    c->n = (int64_t)n + P;
    return 0;
}

void start_timer(hipEvent_t *start, hipEvent_t *stop) {
    (void)hipEventCreate(start);
    (void)hipEventCreate(stop);
    (void)hipEventRecord(*start, 0);
    (void)hipEventSynchronize(*start);
}

float stop_timer(hipEvent_t start, hipEvent_t stop) {
    (void)hipEventRecord(stop, 0);
    (void)hipEventSynchronize(stop);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, start, stop);
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    return ms / 1000;
}

int Speed(Vortex *pos, TVctr *v_inf, size_t s, PVortex *v, TVars *d, TVars nu, tPanel *panels) {
    VX_ARG(pos && v_inf && v && d && panels, "null pointer");
    vx_case *c = compat_case(panels);
    if (!c) return 1;
    if (vx_case_reserve(c, (int64_t)s)) return 1;
    double vinf[2];
    if (read_dev(c->stream, vinf, v_inf, sizeof(vinf))) return 1;
    c->prm.nu = nu;
    if (stage_in(c, pos, (int64_t)s)) return 1;
    c->n = (int64_t)s;
    if (vx_speed(c, vinf[0], vinf[1])) return 1;
    hipLaunchKernelGGL(k_pv_out, dim3(blocks((int64_t)s)), dim3(256), 0, c->stream, c->u.p, c->v.p, (int64_t)s, v);
    VX_CHECK(hipMemcpyAsync(d, c->d.p, sizeof(double) * s, hipMemcpyDeviceToDevice, c->stream));
    VX_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int Step(Vortex *pos, PVortex *V, size_t *n, size_t s, TVars *d_g, PVortex *F_p, TVars *M, tPanel *panels) {
    VX_ARG(pos && V && n && d_g && F_p && M && panels, "null pointer");
    (void)s;
    vx_case *c = compat_case(panels);
    if (!c) return 1;
    const int64_t n0 = (int64_t)*n;
    if (vx_case_reserve(c, ceil32(n0))) return 1;
    if (stage_in(c, pos, n0)) return 1;
    if (n0) {
        hipLaunchKernelGGL(k_pv_in, dim3(blocks(n0)), dim3(256), 0, c->stream, V, n0, c->u.p, c->v.p);
        VX_CHECK(hipGetLastError());
    }
    double f0[3];
    if (read_dev(c->stream, &f0[0], F_p, 2 * sizeof(double)) || read_dev(c->stream, &f0[2], M, sizeof(double))) return 1;
    c->n = n0;
    double out[4];
    if (vx_step(c, out)) return 1;
    if (stage_out(c, pos, 0, n0)) return 1;
    // summ_Kernel accumulates onto the caller's F_p and M before scaling
    // (src/kernel.cu:1144-1157)
    const double res[4] = {out[0], f0[0] * (conf.rho / conf.dt) + out[1], f0[1] * (conf.rho / conf.dt) + out[2],
                           f0[2] * (conf.rho / (2 * conf.dt)) + out[3]};
    VX_CHECK(hipMemcpyAsync(d_g, &res[0], sizeof(double), hipMemcpyHostToDevice, c->stream));
    VX_CHECK(hipMemcpyAsync(F_p, &res[1], 2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    VX_CHECK(hipMemcpyAsync(M, &res[3], sizeof(double), hipMemcpyHostToDevice, c->stream));
    VX_CHECK(hipStreamSynchronize(c->stream));
    *n = (size_t)c->n;
    return 0;
}

int velocity_control(Vortex *pos, TVctr *V_inf, size_t n, Vortex *Contr_points, PVortex *V, size_t n_contr) {
    VX_ARG(pos && V_inf && Contr_points && V, "null pointer");
    vx_case *c = g_compat.c;
    VX_ARG(c, "no case yet: call vort_creation, Speed or Step first");
    double vinf[2];
    if (read_dev(c->stream, vinf, V_inf, sizeof(vinf))) return 1;
    HostPts cp;
    if (read_pts(cp, Contr_points, n_contr)) return 1;
    std::vector<double> px(n_contr), py(n_contr), u(n_contr), w(n_contr);
    for (size_t k = 0; k < n_contr; ++k) {
        px[k] = cp.rec(k)[0];
        py[k] = cp.rec(k)[1];
    }
    const int64_t src = ceil32((int64_t)n);
    if (vx_case_reserve(c, src) || stage_in(c, pos, src)) return 1;
    c->n = (int64_t)n;
    if (vx_velocity_probe(c, px.data(), py.data(), (int64_t)n_contr, vinf[0], vinf[1], u.data(), w.data())) return 1;
    std::vector<PVortex> out(n_contr);
    for (size_t k = 0; k < n_contr; ++k) {
        out[k].v[0] = u[k];
        out[k].v[1] = w[k];
    }
    if (n_contr) VX_CHECK(hipMemcpy(V, out.data(), sizeof(PVortex) * n_contr, hipMemcpyHostToDevice));
    return 0;
}

// ---- second component (src/unit_1.cu:1271-1340) ------------------------------
namespace {

// save_vel_to_file (src/unit_1.cu:616-631)
void compat_write_vels(const HostPts &pos, const std::vector<PVortex> &vel, int stage) {
    char fn[256];
    snprintf(fn, sizeof(fn), "output/vels/Vel%06d_%d.txt", current_step, stage);
    FILE *f = fopen(fn, "w");
    if (!f) return;
    fprintf(f, "%zu\n", pos.m);
    for (size_t i = 0; i < pos.m; ++i)
        fprintf(f, "%zu %lf %lf %lf %lf\n", i, pos.rec(i)[0], pos.rec(i)[1], vel[i].v[0], vel[i].v[1]);
    fclose(f);
}

// step_Kernel on the points with null force pointers: positions and flags of
// the caller's device AoS points moved by the device velocities.
int compat_points_move(vx_case *c, Vortex *pts_dev, HostPts &pts, const std::vector<PVortex> &vel) {
    const size_t m = pts.m;
    std::vector<double> x(m), y(m), g(m), u(m), v(m);
    for (size_t i = 0; i < m; ++i) {
        x[i] = pts.rec(i)[0];
        y[i] = pts.rec(i)[1];
        g[i] = pts.rec(i)[2];
        u[i] = vel[i].v[0];
        v[i] = vel[i].v[1];
    }
    if (vx_points_move(c, x.data(), y.data(), g.data(), u.data(), v.data(), (int64_t)m)) return 1;
    for (size_t i = 0; i < m; ++i) {
        pts.rec(i)[0] = x[i];
        pts.rec(i)[1] = y[i];
        pts.rec(i)[2] = g[i];
    }
    if (m) VX_CHECK(hipMemcpy(pts_dev, pts.raw.data(), pts.raw.size(), hipMemcpyHostToDevice));
    return 0;
}

}  // namespace

int first_speed(Vortex *pos, TVctr *V_inf, size_t n, Vortex *second_pos, PVortex *v_second, size_t *n_second,
                tPanel *panels) {
    VX_ARG(pos && V_inf && second_pos && v_second && n_second && panels, "null pointer");
    if (velocity_control(pos, V_inf, n, second_pos, v_second, *n_second)) return 1;
    vx_case *c = compat_case(panels);
    if (!c) return 1;
    const size_t m = *n_second;
    HostPts pts;
    std::vector<PVortex> vel(m);
    if (read_pts(pts, second_pos, m)) return 1;
    if (m) VX_CHECK(hipMemcpy(vel.data(), v_second, sizeof(PVortex) * m, hipMemcpyDeviceToHost));
    compat_write_vels(pts, vel, 5);
    return compat_points_move(c, second_pos, pts, vel);
}

int second_speed(Vortex *pos, TVctr *V_inf, size_t n, Vortex *second_pos, PVortex *v_second, PVortex *v_env,
                 size_t *n_second, tPanel *panels) {
    VX_ARG(pos && V_inf && second_pos && v_second && v_env && n_second && panels, "null pointer");
    VX_ARG(conf.rel_t != 0.0, "rel_t is nessesary if second component set");
    if (velocity_control(pos, V_inf, n, second_pos, v_env, *n_second)) return 1;
    vx_case *c = compat_case(panels);
    if (!c) return 1;
    const size_t m = *n_second;
    if (current_step == first_step)
        VX_CHECK(hipMemcpy(v_second, v_env, sizeof(PVortex) * m, hipMemcpyDeviceToDevice));
    HostPts pts;
    std::vector<PVortex> env(m), vel(m);
    if (read_pts(pts, second_pos, m)) return 1;
    if (m) {
        VX_CHECK(hipMemcpy(env.data(), v_env, sizeof(PVortex) * m, hipMemcpyDeviceToHost));
        VX_CHECK(hipMemcpy(vel.data(), v_second, sizeof(PVortex) * m, hipMemcpyDeviceToHost));
    }
    compat_write_vels(pts, env, 4);
    std::vector<double> g(m), ue(m), ve(m), u(m), v(m);
    for (size_t i = 0; i < m; ++i) {
        g[i] = pts.rec(i)[2];
        ue[i] = env[i].v[0];
        ve[i] = env[i].v[1];
        u[i] = vel[i].v[0];
        v[i] = vel[i].v[1];
    }
    if (vx_points_relax(g.data(), ue.data(), ve.data(), u.data(), v.data(), (int64_t)m, conf.rel_t, conf.dt))
        return 2;
    for (size_t i = 0; i < m; ++i) {
        vel[i].v[0] = u[i];
        vel[i].v[1] = v[i];
    }
    VX_CHECK(hipMemcpy(v_second, vel.data(), sizeof(PVortex) * m, hipMemcpyHostToDevice));
    compat_write_vels(pts, vel, 5);
    return compat_points_move(c, second_pos, pts, vel);
}

}  // extern "C"
