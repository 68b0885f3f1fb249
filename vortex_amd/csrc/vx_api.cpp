// Host orchestration behind the C-ABI of include/vortex.h.
//
// The call order of one time step follows the reference loop
// (src/main.c:652-785): birth (vort_creation), velocities over the padded
// count s = ceil32(n) (Speed), then move/absorb/delete, sums, compaction and
// the collapse passes (Step).  All device work of a case runs on its own HIP
// stream; the host reads back only the four sums and the vortex count, as the
// reference does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "vx_internal.h"

namespace vx {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

static int64_t ceil32(int64_t n) { return (n + 31) / 32 * 32; }

// Grow capacity in inc_step quanta (incr_vort_quant, src/unit_1.cu:511-558):
// positions are kept, velocities need not be, new slots get the far tail.
static int reserve(vx_case *c, int64_t want) {
    if (want <= c->cap) return 0;
    const int64_t inc = c->prm.inc_step > 0 ? c->prm.inc_step : 8192;
    int64_t cap = c->cap;
    while (cap < want) cap += inc;
    const int64_t old = c->cap;
    hipStream_t s = c->stream;
    if (c->x.ensure((size_t)cap, s, true) || c->y.ensure((size_t)cap, s, true) || c->g.ensure((size_t)cap, s, true) ||
        c->u.ensure((size_t)cap, s, true) || c->v.ensure((size_t)cap, s, true) || c->d.ensure((size_t)cap, s, true))
        return 1;
    c->cap = cap;
    if (launch_fill_tail(c, old, cap)) return 1;
    VX_CHECK(hipMemsetAsync(c->u.p + old, 0, sizeof(double) * (cap - old), s));
    VX_CHECK(hipMemsetAsync(c->v.p + old, 0, sizeof(double) * (cap - old), s));
    VX_CHECK(hipMemsetAsync(c->d.p + old, 0, sizeof(double) * (cap - old), s));
    return 0;
}

}  // namespace vx

using namespace vx;

extern "C" {

int vx_abi_version(void) { return VX_ABI_VERSION; }
const char *vx_last_error(void) { return g_err.c_str(); }

int vx_device_count(int *count) {
    VX_ARG(count, "count");
    int n = 0;
    VX_CHECK(hipGetDeviceCount(&n));
    *count = n;
    return 0;
}

static int case_init(vx_case *c, const vx_params *prm, const vx_panel *panels, int64_t P);
static void graph_release(vx_case *c);

int vx_case_create(const vx_params *prm, const vx_panel *panels, int64_t P, vx_case **out) {
    VX_ARG(prm && panels && out, "null pointer");
    VX_ARG(P >= 2, "need at least two panels");
    VX_ARG(prm->dt > 0 && prm->ve_size > 0, "dt and ve_size must be positive");
    VX_ARG(prm->precision == VX_PREC_F64 || prm->precision == VX_PREC_MIXED, "precision");
    for (int64_t i = 0; i < P; ++i)
        VX_ARG(panels[i].n_of_rpanel < (unsigned)P && panels[i].n_of_lpanel < (unsigned)P, "panel neighbour index");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error("no HIP device");
        return 1;
    }
    vx_case *c = new vx_case();
    const int rc = case_init(c, prm, panels, P);
    if (rc) {
        vx_case_destroy(c);  // releases whatever the failed set-up allocated
        return rc;
    }
    *out = c;
    return 0;
}

static int case_init(vx_case *c, const vx_params *prm, const vx_panel *panels, int64_t P) {
    c->prm = *prm;
    if (c->prm.n_col < 0) c->prm.n_col = 0;
    if (c->prm.n_of_points <= 0) c->prm.n_of_points = 20;
    c->P = P;
    VX_CHECK(hipGetDevice(&c->device));
    VX_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    VX_CHECK(hipHostMalloc((void **)&c->h_count, 128, hipHostMallocDefault));
    VX_CHECK(hipHostMalloc((void **)&c->h_dev, 128, hipHostMallocMapped | hipHostMallocCoherent));
    VX_CHECK(hipHostGetDevicePointer((void **)&c->d_dev, c->h_dev, 0));
    VX_CHECK(hipHostMalloc((void **)&c->h_red, 64, hipHostMallocDefault));
    for (auto &e : c->ev) VX_CHECK(hipEventCreate(&e));
    // panels -> SoA, one allocation: 13 double columns + rp
    std::vector<double> cols((size_t)(13 * P));
    std::vector<int32_t> rp((size_t)P);
    double bxmin = 1e300, bxmax = -1e300, bymin = 1e300, bymax = -1e300, lmax = 0;
    for (int64_t i = 0; i < P; ++i) {
        const vx_panel &p = panels[i];
        const double v[13] = {p.left[0], p.left[1], p.right[0], p.right[1], p.contr[0], p.contr[1], p.birth[0],
                              p.birth[1], p.norm[0], p.norm[1], p.tang[0], p.tang[1], p.length};
        for (int k = 0; k < 13; ++k) cols[(size_t)(k * P + i)] = v[k];
        rp[(size_t)i] = (int32_t)p.n_of_rpanel;
        bxmin = std::min(bxmin, p.birth[0]); bxmax = std::max(bxmax, p.birth[0]);
        bymin = std::min(bymin, p.birth[1]); bymax = std::max(bymax, p.birth[1]);
        lmax = std::max(lmax, p.length);
    }
    c->wall_box = make_double4(bxmin - 5 * lmax, bxmax + 5 * lmax, bymin - 5 * lmax, bymax + 5 * lmax);
    const size_t bytes = sizeof(double) * 13 * P + sizeof(int32_t) * P;
    VX_CHECK(hipMalloc(&c->pan_block, bytes));
    VX_CHECK(hipMemcpy(c->pan_block, cols.data(), sizeof(double) * 13 * P, hipMemcpyHostToDevice));
    VX_CHECK(hipMemcpy((char *)c->pan_block + sizeof(double) * 13 * P, rp.data(), sizeof(int32_t) * P,
                       hipMemcpyHostToDevice));
    double *b = (double *)c->pan_block;
    double **f[13] = {&c->pan.lx, &c->pan.ly, &c->pan.rx, &c->pan.ry, &c->pan.cx, &c->pan.cy, &c->pan.bx,
                      &c->pan.by, &c->pan.nx, &c->pan.ny, &c->pan.tx, &c->pan.ty, &c->pan.len};
    for (int k = 0; k < 13; ++k) *f[k] = b + (size_t)k * P;
    c->pan.rp = (int32_t *)((char *)c->pan_block + sizeof(double) * 13 * P);
    c->pan.P = P;
    // birth RHS segments: runs of at most 128 consecutive panels
    {
        const int64_t K = (P + 127) / 128;
        std::vector<vx::RhsSeg> segs((size_t)K);
        std::vector<int32_t> seg_of((size_t)P);
        for (int64_t sg = 0; sg < K; ++sg) {
            const int64_t f = sg * P / K, e = (sg + 1) * P / K;
            double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
            for (int64_t i = f; i < e; ++i) {
                const vx_panel &p = panels[i];
                x0 = std::min({x0, p.left[0], p.right[0]}); x1 = std::max({x1, p.left[0], p.right[0]});
                y0 = std::min({y0, p.left[1], p.right[1]}); y1 = std::max({y1, p.left[1], p.right[1]});
            }
            const double cx = 0.5 * (x0 + x1), cy = 0.5 * (y0 + y1);
            double r = 0.0;
            for (int64_t i = f; i < e; ++i) {
                const vx_panel &p = panels[i];
                r = std::max({r, std::hypot(p.left[0] - cx, p.left[1] - cy), std::hypot(p.right[0] - cx, p.right[1] - cy)});
                seg_of[(size_t)i] = (int32_t)sg;
            }
            segs[(size_t)sg] = vx::RhsSeg{cx, cy, r, (int32_t)f, (int32_t)(e - f)};
        }
        VX_CHECK(hipMalloc((void **)&c->segs, sizeof(vx::RhsSeg) * K + sizeof(int32_t) * P));
        c->seg_of = (int32_t *)(c->segs + K);
        VX_CHECK(hipMemcpy(c->segs, segs.data(), sizeof(vx::RhsSeg) * K, hipMemcpyHostToDevice));
        VX_CHECK(hipMemcpy(c->seg_of, seg_of.data(), sizeof(int32_t) * P, hipMemcpyHostToDevice));
        c->nseg = K;
    }
    // newborn positions, and whether all of them lie well inside the tree's
    // inclusion margin (400 around the deletion box), which lets the tree
    // include them before their circulations are known
    {
        VX_CHECK(hipMalloc((void **)&c->bpos, sizeof(double) * 2 * P));
        if (launch_birth_positions(c)) return 1;
        std::vector<double> bp((size_t)(2 * P));
        VX_CHECK(hipMemcpyAsync(bp.data(), c->bpos, sizeof(double) * 2 * P, hipMemcpyDeviceToHost, c->stream));
        VX_CHECK(hipStreamSynchronize(c->stream));
        bool near = true;
        for (int64_t i = 0; i < P; ++i) {
            const double bx = bp[(size_t)i], by = bp[(size_t)(P + i)];
            const double ex = std::max(0.0, std::max(c->prm.x_min - bx, bx - c->prm.x_max));
            const double ey = std::max(0.0, std::max(c->prm.y_min - by, by - c->prm.y_max));
            near = near && std::isfinite(bx) && std::isfinite(by) && ex * ex + ey * ey <= 399.0 * 399.0;
        }
        c->births_near = near;
    }
    if (reserve(c, c->prm.inc_step > 0 ? c->prm.inc_step : 8192)) return 1;
    VX_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int vx_case_destroy(vx_case *c) {
    if (!c) return 0;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    vx::DBuf<double> *db[] = {&c->L, &c->M, &c->x, &c->y, &c->g, &c->u, &c->v, &c->d, &c->rp, &c->gsheet,
                              &c->partial, &c->red, &c->tx, &c->ty, &c->tg, &c->px, &c->py, &c->pu, &c->pv, &c->pg,
                              &c->rhs_s, &c->prb_x, &c->prb_y, &c->prb_u, &c->prb_v, &c->prb_part};
    for (auto *p : db) p->release();
    c->xf.release(); c->yf.release(); c->gf.release();
    c->flags.release(); c->scan.release(); c->lists.release(); c->col.release(); c->misc.release();
    c->keys.release(); c->keys_alt.release(); c->vals.release(); c->vals_alt.release(); c->cnode.release();
    c->ptmp.release(); c->gridp.release(); c->dcount.release(); c->colseg.release();
    if (c->side) (void)hipStreamSynchronize(c->side);
    c->wlist.release(); c->wtmp.release(); c->wstage.release();
    tree_release(c);
    lu_release(c);
    graph_release(c);
    c->gr_ring.release();
    if (c->pan_block) (void)hipFree(c->pan_block);
    if (c->segs) (void)hipFree(c->segs);
    if (c->bpos) (void)hipFree(c->bpos);
    if (c->h_count) (void)hipHostFree(c->h_count);
    if (c->h_dev) (void)hipHostFree(c->h_dev);
    if (c->h_red) (void)hipHostFree(c->h_red);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {c->fork_ev, c->join_ev, c->pre_ev, c->born_ev, c->aux_fork, c->aux_done[0], c->aux_done[1]})
        if (e) (void)hipEventDestroy(e);
    for (hipStream_t a : c->aux)
        if (a) (void)hipStreamDestroy(a);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

static bool lu_mode(const vx_case *c) { return (c->prm.flags & VX_FLAG_LU_REFINE) != 0; }

int vx_case_build_matrix(vx_case *c) {
    VX_ARG(c, "case");
    if (launch_panel_matrix(c)) return 1;
    if (lu_mode(c) ? launch_lu_setup(c) : launch_invert(c)) return 1;
    c->have_inverse = !lu_mode(c);
    c->have_matrix = true;
    return 0;
}

int vx_case_set_matrix(vx_case *c, const double *M, int64_t m) {
    VX_ARG(c && M, "null pointer");
    VX_ARG(m == c->P + 1, "matrix size must be P+1");
    if (c->M.ensure((size_t)(m * m), c->stream)) return 1;
    VX_CHECK(hipMemcpyAsync(c->M.p, M, sizeof(double) * m * m, hipMemcpyHostToDevice, c->stream));
    VX_CHECK(hipStreamSynchronize(c->stream));
    c->have_inverse = true;
    if (lu_mode(c) && !c->lu && (launch_panel_matrix(c) || launch_lu_setup(c))) return 1;
    c->have_matrix = true;
    return 0;
}

int vx_case_solve_info(vx_case *c, int32_t *iters, int32_t *converged, double *resid_rel) {
    VX_ARG(c, "case");
    return lu_info(c, iters, converged, resid_rel);
}

int vx_case_get_matrix(vx_case *c, double *M, int64_t m) {
    VX_ARG(c && M, "null pointer");
    VX_ARG(m == c->P + 1 && c->have_matrix && c->have_inverse, "no inverse matrix of that size");
    VX_CHECK(hipMemcpyAsync(M, c->M.p, sizeof(double) * m * m, hipMemcpyDeviceToHost, c->stream));
    VX_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int vx_case_get_panel_matrix(vx_case *c, double *L, int64_t m) {
    VX_ARG(c && L, "null pointer");
    VX_ARG(m == c->P + 1, "matrix size must be P+1");
    if (!c->L.p && launch_panel_matrix(c)) return 1;
    VX_CHECK(hipMemcpyAsync(L, c->L.p, sizeof(double) * m * m, hipMemcpyDeviceToHost, c->stream));
    VX_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int vx_case_reserve(vx_case *c, int64_t capacity) {
    VX_ARG(c && capacity >= 0, "case/capacity");
    if (reserve(c, capacity)) return 1;
    VX_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int vx_case_set_state(vx_case *c, const double *x, const double *y, const double *g, int64_t n) {
    VX_ARG(c && n >= 0 && (n == 0 || (x && y && g)), "state arrays");
    if (reserve(c, ceil32(n + c->P))) return 1;
    if (n) {
        VX_CHECK(hipMemcpyAsync(c->x.p, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
        VX_CHECK(hipMemcpyAsync(c->y.p, y, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
        VX_CHECK(hipMemcpyAsync(c->g.p, g, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    }
    // the padding that the stages read beyond n ([n, ceil32(n + P))) holds
    // the far tail (g = 0), whatever an earlier, larger state left there
    if (launch_fill_tail(c, n, std::min(c->cap, ceil32(n + c->P)))) return 1;
    VX_CHECK(hipStreamSynchronize(c->stream));
    c->n = n;
    return 0;
}

int vx_case_get_state(vx_case *c, double *x, double *y, double *g, int64_t count) {
    VX_ARG(c && count >= 0 && count <= c->cap, "count beyond capacity");
    if (x) VX_CHECK(hipMemcpyAsync(x, c->x.p, sizeof(double) * count, hipMemcpyDeviceToHost, c->stream));
    if (y) VX_CHECK(hipMemcpyAsync(y, c->y.p, sizeof(double) * count, hipMemcpyDeviceToHost, c->stream));
    if (g) VX_CHECK(hipMemcpyAsync(g, c->g.p, sizeof(double) * count, hipMemcpyDeviceToHost, c->stream));
    VX_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int vx_case_get_velocity(vx_case *c, double *u, double *v, double *d, int64_t count) {
    VX_ARG(c && count >= 0 && count <= c->cap, "count beyond capacity");
    if (u) VX_CHECK(hipMemcpyAsync(u, c->u.p, sizeof(double) * count, hipMemcpyDeviceToHost, c->stream));
    if (v) VX_CHECK(hipMemcpyAsync(v, c->v.p, sizeof(double) * count, hipMemcpyDeviceToHost, c->stream));
    if (d) VX_CHECK(hipMemcpyAsync(d, c->d.p, sizeof(double) * count, hipMemcpyDeviceToHost, c->stream));
    VX_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int vx_case_set_velocity(vx_case *c, const double *u, const double *v, const double *d, int64_t count) {
    VX_ARG(c && count >= 0 && count <= c->cap, "count beyond capacity");
    if (u) VX_CHECK(hipMemcpyAsync(c->u.p, u, sizeof(double) * count, hipMemcpyHostToDevice, c->stream));
    if (v) VX_CHECK(hipMemcpyAsync(c->v.p, v, sizeof(double) * count, hipMemcpyHostToDevice, c->stream));
    if (d) {
        VX_CHECK(hipMemcpyAsync(c->d.p, d, sizeof(double) * count, hipMemcpyHostToDevice, c->stream));
        tree_d_changed(c);
    }
    VX_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

// ---- exact checkpoint ----------------------------------------------------------
// File: header (magic, sizes, d_g, the tree's far-target hint, the case
// parameters), then x, y, g over all `cap` slots (the tail beyond n takes
// part in the nearest-3 search), then the caller's blob.  Native byte order.
namespace {
struct CkptHeader {
    char magic[8];
    int64_t P, n, cap, far_hint, user_bytes;
    double d_g;
    vx_params prm;
};
constexpr char kCkptMagic[8] = {'V', 'X', 'C', 'K', 'P', 'T', '0', '1'};

bool same_bits(double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; }
}  // namespace

int vx_case_save(vx_case *c, const char *path, const void *user, int64_t user_bytes) {
    VX_ARG(c && path && user_bytes >= 0 && (user_bytes == 0 || user), "checkpoint arguments");
    CkptHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kCkptMagic, sizeof(h.magic));
    h.P = c->P;
    h.n = c->n;
    h.cap = c->cap;
    h.far_hint = tree_far_hint(c);
    h.user_bytes = user_bytes;
    h.d_g = c->d_g;
    h.prm = c->prm;
    std::vector<double> buf((size_t)(3 * c->cap));
    hipStream_t s = c->stream;
    VX_CHECK(hipMemcpyAsync(buf.data(), c->x.p, sizeof(double) * c->cap, hipMemcpyDeviceToHost, s));
    VX_CHECK(hipMemcpyAsync(buf.data() + c->cap, c->y.p, sizeof(double) * c->cap, hipMemcpyDeviceToHost, s));
    VX_CHECK(hipMemcpyAsync(buf.data() + 2 * c->cap, c->g.p, sizeof(double) * c->cap, hipMemcpyDeviceToHost, s));
    VX_CHECK(hipStreamSynchronize(s));
    FILE *f = std::fopen(path, "wb");
    if (!f) {
        set_error(std::string("checkpoint: cannot write ") + path);
        return 1;
    }
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 && std::fwrite(buf.data(), sizeof(double), buf.size(), f) ==
                                                              buf.size();
    if (ok && user_bytes) ok = std::fwrite(user, 1, (size_t)user_bytes, f) == (size_t)user_bytes;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        set_error(std::string("checkpoint: write failed: ") + path);
        return 1;
    }
    return 0;
}

int vx_case_load(vx_case *c, const char *path, void *user, int64_t user_cap, int64_t *user_bytes) {
    VX_ARG(c && path && user_cap >= 0 && (user_cap == 0 || user), "checkpoint arguments");
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        set_error(std::string("checkpoint: cannot read ") + path);
        return 1;
    }
    CkptHeader h;
    std::vector<double> buf;
    std::string bad;
    if (std::fread(&h, sizeof(h), 1, f) != 1 || std::memcmp(h.magic, kCkptMagic, sizeof(h.magic)) != 0)
        bad = "not a checkpoint of this library";
    else if (h.P != c->P || h.prm.precision != c->prm.precision || !same_bits(h.prm.dt, c->prm.dt) ||
             !same_bits(h.prm.ve_size, c->prm.ve_size) || !same_bits(h.prm.nu, c->prm.nu))
        bad = "checkpoint of another case (panels, precision, dt, ve_size or nu differ)";
    else if (h.n < 0 || h.cap < h.n || h.user_bytes < 0)
        bad = "corrupt checkpoint header";
    else if (h.user_bytes > user_cap)
        bad = "checkpoint: the caller's buffer is smaller than its blob (" + std::to_string(h.user_bytes) + " B)";
    if (bad.empty()) {
        buf.resize((size_t)(3 * h.cap));
        if (std::fread(buf.data(), sizeof(double), buf.size(), f) != buf.size() ||
            (h.user_bytes && std::fread(user, 1, (size_t)h.user_bytes, f) != (size_t)h.user_bytes))
            bad = "truncated checkpoint";
    }
    std::fclose(f);
    if (!bad.empty()) {
        set_error(bad + ": " + path);
        return 2;
    }
    if (reserve(c, h.cap)) return 1;
    hipStream_t s = c->stream;
    if (h.cap) {
        VX_CHECK(hipMemcpyAsync(c->x.p, buf.data(), sizeof(double) * h.cap, hipMemcpyHostToDevice, s));
        VX_CHECK(hipMemcpyAsync(c->y.p, buf.data() + h.cap, sizeof(double) * h.cap, hipMemcpyHostToDevice, s));
        VX_CHECK(hipMemcpyAsync(c->g.p, buf.data() + 2 * h.cap, sizeof(double) * h.cap, hipMemcpyHostToDevice, s));
    }
    // slots the saved case had not grown to yet: the far tail they get on growth
    if (c->cap > h.cap && launch_fill_tail(c, h.cap, c->cap)) return 1;
    VX_CHECK(hipStreamSynchronize(s));
    c->n = h.n;
    c->d_g = h.d_g;
    if ((h.far_hint > 0 || c->tree) && tree_set_far_hint(c, h.far_hint)) return 1;
    if (user_bytes) *user_bytes = h.user_bytes;
    return 0;
}

int64_t vx_case_n(const vx_case *c) { return c ? c->n : -1; }
int64_t vx_case_capacity(const vx_case *c) { return c ? c->cap : -1; }
double vx_case_d_g(const vx_case *c) { return c ? c->d_g : 0.0; }

int vx_case_set_n(vx_case *c, int64_t n) {
    VX_ARG(c && n >= 0 && n <= c->cap, "n beyond capacity");
    c->n = n;
    return 0;
}

int vx_case_set_d_g(vx_case *c, double d_g) {
    VX_ARG(c, "case");
    c->d_g = d_g;
    return 0;
}

int vx_case_sync(vx_case *c) {
    VX_ARG(c, "case");
    VX_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int vx_case_enable_timing(vx_case *c, int on) {
    VX_ARG(c, "case");
    c->timing = on != 0;
    return 0;
}

int vx_case_timings(vx_case *c, vx_timings *t) {
    VX_ARG(c && t, "null pointer");
    *t = c->tm;
    return 0;
}

// ---- stages ----------------------------------------------------------------
int vx_right_part(vx_case *c, int64_t n_src, double vinf_x, double vinf_y, double *R_p) {
    VX_ARG(c && n_src >= 0 && n_src <= c->cap, "n_src beyond capacity");
    if (launch_right_part(c, n_src, vinf_x, vinf_y)) return 1;
    if (R_p) VX_CHECK(hipMemcpyAsync(R_p, c->rp.p, sizeof(double) * c->P, hipMemcpyDeviceToHost, c->stream));
    VX_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int vx_birth(vx_case *c, const double *R_p, double d_g, double *gsheet) {
    VX_ARG(c && c->have_matrix, "no panel matrix");
    if (reserve(c, ceil32(c->n + c->P))) return 1;
    if (R_p) {
        if (c->rp.ensure((size_t)c->P, c->stream)) return 1;
        VX_CHECK(hipMemcpyAsync(c->rp.p, R_p, sizeof(double) * c->P, hipMemcpyHostToDevice, c->stream));
    }
    if (launch_birth(c, d_g)) return 1;
    if (gsheet)
        VX_CHECK(hipMemcpyAsync(gsheet, birth_sheet(c), sizeof(double) * c->P, hipMemcpyDeviceToHost, c->stream));
    VX_CHECK(hipStreamSynchronize(c->stream));
    c->n += c->P;
    return 0;
}

int vx_biot_savart(vx_case *c, int64_t n_tgt, int64_t n_src, double vinf_x, double vinf_y) {
    VX_ARG(c && n_tgt >= 0 && n_tgt <= n_src && n_src <= c->cap, "need n_tgt <= n_src <= capacity");
    if (launch_biot_savart(c, n_tgt, n_src, vinf_x, vinf_y)) return 1;
    VX_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int vx_tree_biot_savart(vx_case *c, int64_t s, double vinf_x, double vinf_y) {
    VX_ARG(c && s >= 0 && s <= c->cap, "count beyond capacity");
    if (launch_tree_biot_savart(c, s, vinf_x, vinf_y)) return 1;
    VX_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int vx_tree_diffusion(vx_case *c) {
    VX_ARG(c, "case");
    if (launch_tree_diffusion(c)) return 1;
    VX_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int vx_tree_stats(vx_case *c, vx_tree_info *info) {
    VX_ARG(c && info, "null pointer");
    return tree_stats(c, info);
}

int vx_tree_census(vx_case *c, int on) {
    VX_ARG(c, "null pointer");
    c->census = on != 0;
    return 0;
}

int vx_tree_order(vx_case *c, uint32_t *keys, int32_t *perm, int64_t count) {
    VX_ARG(c && count >= 0, "case/count");
    return tree_copy_order(c, keys, perm, count);
}

int vx_diffusion(vx_case *c, int64_t n_tgt, int64_t n_src) {
    VX_ARG(c && n_tgt >= 0 && n_tgt <= c->cap && n_src <= c->cap, "counts beyond capacity");
    if (launch_diffusion(c, n_tgt, n_src)) return 1;
    VX_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int vx_diffusion_wall(vx_case *c, int64_t n_tgt) {
    VX_ARG(c && n_tgt >= 0 && n_tgt <= c->cap, "count beyond capacity");
    if (launch_diffusion_wall(c, n_tgt)) return 1;
    VX_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

// Read back the step sums (c->red[0..3]), the device count (slot dn_slot)
// and the collapse error flag; one synchronisation.
static int step_readback(vx_case *c, double out4[4], int *err) {
    if (out4) VX_CHECK(hipMemcpyAsync(c->h_red, c->red.p, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    VX_CHECK(hipMemcpyAsync(c->h_count, c->dcount.p, 9 * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    VX_CHECK(hipStreamSynchronize(c->stream));
    if (out4)
        for (int q = 0; q < 4; ++q) out4[q] = c->h_red[q];
    c->n = c->h_count[c->dn_slot];
    if (err) *err = (int32_t)c->h_count[3];
    return 0;
}

int vx_step_move(vx_case *c, double out4[4]) {
    VX_ARG(c && out4, "null pointer");
    if (launch_set_dn(c, c->n) || launch_step_move(c)) return 1;
    return step_readback(c, out4, nullptr);
}

int vx_compact(vx_case *c) {
    VX_ARG(c, "case");
    if (launch_set_dn(c, c->n) || launch_compact(c, c->n)) return 1;
    return step_readback(c, nullptr, nullptr);
}

int vx_setka(vx_case *c, int kind, int32_t *COL) {
    VX_ARG(c && (kind == 1 || kind == 2), "kind must be 1 or 2");
    if (launch_set_dn(c, c->n) || launch_live_box(c) || launch_setka(c, kind, c->n, false)) return 1;
    if (COL && c->n) VX_CHECK(hipMemcpyAsync(COL, c->col.p, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
    VX_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int vx_collapse(vx_case *c, int kind, const int32_t *COL) {
    VX_ARG(c && (kind == 1 || kind == 2), "kind must be 1 or 2");
    if (COL) {
        for (int64_t i = 0; i < c->n; ++i)
            VX_ARG(COL[i] == -1 || (COL[i] > i && COL[i] < c->n), "COL entries must be -1 or in (i, n)");
        if (c->col.ensure((size_t)(c->n > 0 ? c->n : 1), c->stream)) return 1;
        if (c->n) VX_CHECK(hipMemcpyAsync(c->col.p, COL, sizeof(int32_t) * c->n, hipMemcpyHostToDevice, c->stream));
    }
    if (launch_set_dn(c, c->n) || launch_collapse(c, kind, c->n, false)) return 1;
    int err = 0;
    if (step_readback(c, nullptr, &err)) return 1;
    if (err) {
        set_error("collapse resolution did not converge");
        return 1;
    }
    return 0;
}

// vort_creation (src/unit_1.cu:560-595): right-hand side over the 32-padded
// source count, then the newborns at [n, n+P).
int vx_vort_creation(vx_case *c, double vinf_x, double vinf_y) {
    VX_ARG(c && c->have_matrix, "no panel matrix");
    Timer t(c, &c->tm.creation_ms, 0);
    if (reserve(c, ceil32(c->n + c->P))) return 1;
    // the speed stage's tree build may start before the newborns are written
    // (vx_time_step only, without stage timing)
    const int64_t s_next = ceil32(c->n + c->P);
    const bool ov = c->overlap_births && !c->timing && c->births_near && c->prm.tree_min_n > 0 &&
                    s_next >= c->prm.tree_min_n;
    c->born_at = -1;
    if (ov) {
        if (side_stream(c)) return 1;
        VX_CHECK(hipEventRecord(c->pre_ev, c->stream));
    }
    if (launch_right_part(c, ceil32(c->n), vinf_x, vinf_y) || launch_birth(c, c->d_g)) return 1;
    if (ov) {
        VX_CHECK(hipEventRecord(c->born_ev, c->stream));
        // queued after the birth's kernels (the host issues them first, so
        // the main stream starts at once) but ordered only after pre_ev
        if (tree_build_early(c, s_next, c->n)) return 1;
    }
    c->n += c->P;
    return 0;
}

// Speed (src/unit_1.cu:963-1075): over s = ceil32(n) targets and sources.
int vx_speed(vx_case *c, double vinf_x, double vinf_y) {
    VX_ARG(c, "case");
    const int64_t s = ceil32(c->n);
    if (reserve(c, s)) return 1;
    Timer t(c, &c->tm.speed_ms, 1);
    const bool tree = c->prm.tree_min_n > 0 && s >= c->prm.tree_min_n;
    {
        Timer tb(c, &c->tm.bs_ms, 2);
        if (tree ? launch_tree_biot_savart(c, s, vinf_x, vinf_y) : launch_biot_savart(c, s, s, vinf_x, vinf_y))
            return 1;
        c->tm.bs_launches++;
    }
    // The wall term, the tree's domain-scale diffusion targets and the step's
    // probes over the tree run on the side stream, overlapping the groups'
    // diffusion on the main stream (the side stream's work, slowed down
    // beside it, still ends first); the wall's terms are staged and added
    // after the join.
    int rc = 0;
    {
        Timer td(c, &c->tm.diff_ms, 3);
        if ((tree && tree_diffusion_pre(c)) || side_fork(c)) return 1;
        // the long main-stream kernel is issued first; the side stream's
        // work is ordered after the fork point only
        rc = tree ? tree_diffusion_main(c) : launch_diffusion(c, s, s);
        if (c->timing) (void)hipEventRecord(c->ev[14], c->side);
        if (!rc) rc = launch_wall_side(c, s, c->side);
        if (c->timing) (void)hipEventRecord(c->ev[15], c->side);
        if (!rc && tree) rc = tree_diffusion_far(c, c->side);
        if (!rc && tree && c->prb_in_speed)
            rc = launch_tree_probes(c, c->prb_x.p, c->prb_y.p, c->prb_m, vinf_x, vinf_y, c->prb_u.p, c->prb_v.p,
                                    c->side);
        const int rj = side_join(c);
        if (!rc) rc = rj ? rj : launch_wall_add(c);
    }
    if (!rc && c->timing) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev[14], c->ev[15]) == hipSuccess) c->tm.wall_ms += ms;
    }
    return rc;
}

// Step (src/unit_1.cu:1077-1218).  The count stays on the device until the
// end; every collapse pass is queued, and one whose predecessor of the same
// kind merged nothing searches no grid (launch_setka), which equals the
// reference's early exit when n stops changing.
int vx_step(vx_case *c, double out4[4]) {
    VX_ARG(c && out4, "null pointer");
    Timer t(c, &c->tm.step_ms, 4);
    const int64_t n_ub = c->n;
    if (launch_set_dn(c, n_ub) || launch_step_move(c) || launch_compact(c, n_ub) || launch_step_collapse(c, n_ub))
        return 1;
    int err = 0;
    if (step_readback(c, out4, &err)) return 1;
    if (err) {
        set_error(err & 2 ? "collapse grid node pool full" : "collapse resolution did not converge");
        return 1;
    }
    c->d_g = out4[0];
    return 0;
}

int vx_time_step(vx_case *c, double vinf_x, double vinf_y, double forces4[4]) {
    VX_ARG(c && forces4, "null pointer");
    c->overlap_births = true;
    // velocity_control after the birth (src/main.c:744): direct sums on the
    // main stream beside the tree build, or, when the step's velocities come
    // from the tree, over that tree on the second stream beside the diffusion
    // (the state is the same until Step)
    int rc = vx_vort_creation(c, vinf_x, vinf_y);
    if (!rc) {
        c->prb_in_speed = c->prb_m > 0 && c->prm.tree_min_n > 0 && ceil32(c->n) >= c->prm.tree_min_n;
        rc = (!c->prb_in_speed && launch_step_probes(c, vinf_x, vinf_y)) || vx_speed(c, vinf_x, vinf_y);
        c->prb_in_speed = false;
    }
    c->overlap_births = false;
    if (c->born_at >= 0) {  // an early tree part left behind by a failed stage
        c->born_at = -1;
        (void)side_join(c);
    }
    if (rc || vx_step(c, forces4)) return 1;
    return 0;
}

int vx_run_steps(vx_case *c, int64_t steps, double vinf_x, double vinf_y, double *f4) {
    VX_ARG(c && steps >= 0, "steps");
    double tmp[4];
    for (int64_t k = 0; k < steps; ++k)
        if (vx_time_step(c, vinf_x, vinf_y, f4 ? f4 + 4 * k : tmp)) return 1;
    return 0;
}

// ---- launch-shape tuning ------------------------------------------------------
// Mean milliseconds of `what` over reps runs after one untimed run, by events
// on the case stream (work on the side stream is joined into it).
}  // extern "C"

template <class F>
static int time_ms(vx_case *c, int reps, F what, double *ms) {
    if (what()) return 1;
    hipEvent_t a = nullptr, b = nullptr;
    VX_CHECK(hipEventCreate(&a));
    VX_CHECK(hipEventCreate(&b));
    int rc = hipEventRecord(a, c->stream) != hipSuccess;
    for (int r = 0; r < reps && !rc; ++r) rc = what();
    if (!rc) rc = hipEventRecord(b, c->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess;
    float t = 0.f;
    if (!rc) rc = hipEventElapsedTime(&t, a, b) != hipSuccess;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) {
        if (vx_last_error()[0] == 0) set_error("autotune: timing failed");
        return 1;
    }
    *ms = (double)t / reps;
    return 0;
}

extern "C" {

int vx_case_autotune(vx_case *c, double vinf_x, double vinf_y, vx_tuning *out) {
    VX_ARG(c, "case");
    VX_ARG(!c->timing && !c->devcount, "autotune runs with stage timing off");
    const int64_t s = ceil32(c->n);
    VX_ARG(s > 0, "autotune needs a state");
    if (reserve(c, s)) return 1;
    const bool tree = c->prm.tree_min_n > 0 && s >= c->prm.tree_min_n;
    vx_tuning best = c->tune;
    best.trials = 0;
    constexpr int kReps = 3;
    if (tree) {
        // the whole velocity stage for each bucket depth
        double tdef = 0.0, tbest = 1e300;
        for (int off = 0; off <= 3; ++off) {
            c->tune.leaf_off = off;
            double t = 0.0;
            if (time_ms(c, kReps, [&] { return vx_speed(c, vinf_x, vinf_y); }, &t)) return 1;
            ++best.trials;
            if (off == 1) tdef = t;
            if (t < tbest) {
                tbest = t;
                best.leaf_off = off;
            }
        }
        best.ms_default = tdef;
        best.ms_best = tbest;
    } else {
        // each direct kernel for each number of targets per thread
        const int tpts[4] = {1, 2, 4, 8};
        const int def = direct_tpt_default(c->prm.precision == VX_PREC_MIXED, s);
        double bdef = 0.0, bbest = 1e300, ddef = 0.0, dbest = 1e300;
        for (int q = 0; q < 4; ++q) {
            c->tune.tpt_bs = tpts[q];
            double t = 0.0;
            if (time_ms(c, kReps, [&] { return launch_biot_savart(c, s, s, vinf_x, vinf_y); }, &t)) return 1;
            ++best.trials;
            if (tpts[q] == def) bdef = t;
            if (t < bbest) {
                bbest = t;
                best.tpt_bs = tpts[q];
            }
        }
        c->tune.tpt_bs = best.tpt_bs;
        for (int q = 0; q < 4; ++q) {
            c->tune.tpt_diff = tpts[q];
            double t = 0.0;
            if (time_ms(c, kReps, [&] { return launch_diffusion(c, s, s); }, &t)) return 1;
            ++best.trials;
            if (tpts[q] == def) ddef = t;
            if (t < dbest) {
                dbest = t;
                best.tpt_diff = tpts[q];
            }
        }
        best.ms_default = bdef + ddef;
        best.ms_best = bbest + dbest;
    }
    c->tune = best;
    if (out) *out = best;
    return 0;
}

int vx_case_set_tuning(vx_case *c, const vx_tuning *t) {
    VX_ARG(c && t, "null pointer");
    auto ok_tpt = [](int v) { return v == 0 || v == 1 || v == 2 || v == 4 || v == 8; };
    VX_ARG(ok_tpt(t->tpt_bs) && ok_tpt(t->tpt_diff), "targets per thread must be 0, 1, 2, 4 or 8");
    VX_ARG(t->leaf_off >= -1 && t->leaf_off <= 3, "leaf_off must be -1..3");
    VX_ARG(t->col_share_min >= 0, "col_share_min must be >= 0");
    c->tune.tpt_bs = t->tpt_bs;
    c->tune.tpt_diff = t->tpt_diff;
    c->tune.leaf_off = t->leaf_off;
    c->tune.col_share_min = t->col_share_min;
    return 0;
}

int vx_case_get_tuning(vx_case *c, vx_tuning *t) {
    VX_ARG(c && t, "null pointer");
    *t = c->tune;
    return 0;
}

// One time step on device counts (c->devcount): the loop's order of
// vx_time_step on the direct path, all on the main stream, no host wait,
// ending with the step's record in gr_ring.  Launch sizes depend only on
// c->n_ub and the case, so the same launches repeat every step.
static int device_step(vx_case *c, double vinf_x, double vinf_y) {
    const int64_t n_ub = c->n_ub;
    if (launch_right_part(c, cnt_s(c), vinf_x, vinf_y) || launch_birth(c, 0.0) ||
        launch_step_probes(c, vinf_x, vinf_y))
        return 1;
    const Cnt s = cnt_s(c);
    if (launch_biot_savart(c, s, s, vinf_x, vinf_y) || launch_diffusion(c, s, s) || launch_diffusion_wall(c, s))
        return 1;
    if (launch_set_dn(c, n_ub) || launch_step_move(c) || launch_compact(c, n_ub) ||
        launch_step_collapse(c, n_ub))
        return 1;
    return launch_step_record(c);
}

static void graph_release(vx_case *c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    c->graph_cap = -1;
}

__global__ void k_graph_begin(int64_t *__restrict__ dcount, int64_t n) {
    dcount[1] = n;
    dcount[3] = 0;  // collapse error flag, sticky over the steps
    dcount[9] = 0;  // ring slot
}

// Prepares c for `steps` device-count steps: capacity for every step, the
// record ring, the count, error flag and ring slot on the device, and the
// absorbed circulation in red[0].  Returns with c->devcount set.
static int devcount_begin(vx_case *c, int64_t steps, bool *ring_moved) {
    VX_ARG(c->have_matrix, "no panel matrix");
    VX_ARG(!c->timing, "stage timing synchronises the host; switch it off for graph steps");
    const int64_t n_max = c->n + steps * c->P;
    VX_ARG(c->prm.tree_min_n <= 0 || ceil32(n_max) < c->prm.tree_min_n,
           "graph steps run the direct path: the count must stay below tree_min_n");
    if (reserve(c, ceil32(n_max + c->P) + 32)) return 1;
    hipStream_t s = c->stream;
    if (c->red.ensure(8, s) || c->dcount.ensure(16, s)) return 1;
    *ring_moved = false;
    if (steps > c->gr_cap) {
        if (c->gr_ring.ensure((size_t)(5 * steps), s)) return 1;
        c->gr_cap = (int64_t)(c->gr_ring.n / 5);
        *ring_moved = true;  // the ring's address is part of captured launches
    }
    c->n_ub = c->cap;
    VX_CHECK(hipMemcpyAsync(c->red.p, &c->d_g, sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_graph_begin, dim3(1), dim3(1), 0, s, c->dcount.p, c->n);
    VX_CHECK(hipGetLastError());
    c->devcount = true;
    c->dn_slot = 1;
    return 0;
}

// Reads the record of `steps` device-count steps: forces and counts per step;
// c->n and c->d_g follow the last one.
static int devcount_end(vx_case *c, int64_t steps, double *f4, int64_t *counts) {
    hipStream_t s = c->stream;
    std::vector<double> ring((size_t)(5 * c->gr_cap));
    int64_t flags[16];
    VX_CHECK(hipMemcpyAsync(ring.data(), c->gr_ring.p, sizeof(double) * 5 * c->gr_cap, hipMemcpyDeviceToHost, s));
    VX_CHECK(hipMemcpyAsync(flags, c->dcount.p, sizeof(flags), hipMemcpyDeviceToHost, s));
    VX_CHECK(hipStreamSynchronize(s));
    if (flags[9] != steps) {
        set_error("graph steps: " + std::to_string(flags[9]) + " of " + std::to_string(steps) + " recorded");
        return 1;
    }
    if ((int32_t)flags[3]) {
        set_error((int32_t)flags[3] & 2 ? "collapse grid node pool full" : "collapse resolution did not converge");
        return 1;
    }
    const int64_t *nr = reinterpret_cast<const int64_t *>(ring.data() + 4 * c->gr_cap);
    for (int64_t k = 0; k < steps; ++k) {
        if (f4)
            for (int q = 0; q < 4; ++q) f4[4 * k + q] = ring[(size_t)(4 * k + q)];
        if (counts) counts[k] = nr[k];
    }
    c->n = nr[steps - 1];
    c->d_g = ring[(size_t)(4 * (steps - 1))];
    return 0;
}

static int instantiate(hipGraph_t g, hipGraphExec_t *exec) {
    const hipError_t ei = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        set_error(std::string("graph instantiate: ") + hipGetErrorString(ei));
        *exec = nullptr;
        return 1;
    }
    return 0;
}

// steps time steps of the direct path as replays of one captured graph: the
// first step runs on device counts uncaptured (it sizes every buffer for the
// capacity bound), the second is captured, and the graph is replayed for the
// rest.  The host waits once, at the end.
int vx_run_steps_graph(vx_case *c, int64_t steps, double vinf_x, double vinf_y, double *f4, int64_t *counts) {
    VX_ARG(c && steps >= 0, "steps");
    if (steps == 0) return 0;
    bool moved = false;
    if (devcount_begin(c, steps, &moved)) {
        c->devcount = false;
        return 1;
    }
    if (moved || c->graph_cap != c->cap || c->graph_vinf[0] != vinf_x || c->graph_vinf[1] != vinf_y)
        graph_release(c);
    hipStream_t s = c->stream;
    int rc = device_step(c, vinf_x, vinf_y);
    if (!rc && steps > 1 && !c->graph_exec) {
        VX_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        rc = device_step(c, vinf_x, vinf_y);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (!rc && e != hipSuccess) {
            set_error(std::string("graph capture: ") + hipGetErrorString(e));
            rc = 1;
        }
        if (!rc) rc = instantiate(g, &c->graph_exec);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
        } else {
            c->graph = g;
            c->graph_cap = c->cap;
            c->graph_vinf[0] = vinf_x;
            c->graph_vinf[1] = vinf_y;
        }
    }
    for (int64_t k = 1; !rc && k < steps; ++k) {
        const hipError_t e = hipGraphLaunch(c->graph_exec, s);
        if (e != hipSuccess) {
            set_error(std::string("graph launch: ") + hipGetErrorString(e));
            rc = 1;
        }
    }
    c->devcount = false;
    if (rc) {
        (void)hipStreamSynchronize(s);
        return rc;
    }
    return devcount_end(c, steps, f4, counts);
}

}  // extern "C"

// ---- batches of cases ---------------------------------------------------------
// Independent cases advanced together: every case's device-count step is a
// branch of one graph (captured from a batch stream that forks to the cases'
// streams and joins them back), so one graph launch advances every case by a
// step and the GPU runs the cases' kernels side by side.
struct vx_batch {
    std::vector<vx_case *> cases;
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<int64_t> caps;
    std::vector<double> vinf;
};

static void batch_graph_release(vx_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

extern "C" {

int vx_batch_create(vx_case *const *cases, int64_t k, vx_batch **out) {
    VX_ARG(cases && k > 0 && out, "cases");
    for (int64_t i = 0; i < k; ++i) {
        VX_ARG(cases[i], "null case");
        for (int64_t j = 0; j < i; ++j) VX_ARG(cases[i] != cases[j], "a case appears twice");
    }
    vx_batch *b = new vx_batch();
    b->cases.assign(cases, cases + k);
    b->join.assign((size_t)k, nullptr);
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    for (int64_t i = 0; e == hipSuccess && i < k; ++i) e = hipEventCreateWithFlags(&b->join[(size_t)i], hipEventDisableTiming);
    if (e != hipSuccess) {
        set_error(std::string("batch: ") + hipGetErrorString(e));
        vx_batch_destroy(b);
        return 1;
    }
    *out = b;
    return 0;
}

int vx_batch_destroy(vx_batch *b) {
    if (!b) return 0;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_graph_release(b);
    for (hipEvent_t e : b->join)
        if (e) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

// Every case advances `steps` steps (vinf: 2 per case); forces4 (4 per case
// and step, case-major) and counts (per case and step) are optional.
int vx_batch_run_steps(vx_batch *b, int64_t steps, const double *vinf, double *forces4, int64_t *counts) {
    VX_ARG(b && vinf && steps >= 0, "batch arguments");
    if (steps == 0) return 0;
    const size_t k = b->cases.size();
    bool stale = b->caps.size() != k || b->vinf.size() != 2 * k;
    int rc = 0;
    for (size_t i = 0; i < k && !rc; ++i) {
        vx_case *c = b->cases[i];
        bool moved = false;
        rc = devcount_begin(c, steps, &moved);
        stale = stale || moved || b->caps[i] != c->cap || b->vinf[2 * i] != vinf[2 * i] ||
                b->vinf[2 * i + 1] != vinf[2 * i + 1];
    }
    // the first step of every case on its own stream, uncaptured
    for (size_t i = 0; i < k && !rc; ++i) rc = device_step(b->cases[i], vinf[2 * i], vinf[2 * i + 1]);
    if (!rc && steps > 1) {
        if (stale) batch_graph_release(b);
        // the batch stream waits for every case's first step
        for (size_t i = 0; i < k && !rc; ++i)
            if (hipEventRecord(b->join[i], b->cases[i]->stream) != hipSuccess ||
                hipStreamWaitEvent(b->stream, b->join[i], 0) != hipSuccess)
                rc = 1;
        if (!rc && !b->exec) {
            hipGraph_t g = nullptr;
            if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess ||
                hipEventRecord(b->fork, b->stream) != hipSuccess) {
                set_error("batch: capture could not start");
                rc = 1;
            }
            for (size_t i = 0; i < k && !rc; ++i) {
                vx_case *c = b->cases[i];
                if (hipStreamWaitEvent(c->stream, b->fork, 0) != hipSuccess) rc = 1;
                if (!rc) rc = device_step(c, vinf[2 * i], vinf[2 * i + 1]);
                if (!rc && (hipEventRecord(b->join[i], c->stream) != hipSuccess ||
                            hipStreamWaitEvent(b->stream, b->join[i], 0) != hipSuccess))
                    rc = 1;
            }
            const hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (!rc && e != hipSuccess) {
                set_error(std::string("batch capture: ") + hipGetErrorString(e));
                rc = 1;
            }
            if (!rc) rc = instantiate(g, &b->exec);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
            } else {
                b->graph = g;
                b->caps.resize(k);
                b->vinf.assign(vinf, vinf + 2 * k);
                for (size_t i = 0; i < k; ++i) b->caps[i] = b->cases[i]->cap;
            }
        }
        for (int64_t s = 1; !rc && s < steps; ++s)
            if (hipGraphLaunch(b->exec, b->stream) != hipSuccess) {
                set_error("batch: graph launch failed");
                rc = 1;
            }
        // every case stream then waits for the replays
        if (!rc && hipEventRecord(b->fork, b->stream) != hipSuccess) rc = 1;
        for (size_t i = 0; i < k && !rc; ++i)
            if (hipStreamWaitEvent(b->cases[i]->stream, b->fork, 0) != hipSuccess) rc = 1;
    }
    for (size_t i = 0; i < k; ++i) b->cases[i]->devcount = false;
    if (rc) {
        (void)hipStreamSynchronize(b->stream);
        for (size_t i = 0; i < k; ++i) (void)hipStreamSynchronize(b->cases[i]->stream);
        if (vx_last_error()[0] == 0) set_error("batch: stream ordering failed");
        return 1;
    }
    for (size_t i = 0; i < k; ++i)
        if (devcount_end(b->cases[i], steps, forces4 ? forces4 + 4 * steps * (int64_t)i : nullptr,
                         counts ? counts + steps * (int64_t)i : nullptr))
            return 1;
    return 0;
}

// Second component move (src/unit_1.cu:1295,1334): step_Kernel on the points
// with null force pointers.
int vx_points_move(vx_case *c, double *x, double *y, double *g, const double *u, const double *v, int64_t m) {
    VX_ARG(c && m >= 0 && (m == 0 || (x && y && g && u && v)), "null pointer");
    if (!m) return 0;
    hipStream_t s = c->stream;
    if (c->px.ensure((size_t)m, s) || c->py.ensure((size_t)m, s) || c->pg.ensure((size_t)m, s) ||
        c->pu.ensure((size_t)m, s) || c->pv.ensure((size_t)m, s))
        return 1;
    const size_t b = sizeof(double) * (size_t)m;
    VX_CHECK(hipMemcpyAsync(c->px.p, x, b, hipMemcpyHostToDevice, s));
    VX_CHECK(hipMemcpyAsync(c->py.p, y, b, hipMemcpyHostToDevice, s));
    VX_CHECK(hipMemcpyAsync(c->pg.p, g, b, hipMemcpyHostToDevice, s));
    VX_CHECK(hipMemcpyAsync(c->pu.p, u, b, hipMemcpyHostToDevice, s));
    VX_CHECK(hipMemcpyAsync(c->pv.p, v, b, hipMemcpyHostToDevice, s));
    if (launch_points_move(c, m)) return 1;
    VX_CHECK(hipMemcpyAsync(x, c->px.p, b, hipMemcpyDeviceToHost, s));
    VX_CHECK(hipMemcpyAsync(y, c->py.p, b, hipMemcpyDeviceToHost, s));
    VX_CHECK(hipMemcpyAsync(g, c->pg.p, b, hipMemcpyDeviceToHost, s));
    VX_CHECK(hipStreamSynchronize(s));
    return 0;
}

// second_speed_Kernel (src/kernel.cu:1478-1489): the inertial points' velocity
// relaxes towards the carrier flow, v = (v_env - v_prev) / rel_t * dt + v_prev,
// with v_prev read as FP32 (float2 in the reference); 0 where g = 0.
int vx_points_relax(const double *g, const double *uenv, const double *venv, double *u, double *v, int64_t m,
                    double rel_t, double dt) {
#pragma clang fp contract(off)
    VX_ARG(m >= 0 && (m == 0 || (g && uenv && venv && u && v)), "null pointer");
    VX_ARG(rel_t != 0.0, "rel_t must be non-zero");
    for (int64_t i = 0; i < m; ++i) {
        if (!g[i]) {
            u[i] = 0.0;
            v[i] = 0.0;
            continue;
        }
        const double pu = (double)(float)u[i], pv = (double)(float)v[i];
        u[i] = (uenv[i] - pu) / rel_t * dt + pu;
        v[i] = (venv[i] - pv) / rel_t * dt + pv;
    }
    return 0;
}

int vx_case_set_probes(vx_case *c, const double *px, const double *py, int64_t m) {
    VX_ARG(c && m >= 0 && (m == 0 || (px && py)), "probe points");
    hipStream_t s = c->stream;
    if (m) {
        if (c->prb_x.ensure((size_t)m, s) || c->prb_y.ensure((size_t)m, s) || c->prb_u.ensure((size_t)m, s) ||
            c->prb_v.ensure((size_t)m, s))
            return 1;
        VX_CHECK(hipMemcpyAsync(c->prb_x.p, px, sizeof(double) * m, hipMemcpyHostToDevice, s));
        VX_CHECK(hipMemcpyAsync(c->prb_y.p, py, sizeof(double) * m, hipMemcpyHostToDevice, s));
        VX_CHECK(hipMemsetAsync(c->prb_u.p, 0, sizeof(double) * m, s));
        VX_CHECK(hipMemsetAsync(c->prb_v.p, 0, sizeof(double) * m, s));
        VX_CHECK(hipStreamSynchronize(s));
    }
    c->prb_m = m;
    return 0;
}

int vx_case_get_probes(vx_case *c, double *u, double *v, int64_t m) {
    VX_ARG(c && u && v && m == c->prb_m, "m must equal the probe count");
    if (!m) return 0;
    VX_CHECK(hipMemcpyAsync(u, c->prb_u.p, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
    VX_CHECK(hipMemcpyAsync(v, c->prb_v.p, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
    VX_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

// velocity_control (src/unit_1.cu:1254-1269): sources [0, ceil32(n)).
int vx_velocity_probe(vx_case *c, const double *px, const double *py, int64_t m, double vinf_x, double vinf_y,
                      double *u, double *v) {
    VX_ARG(c && px && py && u && v && m >= 0, "null pointer");
    if (!m) return 0;
    hipStream_t s = c->stream;
    if (c->px.ensure((size_t)m, s) || c->py.ensure((size_t)m, s) || c->pu.ensure((size_t)m, s) ||
        c->pv.ensure((size_t)m, s))
        return 1;
    VX_CHECK(hipMemcpyAsync(c->px.p, px, sizeof(double) * m, hipMemcpyHostToDevice, s));
    VX_CHECK(hipMemcpyAsync(c->py.p, py, sizeof(double) * m, hipMemcpyHostToDevice, s));
    if (launch_probe(c, m, std::min(ceil32(c->n), c->cap), vinf_x, vinf_y)) return 1;
    VX_CHECK(hipMemcpyAsync(u, c->pu.p, sizeof(double) * m, hipMemcpyDeviceToHost, s));
    VX_CHECK(hipMemcpyAsync(v, c->pv.p, sizeof(double) * m, hipMemcpyDeviceToHost, s));
    VX_CHECK(hipStreamSynchronize(s));
    return 0;
}

// vx_velocity_probe on caller-owned device buffers, allocating nothing:
// queued on the case stream, no host wait.
int vx_velocity_probe_noalloc(vx_case *c, const double *px, const double *py, int64_t m, double vinf_x,
                              double vinf_y, double *u, double *v) {
    VX_ARG(c && m >= 0 && (m == 0 || (px && py && u && v)), "null pointer");
    return launch_probe_noalloc(c, px, py, m, vinf_x, vinf_y, u, v);
}

}  // extern "C"
