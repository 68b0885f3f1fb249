// Run surface: run files, profiles, snapshots (see vx_io.h).
#include "vx_io.h"

#include <cinttypes>
#include <cmath>
#include <cstring>

namespace vxio {

namespace {

bool starts(const char *buf, const char *key) {
    const size_t n = strlen(key);
    return strncmp(buf, key, n) == 0 && buf[n] == ' ';
}

template <typename T>
void scan_num(const char *buf, T &out);

template <>
void scan_num<int64_t>(const char *buf, int64_t &out) {
    unsigned long long v = 0;
    if (sscanf(buf, "%*s %llu", &v) == 1) out = (int64_t)v;
}
template <>
void scan_num<double>(const char *buf, double &out) {
    double v = 0;
    if (sscanf(buf, "%*s %lf", &v) == 1) out = v;
}
template <>
void scan_num<int>(const char *buf, int &out) {
    unsigned v = 0;
    if (sscanf(buf, "%*s %u", &v) == 1) out = (int)(v & 0xFF);
}

void scan_str(const char *buf, std::string &out) {
    char s[256] = {0};
    if (sscanf(buf, "%*s %255s", s) == 1) out = s;
}

}  // namespace

int read_conf(const std::string &path, Conf &c, std::string *err) {
    FILE *f = fopen(path.c_str(), "r");
    if (!f) {
        if (err) *err = "unable to open config file " + path;
        return 1;
    }
    char buf[236];  // the reference reads lines through a 236-byte buffer
    while (fgets(buf, sizeof(buf), f)) {
#define STR(k) else if (starts(buf, #k)) scan_str(buf, c.k)
#define NUM(k) else if (starts(buf, #k)) scan_num(buf, c.k)
        if (false) {
        }
        STR(pr_file);
        STR(timings_file);
        STR(kadr_file);
        STR(second_points_file);
        NUM(steps);
        NUM(saving_step);
        NUM(dt);
        NUM(v_inf_incr_steps);
        NUM(inc_step);
        NUM(viscosity);
        NUM(rho);
        NUM(v_inf_x);
        NUM(v_inf_y);
        NUM(n_of_points);
        NUM(x_max);
        NUM(x_min);
        NUM(y_max);
        NUM(y_min);
        NUM(ve_size);
        NUM(n_col);
        NUM(h_col_x);
        NUM(h_col_y);
        NUM(rc_x);
        NUM(rc_y);
        NUM(r_col_diff_sign);
        NUM(r_col_same_sign);
        NUM(matrix_load);
        NUM(max_ve_g);
        NUM(rel_t);
        NUM(steady_flow);
        NUM(tree_depth);
        NUM(theta);
        NUM(no_log_buf);
        NUM(tree_min_n);
        NUM(tree_theta);
        NUM(order);
        NUM(diff_cut);
        NUM(birth_offset);
        NUM(matrix_precise);
        NUM(checkpoint_step);
        STR(checkpoint_file);
        else if (starts(buf, "precision")) {
            std::string v;
            scan_str(buf, v);
            c.precision = (v == "mixed" || v == "1") ? VX_PREC_MIXED : VX_PREC_F64;
        }
#undef STR
#undef NUM
    }
    fclose(f);
    return 0;
}

int read_profile(const std::string &path, std::vector<vx_panel> &pan, std::string *err) {
    FILE *f = fopen(path.c_str(), "r");
    if (!f) {
        if (err) *err = "can't open profile file " + path;
        return 1;
    }
    char buf[255];
    if (!fgets(buf, sizeof(buf), f) || !fgets(buf, sizeof(buf), f)) {
        fclose(f);
        if (err) *err = "profile header missing in " + path;
        return 1;
    }
    unsigned long long P = 0;
    if (fscanf(f, "%llu", &P) != 1 || P < 2 || P > (1ull << 31)) {
        fclose(f);
        if (err) *err = "bad panel count in " + path;
        return 1;
    }
    pan.assign((size_t)P, vx_panel{});
    for (size_t i = 0; i < P; ++i) {
        vx_panel &p = pan[i];
        const int k = fscanf(f, " %u %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %u %u", &p.n, &p.left[0],
                             &p.left[1], &p.right[0], &p.right[1], &p.contr[0], &p.contr[1], &p.birth[0],
                             &p.birth[1], &p.norm[0], &p.norm[1], &p.tang[0], &p.tang[1], &p.length,
                             &p.n_of_lpanel, &p.n_of_rpanel);
        if (k != 16 || p.n_of_lpanel >= P || p.n_of_rpanel >= P) {
            fclose(f);
            if (err) *err = "bad panel record " + std::to_string(i) + " in " + path;
            return 1;
        }
    }
    fclose(f);
    return 0;
}

void profile_box(const std::vector<vx_panel> &pan, double box[4]) {
    box[0] = box[2] = 1e300;
    box[1] = box[3] = -1e300;
    for (const auto &p : pan) {
        box[0] = std::fmin(box[0], p.left[0]);
        box[1] = std::fmax(box[1], p.left[0]);
        box[2] = std::fmin(box[2], p.left[1]);
        box[3] = std::fmax(box[3], p.left[1]);
    }
}

int write_kadr(const std::string &path, const double *x, const double *y, const double *g, int64_t n, double eps) {
    FILE *f = fopen(path.c_str(), "w");
    if (!f) return 1;
    fprintf(f, "%" PRId64 "\n", n);
    for (int64_t i = 0; i < n; ++i)
        fprintf(f, "%" PRId64 " %lf %lf %lf %lf %lf %lf %lf\n", i, eps, x[i], y[i], 0.0, 0.0, 0.0, g[i]);
    fclose(f);
    return 0;
}

int read_kadr(const std::string &path, std::vector<double> &x, std::vector<double> &y, std::vector<double> &g,
              std::string *err) {
    FILE *f = fopen(path.c_str(), "r");
    if (!f) {
        if (err) *err = "error file opening " + path;
        return 1;
    }
    unsigned long long n = 0;
    if (fscanf(f, "%llu", &n) != 1) {
        fclose(f);
        if (err) *err = "bad snapshot header in " + path;
        return 1;
    }
    x.resize(n);
    y.resize(n);
    g.resize(n);
    for (size_t i = 0; i < n; ++i) {
        if (fscanf(f, " %*u %*f %lf %lf %*f %*f %*f %lf", &x[i], &y[i], &g[i]) != 3) {
            fclose(f);
            if (err) *err = "bad snapshot row " + std::to_string(i) + " in " + path;
            return 1;
        }
    }
    fclose(f);
    return 0;
}

int kadr_step(const std::string &path) {
    const size_t p = path.rfind("Kadr");
    if (p == std::string::npos) return -1;
    int v = 0;
    if (sscanf(path.c_str() + p + 4, "%d", &v) != 1) return -1;
    return v;
}

int write_matrix(const std::string &path, const double *M, int64_t m, bool precise) {
    FILE *f = fopen(path.c_str(), "w");
    if (!f) return 1;
    fprintf(f, "%" PRId64 "\n", m);
    for (int64_t i = 0; i < m; ++i) {
        for (int64_t j = 0; j < m; ++j) fprintf(f, precise ? "%.17g\t" : "%lf\t", M[i * m + j]);
        fprintf(f, "\n");
    }
    fclose(f);
    return 0;
}

int read_matrix(const std::string &path, std::vector<double> &M, int64_t *m, std::string *err) {
    FILE *f = fopen(path.c_str(), "r");
    if (!f) {
        if (err) *err = "unable to open matrix file " + path;
        return 1;
    }
    unsigned long long s = 0;
    if (fscanf(f, "%llu", &s) != 1 || s < 1 || s > 100000) {
        fclose(f);
        if (err) *err = "bad matrix size in " + path;
        return 1;
    }
    M.resize(s * s);
    for (size_t k = 0; k < s * s; ++k)
        if (fscanf(f, "%lf", &M[k]) != 1) {
            fclose(f);
            if (err) *err = "short matrix file " + path;
            return 1;
        }
    fclose(f);
    *m = (int64_t)s;
    return 0;
}

vx_params params_from(const Conf &c, const std::vector<vx_panel> &pan) {
    vx_params p{};
    double box[4];
    profile_box(pan, box);
    p.dt = c.dt;
    p.ve_size = c.ve_size;
    p.nu = c.viscosity;
    p.rho = c.rho;
    p.vinf_x = c.v_inf_x;
    p.vinf_y = c.v_inf_y;
    p.x_min = c.x_min;
    p.x_max = c.x_max;
    p.y_min = c.y_min;
    p.y_max = c.y_max;
    p.px_min = box[0];
    p.px_max = box[1];
    p.py_min = box[2];
    p.py_max = box[3];
    p.r_col_diff_sign = c.r_col_diff_sign;
    p.r_col_same_sign = c.r_col_same_sign;
    p.max_ve_g = c.max_ve_g;
    p.rc_x = c.rc_x;
    p.rc_y = c.rc_y;
    p.birth_offset_rel = c.birth_offset;
    p.n_of_points = c.n_of_points;
    p.n_col = c.n_col;
    p.inc_step = c.inc_step > 0 ? c.inc_step : 8192;
    p.precision = c.precision;
    // the reference switches its tree on with tree_depth > 1 (its own theta
    // has another meaning and is not used here)
    p.tree_min_n = (int32_t)(c.tree_min_n > 0 ? c.tree_min_n : (c.tree_depth > 1 ? 1 : 0));
    p.theta = c.tree_theta;
    p.diff_cut = c.diff_cut;
    p.order = c.order;
    p.flags = 0;
    return p;
}

}  // namespace vxio
