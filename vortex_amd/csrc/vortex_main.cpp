// vortex -c file.conf [-l level] [-o logfile]
//
// Drop-in run driver over libvortex_amd.so.  It follows the reference's
// main() (src/main.c:403-830): read the run file and the profile, build or
// load the panel-system inverse, start from an empty wake or a Kadr snapshot
// (the start step is taken from the file name), then per step:
//   ramp V_inf, birth, timings line, every saving_step the Kadr snapshot and
//   the probe velocities, the second component (inertial, then passive
//   tracer points), probe velocities of this step, the force line of the
//   previous step, Speed, Step
// and finally the last snapshot and force line.  Outputs go to
// output/kadrs/KadrNNNNNN.txt, output/vels/contr_VelNNNNNN.txt,
// output/vels/VelNNNNNN_S.txt, output/first|second/KadrNNNNNN.txt and
// output/F_p.txt in the reference's text formats.  With steady_flow the
// wake is frozen (no birth, Speed or Step) and only the points move.
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <cinttypes>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "vortex.h"
#include "vx_io.h"

namespace {

int g_level = 3;
FILE *g_log = nullptr;

void logf(int lev, const char *fmt, ...) {
    if (g_level < lev) return;
    char msg[4096];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, sizeof(msg), fmt, ap);
    va_end(ap);
    timespec t;
    clock_gettime(CLOCK_REALTIME, &t);
    tm bd;
    localtime_r(&t.tv_sec, &bd);
    char ts[64];
    const size_t k = strftime(ts, sizeof(ts), "%d.%m.%Y %H:%M:%S", &bd);
    snprintf(ts + k, sizeof(ts) - k, ".%06ld", t.tv_nsec / 1000);
    static const char sym[] = "AEWIDB";
    fprintf(g_log ? g_log : stdout, "[%s] %c %s\n", ts, sym[lev % 6], msg);
}

int make_dir(const char *d) {
    struct stat st;
    if (stat(d, &st) == 0) return S_ISDIR(st.st_mode) ? 0 : -1;
    return mkdir(d, 0777);
}

int fail(vx_case *c, const char *what) {
    logf(1, "%s: %s", what, vx_last_error());
    if (c) vx_case_destroy(c);
    return 1;
}

// save_vel_to_file (src/unit_1.cu:616-631): output/vels/VelNNNNNN_S.txt.
void write_vels(int64_t step, int stage, const std::vector<double> &x, const std::vector<double> &y,
                const std::vector<double> &u, const std::vector<double> &v) {
    char fn[256];
    snprintf(fn, sizeof(fn), "output/vels/Vel%06" PRId64 "_%d.txt", step, stage);
    FILE *f = fopen(fn, "w");
    if (!f) {
        logf(1, "error file opening %s : %s", fn, strerror(errno));
        return;
    }
    fprintf(f, "%zu\n", x.size());
    for (size_t i = 0; i < x.size(); ++i) fprintf(f, "%zu %lf %lf %lf %lf\n", i, x[i], y[i], u[i], v[i]);
    fclose(f);
}

// One component of tracer points (src/main.c:480-509): positions, circulation
// flag (0 once absorbed or deleted) and velocity.
struct Points {
    std::vector<double> x, y, g, u, v;
    size_t size() const { return x.size(); }
    void resize_vel() {
        u.assign(x.size(), 0.0);
        v.assign(x.size(), 0.0);
    }
};

// The driver's own state in an exact checkpoint's blob: step, the force
// line still to be logged, V_inf, the probe rows since the last save and the
// tracer points.
struct Blob {
    std::vector<char> b;
    template <class T>
    void put(const T &v) {
        const char *p = reinterpret_cast<const char *>(&v);
        b.insert(b.end(), p, p + sizeof(T));
    }
    void put_vec(const std::vector<double> &v) {
        put((int64_t)v.size());
        const char *p = reinterpret_cast<const char *>(v.data());
        b.insert(b.end(), p, p + sizeof(double) * v.size());
    }
};
struct BlobReader {
    const char *p, *e;
    template <class T>
    bool get(T &v) {
        if (e - p < (std::ptrdiff_t)sizeof(T)) return false;
        memcpy(&v, p, sizeof(T));
        p += sizeof(T);
        return true;
    }
    bool get_vec(std::vector<double> &v) {
        int64_t n = 0;
        if (!get(n) || n < 0 || (e - p) / (std::ptrdiff_t)sizeof(double) < n) return false;
        v.resize((size_t)n);
        if (n) memcpy(v.data(), p, sizeof(double) * (size_t)n);
        p += sizeof(double) * (size_t)n;
        return true;
    }
};
void put_points(Blob &bl, const Points &q) {
    bl.put_vec(q.x); bl.put_vec(q.y); bl.put_vec(q.g); bl.put_vec(q.u); bl.put_vec(q.v);
}
bool get_points(BlobReader &r, Points &q) {
    return r.get_vec(q.x) && r.get_vec(q.y) && r.get_vec(q.g) && r.get_vec(q.u) && r.get_vec(q.v);
}

}  // namespace

int main(int argc, char **argv) {
    std::string conf_path, log_path;
    bool print_only = false;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "-p")) {
            print_only = true;
            continue;
        }
        if (argv[i][0] != '-' || i + 1 >= argc) continue;
        switch (argv[i][1]) {
            case 'c': conf_path = argv[++i]; break;
            case 'l': g_level = atoi(argv[++i]); break;
            case 'o': log_path = argv[++i]; break;
        }
    }
    if (conf_path.empty()) {
        printf("Usage:\nvortex -c file.conf\n");
        return 1;
    }
    vxio::Conf conf;
    std::string err;
    if (vxio::read_conf(conf_path, conf, &err)) {
        logf(1, "%s", err.c_str());
        return 1;
    }
    if (print_only) {  // parsed run file and profile, no device work
        std::vector<vx_panel> pan;
        const int rc = vxio::read_profile(conf.pr_file, pan, &err);
        const vx_params p = vxio::params_from(conf, pan);
        printf("{\"pr_file\": \"%s\", \"kadr_file\": \"%s\", \"timings_file\": \"%s\", \"steps\": %" PRId64
               ", \"saving_step\": %" PRId64 ", \"v_inf_incr_steps\": %" PRId64 ", \"matrix_load\": %d, "
               "\"profile_ok\": %d, \"P\": %zu, \"params\": {",
               conf.pr_file.c_str(), conf.kadr_file.c_str(), conf.timings_file.c_str(), conf.steps, conf.saving_step,
               conf.v_inf_incr_steps, conf.matrix_load, rc == 0, pan.size());
        printf("\"dt\": %.17g, \"ve_size\": %.17g, \"nu\": %.17g, \"rho\": %.17g, \"vinf_x\": %.17g, "
               "\"vinf_y\": %.17g, \"x_min\": %.17g, \"x_max\": %.17g, \"y_min\": %.17g, \"y_max\": %.17g, "
               "\"px_min\": %.17g, \"px_max\": %.17g, \"py_min\": %.17g, \"py_max\": %.17g, "
               "\"r_col_diff_sign\": %.17g, \"r_col_same_sign\": %.17g, \"max_ve_g\": %.17g, \"rc_x\": %.17g, "
               "\"rc_y\": %.17g, \"n_of_points\": %" PRId64 ", \"n_col\": %" PRId64 ", \"inc_step\": %" PRId64
               ", \"precision\": %d, \"tree_min_n\": %d}}\n",
               p.dt, p.ve_size, p.nu, p.rho, p.vinf_x, p.vinf_y, p.x_min, p.x_max, p.y_min, p.y_max, p.px_min,
               p.px_max, p.py_min, p.py_max, p.r_col_diff_sign, p.r_col_same_sign, p.max_ve_g, p.rc_x, p.rc_y,
               p.n_of_points, p.n_col, p.inc_step, p.precision, p.tree_min_n);
        return rc;
    }
    if (make_dir("output") || make_dir("output/kadrs") || make_dir("output/vels")) {
        logf(1, "cannot create output directories: %s", strerror(errno));
        return 1;
    }
    if (!log_path.empty()) g_log = fopen(log_path.c_str(), "w");
    FILE *timings = conf.timings_file.empty() ? nullptr : fopen(conf.timings_file.c_str(), "w");
    const bool have_points = !conf.second_points_file.empty();
    if (have_points && conf.rel_t == 0.0) {
        logf(1, "rel_t is nessesary if second component set");
        return 1;
    }
    if (!have_points && conf.steady_flow) {
        logf(1, "steady_flow and no second component");
        return 1;
    }

    std::vector<vx_panel> pan;
    if (vxio::read_profile(conf.pr_file, pan, &err)) {
        logf(1, "%s", err.c_str());
        return 1;
    }
    const int64_t P = (int64_t)pan.size();
    vx_params prm = vxio::params_from(conf, pan);
    logf(3, "profile gabarites: x_min = %lf x_max = %lf y_min = %lf y_max = %lf", prm.px_min, prm.px_max,
         prm.py_min, prm.py_max);
    vx_case *c = nullptr;
    if (vx_case_create(&prm, pan.data(), P, &c)) return fail(nullptr, "case creation");

    if (conf.matrix_load) {
        std::vector<double> M;
        int64_t m = 0;
        if (vxio::read_matrix("M.txt", M, &m, &err)) {
            logf(1, "%s", err.c_str());
            vx_case_destroy(c);
            return 1;
        }
        if (m != P + 1) {
            logf(1, "martix and profile mismatch");
            vx_case_destroy(c);
            return 1;
        }
        if (vx_case_set_matrix(c, M.data(), m)) return fail(c, "matrix load");
        logf(3, "Matrix loaded!");
    } else {
        if (vx_case_build_matrix(c)) return fail(c, "Matrix creation error!");
        std::vector<double> M((size_t)((P + 1) * (P + 1)));
        if (vx_case_get_matrix(c, M.data(), P + 1)) return fail(c, "matrix read-back");
        vxio::write_matrix("M.txt", M.data(), P + 1, conf.matrix_precise != 0);
        logf(3, "Matrix created!");
    }

    // second component: inertial (second) and passive (first) copies of the
    // points (src/main.c:480-509)
    Points first, second;
    if (have_points) {
        if (vxio::read_kadr(conf.second_points_file, second.x, second.y, second.g, &err)) {
            logf(1, "%s", err.c_str());
            vx_case_destroy(c);
            return 1;
        }
        first = second;
        first.resize_vel();
        second.resize_vel();
        if (make_dir("output/first") || make_dir("output/second")) {
            logf(1, "cannot create output directories: %s", strerror(errno));
            vx_case_destroy(c);
            return 1;
        }
    }

    int64_t step = 0;
    // probe points (src/main.c:521-526)
    const int64_t NPROBE = 500;
    const int64_t save = conf.saving_step > 0 ? conf.saving_step : 0;
    std::vector<double> px(NPROBE), py(NPROBE), pu(NPROBE), pv(NPROBE);
    for (int64_t k = 0; k < NPROBE; ++k) {
        px[k] = -0.15;
        py[k] = 0.01 + 0.002 * (double)k;
    }
    std::vector<double> vbuf((size_t)(2 * NPROBE * (save > 0 ? save : 1)), 0.0);
    int64_t vslot = 0;
    double vinf[2] = {conf.v_inf_x, conf.v_inf_y}, dv[2] = {0.0, 0.0};
    if (conf.v_inf_incr_steps) {
        dv[0] = conf.v_inf_x / (double)conf.v_inf_incr_steps;
        dv[1] = conf.v_inf_y / (double)conf.v_inf_incr_steps;
        vinf[0] = vinf[1] = 0.0;
    }
    double forces[4] = {0.0, 0.0, 0.0, 0.0};
    bool from_ckpt = false;
    if (!conf.checkpoint_file.empty()) {
        // exact restart: the case state and this driver's own
        std::vector<char> blob((size_t)(64 << 20));
        int64_t got = 0;
        if (vx_case_load(c, conf.checkpoint_file.c_str(), blob.data(), (int64_t)blob.size(), &got))
            return fail(c, "checkpoint load");
        BlobReader r{blob.data(), blob.data() + got};
        std::vector<double> vb;
        if (!r.get(step) || !r.get(forces) || !r.get(vinf) || !r.get(vslot) || !r.get_vec(vb) ||
            !get_points(r, second) || !get_points(r, first) || vb.size() != vbuf.size() ||
            second.size() != first.size() || vslot < 0 || vslot > (save > 0 ? save : 1)) {
            logf(1, "checkpoint %s: driver state does not match this run file", conf.checkpoint_file.c_str());
            vx_case_destroy(c);
            return 1;
        }
        vbuf = vb;
        conf.steps += step;
        from_ckpt = true;
        logf(3, "restart from checkpoint %s at step %" PRId64, conf.checkpoint_file.c_str(), step);
    }
    const int64_t ckpt = conf.checkpoint_step > 0 ? conf.checkpoint_step : 0;
    if (ckpt && make_dir("output/checkpoints")) {
        logf(1, "cannot create output/checkpoints: %s", strerror(errno));
        vx_case_destroy(c);
        return 1;
    }
    auto save_ckpt = [&](int64_t at) -> int {
        Blob bl;
        bl.put(at);
        bl.put(forces);
        bl.put(vinf);
        bl.put(vslot);
        bl.put_vec(vbuf);
        put_points(bl, second);
        put_points(bl, first);
        char fn[256];
        snprintf(fn, sizeof(fn), "output/checkpoints/Ckpt%06" PRId64 ".bin", at);
        return vx_case_save(c, fn, bl.b.data(), (int64_t)bl.b.size());
    };
    if (!from_ckpt && !conf.kadr_file.empty()) {
        std::vector<double> x, y, g;
        if (vxio::read_kadr(conf.kadr_file, x, y, g, &err)) {
            logf(1, "%s", err.c_str());
            vx_case_destroy(c);
            return 1;
        }
        if (vx_case_set_state(c, x.data(), y.data(), g.data(), (int64_t)x.size())) return fail(c, "restart");
        const int k = vxio::kadr_step(conf.kadr_file);
        if (k >= 0) {
            step = k;
            conf.steps += k;
        }
    }
    const int64_t first_step = step;
    if (step && have_points && !from_ckpt) {
        // restart of the points from their own snapshots (src/main.c:606-648)
        char fs[256], ff[256];
        snprintf(fs, sizeof(fs), "output/second/Kadr%06" PRId64 ".txt", step);
        snprintf(ff, sizeof(ff), "output/first/Kadr%06" PRId64 ".txt", step);
        struct stat st;
        if (stat(fs, &st) == 0 && !S_ISDIR(st.st_mode) && stat(ff, &st) == 0 && !S_ISDIR(st.st_mode)) {
            Points s2, f2;
            if (vxio::read_kadr(fs, s2.x, s2.y, s2.g, &err) || vxio::read_kadr(ff, f2.x, f2.y, f2.g, &err) ||
                s2.size() != f2.size()) {
                logf(1, "bad point snapshots at step %" PRId64 ": %s", step, err.c_str());
                vx_case_destroy(c);
                return 1;
            }
            second = s2;
            first = f2;
            first.resize_vel();
            second.resize_vel();
            logf(3, "new n_of_second = %zu", second.size());
        }
    }

    // a checkpoint restart continues the force history
    FILE *fp = fopen("output/F_p.txt", from_ckpt ? "a" : "w");
    vx_case_enable_timing(c, 1);
    const double eps_out = 0.008;  // src/main.c:517
    std::vector<double> hx, hy, hg;
    logf(2, "Start from step %" PRId64, step);
    for (; step < conf.steps; ++step) {
        if (ckpt && step % ckpt == 0 && step != first_step && save_ckpt(step)) return fail(c, "checkpoint");
        if (step < conf.v_inf_incr_steps) {
            vinf[0] += dv[0];
            vinf[1] += dv[1];
        }
        if (!conf.steady_flow && vx_vort_creation(c, vinf[0], vinf[1])) return fail(c, "Creation ERROR!");
        vx_timings tm;
        vx_case_timings(c, &tm);
        FILE *tf = timings ? timings : stdout;
        fprintf(tf, "%" PRId64 "\tN = %" PRId64 "\tCreation time = %f\tSpeed time = %f\tStep time = %f\n", step,
                vx_case_n(c), tm.creation_ms / 1e3, tm.speed_ms / 1e3, tm.step_ms / 1e3);
        if (save > 0 && step % save == 0 && !conf.steady_flow) {
            logf(2, "Output %" PRId64, step);
            const int64_t n = vx_case_n(c);
            hx.resize(n); hy.resize(n); hg.resize(n);
            if (vx_case_get_state(c, hx.data(), hy.data(), hg.data(), n)) return fail(c, "Saving ERROR at POS copy");
            char fn[256];
            snprintf(fn, sizeof(fn), "output/kadrs/Kadr%06" PRId64 ".txt", step);
            vxio::write_kadr(fn, hx.data(), hy.data(), hg.data(), n, eps_out);
            snprintf(fn, sizeof(fn), "output/vels/contr_Vel%06" PRId64 ".txt", step);
            if (FILE *vf = fopen(fn, "w")) {
                for (int64_t i = 0; i < NPROBE * save; ++i)
                    fprintf(vf, "%" PRId64 " %lf %lf %lf %lf\n", i, px[i % NPROBE], py[i % NPROBE], vbuf[2 * i],
                            vbuf[2 * i + 1]);
                fclose(vf);
            }
            vslot = 0;
        }
        if (save > 0 && step % save == 0 && have_points) {
            char fn[256];
            snprintf(fn, sizeof(fn), "output/second/Kadr%06" PRId64 ".txt", step);
            vxio::write_kadr(fn, second.x.data(), second.y.data(), second.g.data(), (int64_t)second.size(), eps_out);
            snprintf(fn, sizeof(fn), "output/first/Kadr%06" PRId64 ".txt", step);
            vxio::write_kadr(fn, first.x.data(), first.y.data(), first.g.data(), (int64_t)first.size(), eps_out);
        }
        if (have_points && second.size()) {
            // second_speed (src/unit_1.cu:1300-1340): flow velocity at the
            // inertial points, relaxation of their velocity, move
            const int64_t m = (int64_t)second.size();
            std::vector<double> ue(m), ve(m);
            if (vx_velocity_probe(c, second.x.data(), second.y.data(), m, vinf[0], vinf[1], ue.data(), ve.data()))
                return fail(c, "second Speed evaluation ERROR!");
            if (step == first_step && !from_ckpt) {
                second.u = ue;
                second.v = ve;
            }
            write_vels(step, 4, second.x, second.y, ue, ve);
            if (vx_points_relax(second.g.data(), ue.data(), ve.data(), second.u.data(), second.v.data(), m,
                                conf.rel_t, conf.dt))
                return fail(c, "second Speed evaluation ERROR!");
            write_vels(step, 5, second.x, second.y, second.u, second.v);
            if (vx_points_move(c, second.x.data(), second.y.data(), second.g.data(), second.u.data(),
                               second.v.data(), m))
                return fail(c, "second Speed evaluation ERROR!");
            // first_speed (src/unit_1.cu:1271-1298): passive points move with the flow
            if (vx_velocity_probe(c, first.x.data(), first.y.data(), m, vinf[0], vinf[1], first.u.data(),
                                  first.v.data()))
                return fail(c, "second Speed evaluation ERROR!");
            write_vels(step, 5, first.x, first.y, first.u, first.v);
            if (vx_points_move(c, first.x.data(), first.y.data(), first.g.data(), first.u.data(), first.v.data(),
                               m))
                return fail(c, "second Speed evaluation ERROR!");
        }
        if (conf.steady_flow) continue;
        if (vx_velocity_probe(c, px.data(), py.data(), NPROBE, vinf[0], vinf[1], pu.data(), pv.data()))
            return fail(c, "velocity control");
        if (save > 0 && vslot < save) {
            for (int64_t k = 0; k < NPROBE; ++k) {
                vbuf[2 * (vslot * NPROBE + k)] = pu[k];
                vbuf[2 * (vslot * NPROBE + k) + 1] = pv[k];
            }
            ++vslot;
        }
        if (fp) {
            fprintf(fp, "%" PRId64 " %lf %lf %lf\n", step, forces[1], forces[2], forces[3]);
            fflush(fp);
        }
        if (vx_speed(c, vinf[0], vinf[1])) return fail(c, "Speed evaluation ERROR!");
        if (vx_step(c, forces)) return fail(c, "Moving ERROR!");
    }
    if (ckpt && conf.steps % ckpt == 0 && conf.steps != first_step && save_ckpt(conf.steps))
        return fail(c, "checkpoint");
    {
        vx_timings tm;
        vx_case_timings(c, &tm);
        fprintf(timings ? timings : stdout, "%" PRId64 "\tN = %" PRId64 "\tCreation time = %f\tSpeed time = %f\tStep time = %f\n",
                conf.steps, vx_case_n(c), tm.creation_ms / 1e3, tm.speed_ms / 1e3, tm.step_ms / 1e3);
        const int64_t n = vx_case_n(c);
        hx.resize(n); hy.resize(n); hg.resize(n);
        if (vx_case_get_state(c, hx.data(), hy.data(), hg.data(), n)) return fail(c, "Saving ERROR at POS copy");
        char fn[256];
        snprintf(fn, sizeof(fn), "output/kadrs/Kadr%06" PRId64 ".txt", conf.steps);
        if (!conf.steady_flow) {
            vxio::write_kadr(fn, hx.data(), hy.data(), hg.data(), n, eps_out);
            if (fp) fprintf(fp, "%" PRId64 " %lf %lf %lf\n", conf.steps, forces[1], forces[2], forces[3]);
        }
        if (have_points) {
            snprintf(fn, sizeof(fn), "output/second/Kadr%06" PRId64 ".txt", conf.steps);
            vxio::write_kadr(fn, second.x.data(), second.y.data(), second.g.data(), (int64_t)second.size(), eps_out);
            snprintf(fn, sizeof(fn), "output/first/Kadr%06" PRId64 ".txt", conf.steps);
            vxio::write_kadr(fn, first.x.data(), first.y.data(), first.g.data(), (int64_t)first.size(), eps_out);
        }
    }
    if (fp) fclose(fp);
    if (timings) fclose(timings);
    logf(3, "ready!");
    vx_case_destroy(c);
    if (g_log) fclose(g_log);
    return 0;
}
