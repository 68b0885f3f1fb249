// Run surface of the drop-in solver: run files, profiles, snapshots, force
// history.  Formats follow the reference so that its .conf files, profile
// files and restart snapshots are read unchanged and its outputs are written
// in the same layout.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "vortex.h"

namespace vxio {

// Keys of read_config (src/main.c:69-115) plus the ones this solver adds
// (the reference ignores unknown keys, so run files stay compatible).
struct Conf {
    std::string pr_file, timings_file, kadr_file, second_points_file;
    int64_t steps = 0, saving_step = 1, v_inf_incr_steps = 1, inc_step = 8192, n_of_points = 20, n_col = 0;
    double dt = 0, viscosity = 0, rho = 0, v_inf_x = 0, v_inf_y = 0;
    double x_max = 0, x_min = 0, y_max = 0, y_min = 0, ve_size = 0;
    double h_col_x = 0, h_col_y = 0, rc_x = 0, rc_y = 0, r_col_diff_sign = 0, r_col_same_sign = 0;
    double max_ve_g = 0, rel_t = 0, theta = 0;
    int64_t tree_depth = 0;
    int matrix_load = 0, steady_flow = 0, no_log_buf = 0;
    // additions
    int precision = VX_PREC_F64;  // "precision f64|mixed"
    int64_t tree_min_n = 0;       // "tree_min_n N"
    int order = 16;               // "order 16|32"
    double tree_theta = 0;        // "tree_theta t": opening angle of the Morton tree
    double diff_cut = 0;          // "diff_cut c"
    double birth_offset = 0;      // "birth_offset r" (multiple of the panel length)
    int matrix_precise = 0;       // "matrix_precise 1": M.txt with 17 digits
    int64_t checkpoint_step = 0;  // "checkpoint_step k": exact checkpoint every k steps
    std::string checkpoint_file;  // "checkpoint_file path": restart from an exact checkpoint
};

// Parses "key value # comment" lines; a key matches when the line starts with
// "key " (the reference's strncmp rule); unknown keys are ignored.  Returns 0,
// or 1 when the file cannot be opened.
int read_conf(const std::string &path, Conf &c, std::string *err);

// Profile file (src/main.c:220-256): two header lines, panel count, then P
// records of 16 columns; anything after the records is ignored.
int read_profile(const std::string &path, std::vector<vx_panel> &pan, std::string *err);
void profile_box(const std::vector<vx_panel> &pan, double box[4]);  // from the left end points

// Kadr snapshot (src/main.c:175-203): "N" then rows "i eps x y 0 0 0 g".
int write_kadr(const std::string &path, const double *x, const double *y, const double *g, int64_t n, double eps);
int read_kadr(const std::string &path, std::vector<double> &x, std::vector<double> &y, std::vector<double> &g,
              std::string *err);
// Step number encoded in a snapshot name "...KadrNNNNNN..." (src/main.c:594-605); -1 if none.
int kadr_step(const std::string &path);

// M.txt (src/unit_1.cu:47-65, 404-456): "size" then rows of tab-separated values.
int write_matrix(const std::string &path, const double *M, int64_t m, bool precise);
int read_matrix(const std::string &path, std::vector<double> &M, int64_t *m, std::string *err);

vx_params params_from(const Conf &c, const std::vector<vx_panel> &pan);

}  // namespace vxio
