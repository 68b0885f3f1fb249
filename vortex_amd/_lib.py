"""ctypes binding of libvortex_amd.so (include/vortex.h).

The shared library is built in-tree (vortex_amd/libvortex_amd.so) by
``__graft_entry__.build()`` / ``make -C vortex_amd/csrc``.  There is no
fallback: if the library is missing or cannot be loaded, every entry point
raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvortex_amd.so")

VX_PREC_F64 = 0
VX_PREC_MIXED = 1
VX_FLAG_EXACT_NEAREST = 1
VX_FLAG_LU_REFINE = 2

EXPORTS = [
    "vx_abi_version", "vx_last_error", "vx_device_count",
    "vx_case_create", "vx_case_destroy", "vx_case_build_matrix", "vx_case_set_matrix",
    "vx_case_get_matrix", "vx_case_get_panel_matrix", "vx_case_reserve", "vx_case_set_state",
    "vx_case_get_state", "vx_case_get_velocity", "vx_case_set_velocity", "vx_case_n",
    "vx_case_capacity", "vx_case_set_n", "vx_case_d_g", "vx_case_set_d_g", "vx_case_sync", "vx_case_save", "vx_case_load",
    "vx_vort_creation", "vx_speed", "vx_step", "vx_time_step", "vx_run_steps",
    "vx_velocity_probe", "vx_right_part", "vx_birth", "vx_biot_savart", "vx_diffusion",
    "vx_diffusion_wall", "vx_step_move", "vx_compact", "vx_setka", "vx_collapse",
    "vx_case_timings", "vx_case_enable_timing", "vx_tree_biot_savart", "vx_tree_diffusion",
    "vx_tree_stats", "vx_points_move", "vx_points_relax", "vx_tree_order", "vx_tree_census",
    "vx_case_set_probes", "vx_case_get_probes", "vx_velocity_probe_noalloc",
    "vx_case_solve_info", "vx_run_steps_graph", "vx_batch_create", "vx_batch_run_steps", "vx_batch_destroy",
    "vx_case_autotune", "vx_case_set_tuning", "vx_case_get_tuning",
]


class Params(C.Structure):
    """Mirror of vx_params (include/vortex.h)."""
    _fields_ = [(n, C.c_double) for n in (
        "dt", "ve_size", "nu", "rho", "vinf_x", "vinf_y",
        "x_min", "x_max", "y_min", "y_max",
        "px_min", "px_max", "py_min", "py_max",
        "r_col_diff_sign", "r_col_same_sign", "max_ve_g", "rc_x", "rc_y",
        "birth_offset_rel")] + [
        ("n_of_points", C.c_int64), ("n_col", C.c_int64), ("inc_step", C.c_int64),
        ("precision", C.c_int32), ("tree_min_n", C.c_int32),
        ("theta", C.c_double), ("diff_cut", C.c_double), ("order", C.c_int32), ("flags", C.c_int32)]


class TreeInfo(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("m", "leaves", "nodes", "fallback", "bs_pairs", "m2p_evals",
                                          "diff_pairs")] + [("order", C.c_int32), ("heavy", C.c_int32),
                                                            ("m2l", C.c_int64), ("diff_in_radius", C.c_int64)]


class Tuning(C.Structure):
    """Mirror of vx_tuning (include/vortex.h)."""
    _fields_ = [("tpt_bs", C.c_int32), ("tpt_diff", C.c_int32), ("leaf_off", C.c_int32), ("trials", C.c_int32),
                ("ms_default", C.c_double), ("ms_best", C.c_double), ("col_share_min", C.c_int64)]


class Timings(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("creation_ms", "speed_ms", "step_ms", "bs_ms",
                                           "diff_ms", "wall_ms")] + [("bs_launches", C.c_int64),
                                                                      ("tree_build_ms", C.c_double),
                                                                      ("bs_kernel_ms", C.c_double),
                                                                      ("diff_kernel_ms", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(make -C vortex_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    vp, i64, d, dp, ip = C.c_void_p, C.c_int64, C.c_double, C.c_void_p, C.c_void_p
    sig = {
        "vx_abi_version": ([], C.c_int),
        "vx_last_error": ([], C.c_char_p),
        "vx_device_count": ([C.POINTER(C.c_int)], C.c_int),
        "vx_case_create": ([C.POINTER(Params), vp, i64, C.POINTER(vp)], C.c_int),
        "vx_case_destroy": ([vp], C.c_int),
        "vx_case_build_matrix": ([vp], C.c_int),
        "vx_case_set_matrix": ([vp, dp, i64], C.c_int),
        "vx_case_get_matrix": ([vp, dp, i64], C.c_int),
        "vx_case_get_panel_matrix": ([vp, dp, i64], C.c_int),
        "vx_case_reserve": ([vp, i64], C.c_int),
        "vx_case_set_state": ([vp, dp, dp, dp, i64], C.c_int),
        "vx_case_get_state": ([vp, dp, dp, dp, i64], C.c_int),
        "vx_case_get_velocity": ([vp, dp, dp, dp, i64], C.c_int),
        "vx_case_set_velocity": ([vp, dp, dp, dp, i64], C.c_int),
        "vx_case_n": ([vp], i64),
        "vx_case_capacity": ([vp], i64),
        "vx_case_set_n": ([vp, i64], C.c_int),
        "vx_case_d_g": ([vp], d),
        "vx_case_set_d_g": ([vp, d], C.c_int),
        "vx_case_sync": ([vp], C.c_int),
        "vx_case_save": ([vp, C.c_char_p, vp, i64], C.c_int),
        "vx_case_load": ([vp, C.c_char_p, vp, i64, C.POINTER(C.c_int64)], C.c_int),
        "vx_vort_creation": ([vp, d, d], C.c_int),
        "vx_speed": ([vp, d, d], C.c_int),
        "vx_step": ([vp, dp], C.c_int),
        "vx_time_step": ([vp, d, d, dp], C.c_int),
        "vx_run_steps": ([vp, i64, d, d, dp], C.c_int),
        "vx_velocity_probe": ([vp, dp, dp, i64, d, d, dp, dp], C.c_int),
        "vx_right_part": ([vp, i64, d, d, dp], C.c_int),
        "vx_birth": ([vp, dp, d, dp], C.c_int),
        "vx_biot_savart": ([vp, i64, i64, d, d], C.c_int),
        "vx_diffusion": ([vp, i64, i64], C.c_int),
        "vx_diffusion_wall": ([vp, i64], C.c_int),
        "vx_step_move": ([vp, dp], C.c_int),
        "vx_compact": ([vp], C.c_int),
        "vx_setka": ([vp, C.c_int, ip], C.c_int),
        "vx_collapse": ([vp, C.c_int, ip], C.c_int),
        "vx_case_timings": ([vp, C.POINTER(Timings)], C.c_int),
        "vx_case_enable_timing": ([vp, C.c_int], C.c_int),
        "vx_tree_biot_savart": ([vp, i64, d, d], C.c_int),
        "vx_tree_diffusion": ([vp], C.c_int),
        "vx_tree_stats": ([vp, C.POINTER(TreeInfo)], C.c_int),
        "vx_points_move": ([vp, dp, dp, dp, dp, dp, i64], C.c_int),
        "vx_points_relax": ([dp, dp, dp, dp, dp, i64, d, d], C.c_int),
        "vx_tree_order": ([vp, dp, dp, i64], C.c_int),
        "vx_tree_census": ([vp, C.c_int], C.c_int),
        "vx_case_set_probes": ([vp, dp, dp, i64], C.c_int),
        "vx_case_get_probes": ([vp, dp, dp, i64], C.c_int),
        "vx_velocity_probe_noalloc": ([vp, vp, vp, i64, d, d, vp, vp], C.c_int),
        "vx_run_steps_graph": ([vp, i64, d, d, dp, dp], C.c_int),
        "vx_batch_create": ([vp, i64, C.POINTER(vp)], C.c_int),
        "vx_batch_run_steps": ([vp, i64, dp, dp, dp], C.c_int),
        "vx_batch_destroy": ([vp], C.c_int),
        "vx_case_autotune": ([vp, d, d, C.POINTER(Tuning)], C.c_int),
        "vx_case_set_tuning": ([vp, C.POINTER(Tuning)], C.c_int),
        "vx_case_get_tuning": ([vp, C.POINTER(Tuning)], C.c_int),
        "vx_case_solve_info": ([vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_double)], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


class VortexError(RuntimeError):
    pass


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().vx_last_error().decode(errors="replace")
        raise VortexError(f"libvortex_amd error {rc}: {msg}")


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)
