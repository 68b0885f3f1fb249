"""Python host interface to one flow case on the GPU.

``Case`` mirrors the reference's host routines (include/unit_1.cuh:20-52):
``build_matrix`` (matr_creation), ``vort_creation``, ``speed`` (Speed),
``step`` (Step), ``velocity_probe`` (velocity_control) and ``time_step`` (one
iteration of src/main.c:652-785), plus stage-level calls for testing.  All
arithmetic runs in libvortex_amd.so; this module only moves host arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import geometry as geo
from ._lib import (VX_FLAG_EXACT_NEAREST, VX_FLAG_LU_REFINE, VX_PREC_F64, VX_PREC_MIXED, Params, Timings, Tuning,
                   check, lib, ptr)

__all__ = ["Case", "Batch", "make_params", "VX_PREC_F64", "VX_PREC_MIXED", "VX_FLAG_EXACT_NEAREST", "VX_FLAG_LU_REFINE"]


def make_params(panels: np.ndarray, *, dt=0.01, ve_size=None, nu=1e-3, rho=1.0, vinf=(1.0, 0.0),
                box=(-2.0, 10.0, -10.0, 10.0), r_col_diff_sign=1.0, r_col_same_sign=0.3,
                max_ve_g=None, rc=(0.0, 0.0), birth_offset_rel=0.0, n_of_points=20, n_col=2,
                inc_step=8192, precision=VX_PREC_F64, tree_min_n=0, theta=0.5, diff_cut=0.0,
                order=16, flags=0) -> Params:
    """Case parameters with the profile box taken from the panels' left end points."""
    bx = geo.profile_bbox(panels)
    ve = float(panels["length"].mean()) if ve_size is None else ve_size
    if max_ve_g is None:
        max_ve_g = 2.0 * ve  # a few times a newborn's circulation at U = 1
    return Params(dt=dt, ve_size=ve, nu=nu, rho=rho, vinf_x=vinf[0], vinf_y=vinf[1],
                  x_min=box[0], x_max=box[1], y_min=box[2], y_max=box[3],
                  px_min=bx[0], px_max=bx[1], py_min=bx[2], py_max=bx[3],
                  r_col_diff_sign=r_col_diff_sign, r_col_same_sign=r_col_same_sign,
                  max_ve_g=max_ve_g, rc_x=rc[0], rc_y=rc[1], birth_offset_rel=birth_offset_rel,
                  n_of_points=n_of_points, n_col=n_col, inc_step=inc_step,
                  precision=precision, tree_min_n=tree_min_n, theta=theta, diff_cut=diff_cut,
                  order=order, flags=flags)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def points_relax(g, uenv, venv, u, v, rel_t, dt):
    """Inertial points' velocity relaxed towards the flow (host arithmetic);
    returns new (u, v)."""
    g, uenv, venv = _f64(g), _f64(uenv), _f64(venv)
    u, v = _f64(u).copy(), _f64(v).copy()
    check(lib().vx_points_relax(ptr(g), ptr(uenv), ptr(venv), ptr(u), ptr(v), g.size, float(rel_t), float(dt)))
    return u, v


class Case:
    def __init__(self, params: Params, panels: np.ndarray):
        self.panels = np.ascontiguousarray(panels)
        assert self.panels.dtype.itemsize == 120
        self.params = params
        self.P = int(self.panels.size)
        h = C.c_void_p()
        check(lib().vx_case_create(C.byref(params), ptr(self.panels), self.P, C.byref(h)))
        self._h = h

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib().vx_case_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- matrix -----------------------------------------------------------
    def build_matrix(self):
        check(lib().vx_case_build_matrix(self._h))

    def set_matrix(self, M):
        M = _f64(M)
        if M.ndim != 2 or M.shape[0] != M.shape[1]:
            raise ValueError("matrix must be square")
        check(lib().vx_case_set_matrix(self._h, ptr(M), M.shape[0]))

    def matrix(self):
        M = np.zeros((self.P + 1, self.P + 1))
        check(lib().vx_case_get_matrix(self._h, ptr(M), self.P + 1))
        return M

    def panel_matrix(self):
        L = np.zeros((self.P + 1, self.P + 1))
        check(lib().vx_case_get_panel_matrix(self._h, ptr(L), self.P + 1))
        return L

    # -- state ------------------------------------------------------------
    @property
    def n(self) -> int:
        return int(lib().vx_case_n(self._h))

    @n.setter
    def n(self, value: int):
        check(lib().vx_case_set_n(self._h, int(value)))

    @property
    def capacity(self) -> int:
        return int(lib().vx_case_capacity(self._h))

    @property
    def d_g(self) -> float:
        return float(lib().vx_case_d_g(self._h))

    @d_g.setter
    def d_g(self, v: float):
        check(lib().vx_case_set_d_g(self._h, float(v)))

    def reserve(self, cap: int):
        check(lib().vx_case_reserve(self._h, int(cap)))

    def set_state(self, x, y, g):
        x, y, g = _f64(x), _f64(y), _f64(g)
        check(lib().vx_case_set_state(self._h, ptr(x), ptr(y), ptr(g), x.size))

    def state(self, count=None):
        count = self.n if count is None else count
        x, y, g = np.zeros(count), np.zeros(count), np.zeros(count)
        check(lib().vx_case_get_state(self._h, ptr(x), ptr(y), ptr(g), count))
        return x, y, g

    def velocity(self, count=None):
        count = self.n if count is None else count
        u, v, d = np.zeros(count), np.zeros(count), np.zeros(count)
        check(lib().vx_case_get_velocity(self._h, ptr(u), ptr(v), ptr(d), count))
        return u, v, d

    def set_velocity(self, u, v, d=None):
        u, v = _f64(u), _f64(v)
        dd = None if d is None else _f64(d)
        check(lib().vx_case_set_velocity(self._h, ptr(u), ptr(v), ptr(dd), u.size))

    def sync(self):
        check(lib().vx_case_sync(self._h))

    def save(self, path, user=b""):
        """Exact checkpoint of the case state plus an opaque blob (vx_case_save)."""
        buf = C.create_string_buffer(bytes(user), len(user)) if user else None
        check(lib().vx_case_save(self._h, str(path).encode(), buf, len(user)))

    def load(self, path, user_cap=1 << 20) -> bytes:
        """Restore a checkpoint written by save(); returns its blob."""
        buf = C.create_string_buffer(max(1, user_cap))
        got = C.c_int64()
        check(lib().vx_case_load(self._h, str(path).encode(), buf, user_cap, C.byref(got)))
        return buf.raw[:got.value]

    # -- stages -----------------------------------------------------------
    def vort_creation(self, vinf):
        check(lib().vx_vort_creation(self._h, vinf[0], vinf[1]))

    def speed(self, vinf):
        check(lib().vx_speed(self._h, vinf[0], vinf[1]))

    def step(self):
        out = np.zeros(4)
        check(lib().vx_step(self._h, ptr(out)))
        return out

    def time_step(self, vinf):
        out = np.zeros(4)
        check(lib().vx_time_step(self._h, vinf[0], vinf[1], ptr(out)))
        return out

    def run_steps(self, k, vinf):
        out = np.zeros(4 * max(k, 1))
        check(lib().vx_run_steps(self._h, int(k), vinf[0], vinf[1], ptr(out)))
        return out[:4 * k].reshape(k, 4)

    def autotune(self, vinf=(1.0, 0.0)) -> dict:
        """Time the launch-shape choices on the current state and keep the
        fastest (vx_case_autotune)."""
        t = Tuning()
        check(lib().vx_case_autotune(self._h, float(vinf[0]), float(vinf[1]), C.byref(t)))
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def set_tuning(self, tpt_bs=0, tpt_diff=0, leaf_off=-1, col_share_min=0):
        t = Tuning(tpt_bs=tpt_bs, tpt_diff=tpt_diff, leaf_off=leaf_off, col_share_min=col_share_min)
        check(lib().vx_case_set_tuning(self._h, C.byref(t)))

    def tuning(self) -> dict:
        t = Tuning()
        check(lib().vx_case_get_tuning(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def run_steps_graph(self, k, vinf):
        """k time steps as replays of one captured HIP graph (direct path);
        returns the forces (k x 4) and the count after each step."""
        f4 = np.zeros((k, 4))
        counts = np.zeros(k, dtype=np.int64)
        check(lib().vx_run_steps_graph(self._h, int(k), float(vinf[0]), float(vinf[1]), ptr(f4), ptr(counts)))
        return f4, counts

    def velocity_probe(self, px, py, vinf):
        px, py = _f64(px), _f64(py)
        u, v = np.zeros(px.size), np.zeros(px.size)
        check(lib().vx_velocity_probe(self._h, ptr(px), ptr(py), px.size, vinf[0], vinf[1],
                                      ptr(u), ptr(v)))
        return u, v

    def points_move(self, x, y, g, u, v):
        """Move tracer points over dt (absorbed: g = 0; out of the box: deleted).
        Returns new (x, y, g)."""
        x, y, g = _f64(x).copy(), _f64(y).copy(), _f64(g).copy()
        u, v = _f64(u), _f64(v)
        check(lib().vx_points_move(self._h, ptr(x), ptr(y), ptr(g), ptr(u), ptr(v), x.size))
        return x, y, g

    def right_part(self, n_src, vinf):
        R = np.zeros(self.P)
        check(lib().vx_right_part(self._h, int(n_src), vinf[0], vinf[1], ptr(R)))
        return R

    def birth(self, R_p, d_g):
        R_p = _f64(R_p)
        gs = np.zeros(self.P)
        check(lib().vx_birth(self._h, ptr(R_p), float(d_g), ptr(gs)))
        return gs

    def biot_savart(self, n_tgt, n_src, vinf=(0.0, 0.0)):
        check(lib().vx_biot_savart(self._h, int(n_tgt), int(n_src), vinf[0], vinf[1]))

    def diffusion(self, n_tgt, n_src):
        check(lib().vx_diffusion(self._h, int(n_tgt), int(n_src)))

    def tree_biot_savart(self, s, vinf=(0.0, 0.0)):
        check(lib().vx_tree_biot_savart(self._h, int(s), vinf[0], vinf[1]))

    def tree_diffusion(self):
        check(lib().vx_tree_diffusion(self._h))

    def tree_census(self, on=True):
        """Count the diffusion pairs within each target's radius from the next
        tree diffusion on (tree_stats()["diff_in_radius"])."""
        check(lib().vx_tree_census(self._h, int(bool(on))))

    def tree_stats(self) -> dict:
        from ._lib import TreeInfo
        t = TreeInfo()
        check(lib().vx_tree_stats(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in TreeInfo._fields_}

    def solve_info(self) -> dict:
        """Refinement corrections, convergence and relative residual of the
        last panel-system solve (VX_FLAG_LU_REFINE cases)."""
        it, ok, r = C.c_int32(), C.c_int32(), C.c_double()
        check(lib().vx_case_solve_info(self._h, C.byref(it), C.byref(ok), C.byref(r)))
        return {"iters": it.value, "converged": bool(ok.value), "resid_rel": r.value}

    def set_probes(self, px, py):
        """Probe points kept on the device; time_step computes their velocities."""
        px, py = _f64(px), _f64(py)
        check(lib().vx_case_set_probes(self._h, ptr(px), ptr(py), px.size))
        self._nprobe = px.size

    def probes(self):
        m = getattr(self, "_nprobe", 0)
        u, v = np.zeros(m), np.zeros(m)
        check(lib().vx_case_get_probes(self._h, ptr(u), ptr(v), m))
        return u, v

    def tree_order(self, count):
        """Sorted Morton keys and slot permutation of the last tree build."""
        keys = np.zeros(count, np.uint32)
        perm = np.zeros(count, np.int32)
        check(lib().vx_tree_order(self._h, ptr(keys), ptr(perm), count))
        return keys, perm

    def diffusion_wall(self, n_tgt):
        check(lib().vx_diffusion_wall(self._h, int(n_tgt)))

    def step_move(self):
        out = np.zeros(4)
        check(lib().vx_step_move(self._h, ptr(out)))
        return out

    def compact(self):
        check(lib().vx_compact(self._h))

    def setka(self, kind):
        col = np.zeros(max(self.n, 1), dtype=np.int32)
        check(lib().vx_setka(self._h, int(kind), ptr(col)))
        return col[:self.n]

    def collapse(self, kind, col=None):
        c = None if col is None else np.ascontiguousarray(col, dtype=np.int32)
        check(lib().vx_collapse(self._h, int(kind), ptr(c)))

    def enable_timing(self, on=True):
        check(lib().vx_case_enable_timing(self._h, 1 if on else 0))

    def timings(self) -> dict:
        t = Timings()
        check(lib().vx_case_timings(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Timings._fields_}


class Batch:
    """Independent cases advanced together by one graph launch per step
    (vx_batch_*).  The cases stay owned by the caller."""

    def __init__(self, cases):
        self.cases = list(cases)
        arr = (C.c_void_p * len(self.cases))(*[c._h.value for c in self.cases])
        h = C.c_void_p()
        check(lib().vx_batch_create(arr, len(self.cases), C.byref(h)))
        self._h = h

    def run_steps(self, k, vinf):
        """vinf: one (vx, vy) per case; returns forces (cases x k x 4) and counts (cases x k)."""
        v = np.ascontiguousarray(np.asarray(vinf, dtype=np.float64).reshape(len(self.cases), 2))
        f4 = np.zeros((len(self.cases), k, 4))
        counts = np.zeros((len(self.cases), k), dtype=np.int64)
        check(lib().vx_batch_run_steps(self._h, int(k), ptr(v), ptr(f4), ptr(counts)))
        return f4, counts

    def close(self):
        if getattr(self, "_h", None):
            lib().vx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
