"""ctypes binding of the FP64 CPU oracle (oracle/build/libvx_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path never imports
this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libvx_ref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_i64 = C.c_int64
_d = C.c_double


class Params(C.Structure):
    """Mirror of vxr_params in vx_ref.h."""
    _fields_ = [(n, C.c_double) for n in (
        "dt", "ve_size", "nu", "rho", "vinf_x", "vinf_y",
        "x_min", "x_max", "y_min", "y_max",
        "px_min", "px_max", "py_min", "py_max",
        "r_col_diff_sign", "r_col_same_sign", "max_ve_g", "rc_x", "rc_y",
        "birth_offset_rel")] + [("n_of_points", C.c_int64), ("n_col", C.c_int64)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None
_lib_path = LIB_PATH


def host_has_avx512() -> bool:
    try:
        flags = open("/proc/cpuinfo").read()
    except OSError:
        return False
    return all(f in flags for f in ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"))


def select_build(fast: bool) -> str:
    """Load the x86-64-v4 build of the same source (bench.py's CPU baseline)
    when fast and the host supports AVX-512; returns the path in use."""
    global _lib, _lib_path
    want = LIB_PATH
    if fast and host_has_avx512():
        v4 = os.path.join(os.path.dirname(LIB_PATH), "libvx_ref_v4.so")
        if os.path.exists(v4):
            want = v4
    if want != _lib_path:
        _lib, _lib_path = None, want
    lib()
    return _lib_path


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_lib_path):
            build()
        L = C.CDLL(_lib_path)
        L.vxr_panel_matrix.argtypes = [_vp, _i64, _dp]
        L.vxr_invert.argtypes = [_dp, _i64, _dp]
        L.vxr_invert.restype = C.c_int
        L.vxr_invert_mode.argtypes = [_dp, _i64, _dp, C.c_int]
        L.vxr_invert_mode.restype = C.c_int
        L.vxr_right_part.argtypes = [_dp, _dp, _dp, _i64, _vp, _i64, _d, _d, _dp]
        L.vxr_birth.argtypes = [_dp, _i64, _dp, _d, _vp, _d, _dp, _dp, _dp, _vp]
        L.vxr_biot_savart.argtypes = [_dp, _dp, _dp, _i64, _i64, _d, _d, _d, _dp, _dp, _vp]
        L.vxr_velocity_probe.argtypes = [_dp, _dp, _dp, _i64, _dp, _dp, _i64, _d, _d, _d, _dp, _dp]
        L.vxr_diffusion.argtypes = [_dp, _dp, _dp, _dp, _i64, _i64, _d, _d, _dp, _dp]
        L.vxr_diffusion_wall.argtypes = [_dp, _dp, _dp, _i64, _vp, _i64, _d, _i64, _dp, _dp]
        L.vxr_step_move.argtypes = [_dp, _dp, _dp, _dp, _dp, _i64, _i64, _vp, C.POINTER(Params),
                                    _vp, _vp, _vp, _vp, _dp]
        L.vxr_hitting.argtypes = [_vp, _i64, C.POINTER(Params), _d, _d, _d, _d, C.POINTER(C.c_int)]
        L.vxr_hitting.restype = C.c_int
        L.vxr_points_move.argtypes = [_dp, _dp, _dp, _dp, _dp, _i64, _i64, _vp, C.POINTER(Params)]
        L.vxr_points_relax.argtypes = [_dp, _dp, _dp, _dp, _dp, _i64, _d, _d]
        L.vxr_compact.argtypes = [_dp, _dp, _dp, _i64]
        L.vxr_compact.restype = _i64
        L.vxr_setka.argtypes = [_dp, _dp, _dp, _i64, C.c_int, _d, _d, _ip]
        L.vxr_collapse.argtypes = [_dp, _dp, _dp, _ip, _i64, C.c_int]
        L.vxr_Step.argtypes = [_dp, _dp, _dp, _dp, _dp, C.POINTER(_i64), _vp, _i64,
                               C.POINTER(Params), _dp]
        L.vxr_Speed.argtypes = [_dp, _dp, _dp, _i64, _vp, _i64, C.POINTER(Params), _d, _d,
                                _dp, _dp, _dp]
        L.vxr_vort_creation.argtypes = [_dp, _dp, _dp, _i64, _dp, _vp, _i64, C.POINTER(Params),
                                        _d, _d, _d]
        L.vxr_time_step.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, C.POINTER(_i64), _dp, _vp, _i64,
                                    C.POINTER(Params), _d, _d, C.POINTER(C.c_double), _dp]
        L.vxr_num_threads.restype = C.c_int
        L.vxr_set_num_threads.argtypes = [C.c_int]
        L.vxr_sizeof_panel.restype = C.c_size_t
        _u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
        _i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.vxr_spread16.argtypes = [C.c_uint32]
        L.vxr_spread16.restype = C.c_uint32
        L.vxr_morton_keys.argtypes = [_dp, _dp, _dp, _i64, _dp, _u32p, _dp]
        L.vxr_morton_keys.restype = _i64
        L.vxr_stable_order.argtypes = [_u32p, _i64, _i32p]
        _lib = L
    return _lib


def _pp(pan: np.ndarray):
    assert pan.dtype.itemsize == 120 and pan.flags["C_CONTIGUOUS"]
    return pan.ctypes.data


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def panel_matrix(pan):
    P = pan.size
    L = np.zeros((P + 1) * (P + 1))
    lib().vxr_panel_matrix(_pp(pan), P, L)
    return L.reshape(P + 1, P + 1)


def invert(A, mode=0):
    """mode 0: reference pivot rule; 1: standard partial pivoting;
    -1: reference rule, falling back to 1 where it reports a tiny pivot."""
    A0 = _f(A)
    m = A0.shape[0]
    Minv = np.zeros((m, m))
    if mode == -1:
        A = A0.copy()
        if lib().vxr_invert_mode(A.reshape(-1), m, Minv.reshape(-1), 0) == 0:
            return Minv
        mode = 1
    A = A0.copy()
    rc = lib().vxr_invert_mode(A.reshape(-1), m, Minv.reshape(-1), mode)
    if rc:
        raise np.linalg.LinAlgError("singular panel matrix (pivot below 1e-12)")
    return Minv


def right_part(x, y, g, pan, vinf):
    R = np.zeros(pan.size)
    lib().vxr_right_part(_f(x), _f(y), _f(g), len(x), _pp(pan), pan.size, vinf[0], vinf[1], R)
    return R


def birth(M, R_p, d_g, pan, birth_offset_rel=0.0):
    P = pan.size
    xo, yo, go = np.zeros(P), np.zeros(P), np.zeros(P)
    gs = np.zeros(P)
    lib().vxr_birth(_f(M).reshape(-1), P, _f(R_p), d_g, _pp(pan), birth_offset_rel, xo, yo, go,
                    gs.ctypes.data)
    return xo, yo, go, gs


def biot_savart(x, y, g, eps2, vinf=(0.0, 0.0), n_tgt=None):
    n = len(x)
    nt = n if n_tgt is None else n_tgt
    u, v, d = np.zeros(nt), np.zeros(nt), np.zeros(nt)
    lib().vxr_biot_savart(_f(x), _f(y), _f(g), nt, n, eps2, vinf[0], vinf[1], u, v, d.ctypes.data)
    return u, v, d


def velocity_probe(x, y, g, px, py, eps2, vinf=(0.0, 0.0)):
    m = len(px)
    u, v = np.zeros(m), np.zeros(m)
    lib().vxr_velocity_probe(_f(x), _f(y), _f(g), len(x), _f(px), _f(py), m, eps2,
                             vinf[0], vinf[1], u, v)
    return u, v


def diffusion(x, y, g, d, nu, ve_size, u=None, v=None, n_tgt=None):
    n = len(x)
    nt = n if n_tgt is None else n_tgt
    u = np.zeros(nt) if u is None else _f(u).copy()
    v = np.zeros(nt) if v is None else _f(v).copy()
    lib().vxr_diffusion(_f(x), _f(y), _f(g), _f(d), nt, n, nu, ve_size, u, v)
    return u, v


def diffusion_wall(x, y, d, pan, nu, n_of_points, u=None, v=None):
    n = len(x)
    u = np.zeros(n) if u is None else _f(u).copy()
    v = np.zeros(n) if v is None else _f(v).copy()
    lib().vxr_diffusion_wall(_f(x), _f(y), _f(d), n, _pp(pan), pan.size, nu, n_of_points, u, v)
    return u, v


def step_move(x, y, g, u, v, n, pan, prm: Params):
    x, y, g = _f(x).copy(), _f(y).copy(), _f(g).copy()
    dg, fx, fy, m = (np.zeros(max(n, 1)) for _ in range(4))
    out = np.zeros(4)
    lib().vxr_step_move(x, y, g, _f(u), _f(v), n, pan.size, _pp(pan), C.byref(prm),
                        dg.ctypes.data, fx.ctypes.data, fy.ctypes.data, m.ctypes.data, out)
    return x, y, g, out, (dg[:n], fx[:n], fy[:n], m[:n])


def hitting(pan, prm: Params, new, old):
    hp = C.c_int(-7)
    r = lib().vxr_hitting(_pp(pan), pan.size, C.byref(prm), new[0], new[1], old[0], old[1],
                          C.byref(hp))
    return bool(r), hp.value


def points_move(x, y, g, u, v, pan, prm: Params):
    x, y, g = _f(x).copy(), _f(y).copy(), _f(g).copy()
    lib().vxr_points_move(x, y, g, _f(u), _f(v), x.size, pan.size, _pp(pan), C.byref(prm))
    return x, y, g


def points_relax(g, uenv, venv, u, v, rel_t, dt):
    u, v = _f(u).copy(), _f(v).copy()
    lib().vxr_points_relax(_f(g), _f(uenv), _f(venv), u, v, u.size, rel_t, dt)
    return u, v


def compact(x, y, g, n=None):
    x, y, g = _f(x).copy(), _f(y).copy(), _f(g).copy()
    n = len(x) if n is None else n
    nn = lib().vxr_compact(x, y, g, n)
    return x, y, g, int(nn)


def setka(x, y, g, kind, r2, max_ve_g, n=None):
    n = len(x) if n is None else n
    COL = np.zeros(max(n, 1), dtype=np.int32)
    lib().vxr_setka(_f(x), _f(y), _f(g), n, kind, r2, max_ve_g, COL)
    return COL[:n]


def collapse(x, y, g, COL, kind):
    x, y, g = _f(x).copy(), _f(y).copy(), _f(g).copy()
    COL = np.ascontiguousarray(COL, dtype=np.int32).copy()
    lib().vxr_collapse(x, y, g, COL, len(COL), kind)
    return x, y, g


def Step(x, y, g, u, v, n, pan, prm: Params):
    x, y, g = _f(x).copy(), _f(y).copy(), _f(g).copy()
    nn = _i64(n)
    out = np.zeros(4)
    lib().vxr_Step(x, y, g, _f(u), _f(v), C.byref(nn), _pp(pan), pan.size, C.byref(prm), out)
    return x, y, g, int(nn.value), out


def Speed(x, y, g, s, pan, prm: Params, vinf):
    u, v, d = np.zeros(s), np.zeros(s), np.zeros(s)
    lib().vxr_Speed(_f(x), _f(y), _f(g), s, _pp(pan), pan.size, C.byref(prm), vinf[0], vinf[1],
                    u, v, d)
    return u, v, d


def vort_creation(x, y, g, n, M, pan, prm: Params, vinf, d_g):
    x, y, g = _f(x).copy(), _f(y).copy(), _f(g).copy()
    lib().vxr_vort_creation(x, y, g, n, _f(M).reshape(-1), _pp(pan), pan.size, C.byref(prm),
                            vinf[0], vinf[1], d_g)
    return x, y, g


def morton_keys(x, y, g, box):
    """Tree keys of slots [0, len(x)): (keys uint32, included count, key box)."""
    x, y, g = _f(x), _f(y), _f(g)
    keys = np.zeros(x.size, np.uint32)
    ob = np.zeros(4)
    m = lib().vxr_morton_keys(x, y, g, x.size, _f(box), keys, ob)
    return keys, int(m), ob


def stable_order(keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    perm = np.zeros(keys.size, np.int32)
    lib().vxr_stable_order(keys, keys.size, perm)
    return perm


class State:
    """Host state of one case for the oracle time loop (capacity-managed)."""

    def __init__(self, pan, prm: Params, M, capacity=8192, tail_seed=12345):
        self.pan, self.prm, self.M = pan, prm, _f(M).reshape(-1)
        self.P = pan.size
        self.n = 0
        self.d_g = 0.0
        self.x = np.zeros(0); self.y = np.zeros(0); self.g = np.zeros(0)
        self._grow(capacity, tail_seed)

    def _grow(self, cap, seed=12345):
        old = self.x.size
        from vortex_amd.state import far_tail
        tx, ty = far_tail(old, cap - old, seed)
        self.x = np.concatenate([self.x, tx])
        self.y = np.concatenate([self.y, ty])
        self.g = np.concatenate([self.g, np.zeros(cap - old)])
        self.u = np.zeros(cap); self.v = np.zeros(cap); self.d = np.zeros(cap)

    def step(self, vinf):
        need = (self.n + self.P + 31) // 32 * 32
        if need > self.x.size:
            inc = 8192
            cap = self.x.size
            while cap < need:
                cap += inc
            self._grow(cap)
        n = _i64(self.n)
        dg = C.c_double(self.d_g)
        f4 = np.zeros(4)
        lib().vxr_time_step(self.x, self.y, self.g, self.u, self.v, self.d, C.byref(n), self.M,
                            _pp(self.pan), self.P, C.byref(self.prm), vinf[0], vinf[1],
                            C.byref(dg), f4)
        self.n = int(n.value)
        self.d_g = dg.value
        return f4
