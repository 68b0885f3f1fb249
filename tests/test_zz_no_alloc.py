"""vx_velocity_probe_noalloc (caller-owned device buffers, partials in the
library's file-scope device scratch) against vx_velocity_probe on the same
state: two calls in a row, the second with fewer probes, both precisions.
The first call's probe count makes the partials exceed the scratch at the
default split, so it also takes the capped chunking (the FP64 sums then
group differently: agreement to 1e-12 of the largest magnitude, the bar for
FP64 sums).  The device buffers come from the HIP runtime the library itself
uses (hipMalloc through ctypes); the case stream is synchronised with
vx_case_sync before the outputs are copied back."""
import ctypes as C

import numpy as np
import pytest

from vortex_amd import geometry as geo
from vortex_amd._lib import check, lib
from vortex_amd.solver import Case, make_params, VX_PREC_F64, VX_PREC_MIXED

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip():
    lib()  # the library loads the runtime it was linked against
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DevArray:
    def __init__(self, h, n):
        self.h, self.n, self.p = h, n, C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), 8 * n) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n and self.h.hipMemcpy(self.p, a.ctypes.data, 8 * self.n, H2D) == 0

    def get(self):
        a = np.empty(self.n)
        assert self.h.hipMemcpy(a.ctypes.data, self.p, 8 * self.n, D2H) == 0
        return a

    def free(self):
        if self.p:
            self.h.hipFree(self.p)
            self.p = C.c_void_p()


def state(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.5, 6.0, n)
    y = rng.uniform(-2.0, 2.0, n)
    keep = x * x + y * y > 0.3
    x, y = x[keep], y[keep]
    g = rng.uniform(-0.004, 0.004, x.size)
    return x, y, g


@pytest.mark.parametrize("prec", [VX_PREC_F64, VX_PREC_MIXED])
def test_noalloc_probe_matches_probe(prec):
    h = hip()
    pan = geo.circle(200, 0.5)
    c = Case(make_params(pan, precision=prec, ve_size=0.004), pan)
    x, y, g = state(30000, 3)
    c.set_state(x, y, g)
    rng = np.random.default_rng(4)
    vinf = (1.0, 0.25)
    for m in (40000, 700):
        print(f"[noalloc] prec={prec} m={m}", flush=True)
        px = rng.uniform(-1.0, 7.0, m)
        py = rng.uniform(-2.5, 2.5, m)
        ur, vr = c.velocity_probe(px, py, vinf)
        bufs = [DevArray(h, m) for _ in range(4)]
        dpx, dpy, du, dv = bufs
        dpx.put(px)
        dpy.put(py)
        du.put(np.full(m, np.nan))
        dv.put(np.full(m, np.nan))
        check(lib().vx_velocity_probe_noalloc(c._h, dpx.p, dpy.p, m, vinf[0], vinf[1], du.p, dv.p))
        check(lib().vx_case_sync(c._h))
        u, v = du.get(), dv.get()
        for b in bufs:
            b.free()
        scale = max(np.max(np.abs(ur)), np.max(np.abs(vr)))
        assert np.max(np.abs(u - ur)) <= 1e-12 * scale and np.max(np.abs(v - vr)) <= 1e-12 * scale
    # more probes than the scratch holds: an argument error before any launch
    one = DevArray(h, 1)
    assert lib().vx_velocity_probe_noalloc(c._h, one.p, one.p, (1 << 20) + 1, 0.0, 0.0, one.p, one.p) == 2
    one.free()
    c.close()
