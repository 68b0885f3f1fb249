"""Launch-shape autotuning (vx_case_autotune).

Direct path: every number of targets per thread sums each target's sources
in the same order, so velocities, d and diffusion are bit-identical before
and after tuning.  Tree path: the chosen bucket depth keeps the tree's error
bound (3e-4 of the largest velocity in mixed precision against the FP64
direct oracle on sampled targets)."""
import math

import numpy as np
import pytest

from vortex_amd import geometry as geo
from vortex_amd.solver import Case, make_params, VX_PREC_F64, VX_PREC_MIXED
from vortex_amd.state import uniform_wake

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [VX_PREC_F64, VX_PREC_MIXED])
def test_autotune_direct_path_keeps_results(prec):
    pan = geo.circle(200, 0.5)
    n = 20000
    x, y, g = uniform_wake(n, seed=5)
    c = Case(make_params(pan, precision=prec, ve_size=0.004), pan)
    c.set_state(x, y, g)
    c.biot_savart(n, n, (1.0, 0.0))
    c.diffusion(n, n)
    ref = c.velocity(n)
    t = c.autotune((1.0, 0.0))
    print(f"[tune] direct prec={prec}: {t}", flush=True)
    assert t["tpt_bs"] in (1, 2, 4, 8) and t["tpt_diff"] in (1, 2, 4, 8) and t["trials"] == 8
    assert 0 < t["ms_best"] <= t["ms_default"] * 1.0001
    c.biot_savart(n, n, (1.0, 0.0))
    c.diffusion(n, n)
    got = c.velocity(n)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    c.close()


def test_autotune_tree_depth_keeps_the_error_bound(ref):
    n = 1 << 18
    x, y, g = uniform_wake(n, seed=2)
    eps = math.pi / 1000
    pan = geo.circle(1000, 0.5)
    c = Case(make_params(pan, precision=VX_PREC_MIXED, ve_size=eps, tree_min_n=1), pan)
    c.set_state(x, y, g)
    t = c.autotune((0.0, 0.0))
    print(f"[tune] tree: {t}", flush=True)
    assert 0 <= t["leaf_off"] <= 3 and t["trials"] == 4
    c.tree_biot_savart(n, (0.0, 0.0))
    u, v, d = c.velocity(n)
    c.close()
    sample = 512
    ur, vr, dr = ref.biot_savart(x, y, g, eps * eps, (0.0, 0.0), n_tgt=sample)
    scale = max(np.max(np.abs(ur)), np.max(np.abs(vr)))
    assert np.max(np.abs(u[:sample] - ur)) < 3e-4 * scale and np.max(np.abs(v[:sample] - vr)) < 3e-4 * scale
    assert np.max(np.abs(d[:sample] - dr) / dr) < 1e-4
