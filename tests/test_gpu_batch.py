"""Batched cases (vx_batch_*) against each case run alone.

A batch in the shape of configuration C3: flat plates with panel counts drawn
from {200, 1000, 2040, 3320}, angle of attack uniform in [-10, 20] degrees,
Re log-uniform in [500, 5000], dt = 0.005 (draws from numpy's PCG64, seed 3).
Each case's forces, counts and final state after the batched steps equal its
own stream steps (vx_time_step) bit for bit, in FP64 and in mixed precision,
over two calls (the second replays the kept graph)."""
import numpy as np
import pytest

from vortex_amd import geometry as geo
from vortex_amd.solver import Batch, Case, make_params, VX_PREC_F64, VX_PREC_MIXED

pytestmark = pytest.mark.gpu


def c3_cases(k, prec, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(k):
        P = int(rng.choice([200, 1000, 2040, 3320]))
        alpha = float(rng.uniform(-10.0, 20.0))
        re = float(np.exp(rng.uniform(np.log(500.0), np.log(5000.0))))
        pan = geo.flat_plate(P, chord=1.0, thickness=0.02, alpha_deg=alpha)
        prm = make_params(pan, dt=0.005, nu=1.0 / re, r_col_diff_sign=1.0, r_col_same_sign=0.3, precision=prec)
        out.append((pan, prm))
    return out


@pytest.mark.parametrize("prec", [VX_PREC_F64, VX_PREC_MIXED])
def test_batch_equals_single_cases(prec):
    spec = c3_cases(6, prec)
    batch_cases, solo = [], []
    for pan, prm in spec:
        for lst in (batch_cases, solo):
            c = Case(prm, pan)
            c.build_matrix()
            lst.append(c)
    b = Batch(batch_cases)
    vinf = [(1.0, 0.0)] * len(spec)
    for call, k in enumerate((3, 4)):
        print(f"[batch] prec={prec} call={call} steps={k}", flush=True)
        fb, nb = b.run_steps(k, vinf)
        for i, c in enumerate(solo):
            fs = np.array([c.time_step((1.0, 0.0)) for _ in range(k)])
            assert np.array_equal(fs, fb[i]), (call, i, np.max(np.abs(fs - fb[i])))
            assert batch_cases[i].n == c.n == nb[i, -1]
            xa, ya, ga = c.state()
            xb, yb, gb = batch_cases[i].state()
            assert np.array_equal(xa, xb) and np.array_equal(ya, yb) and np.array_equal(ga, gb)
    b.close()
    for c in batch_cases + solo:
        c.close()
