"""Graph-replayed time steps (vx_run_steps_graph) against stream steps
(vx_run_steps) on the same start: forces, counts and the whole state bit for
bit, in both precisions, over two calls (the second replays the kept graph),
with the capacity growing in between (the graph is captured again)."""
import numpy as np
import pytest

from vortex_amd import geometry as geo
from vortex_amd.solver import Case, make_params, VX_FLAG_LU_REFINE, VX_PREC_F64, VX_PREC_MIXED

pytestmark = pytest.mark.gpu
VINF = (1.0, 0.0)


def c0_case(prec, flags=0):
    pan = geo.circle(200, 0.5)
    prm = make_params(pan, dt=0.01, nu=1e-3, r_col_diff_sign=1.0, r_col_same_sign=0.3, max_ve_g=0.00628162,
                      n_col=2, birth_offset_rel=0.5, precision=prec, flags=flags, inc_step=1024)
    c = Case(prm, pan)
    c.build_matrix()
    return c


@pytest.mark.parametrize("prec,flags", [(VX_PREC_F64, 0), (VX_PREC_MIXED, 0), (VX_PREC_MIXED, VX_FLAG_LU_REFINE)])
def test_graph_steps_equal_stream_steps(prec, flags):
    a, b = c0_case(prec, flags), c0_case(prec, flags)
    for call, k in enumerate((7, 12)):
        print(f"[graph] prec={prec} flags={flags} call={call} steps={k}", flush=True)
        fa = np.array([a.time_step(VINF) for _ in range(k)])
        fb, counts = b.run_steps_graph(k, VINF)
        assert a.n == b.n, (call, a.n, b.n)
        assert counts[-1] == b.n
        assert np.array_equal(fa, fb), (call, np.max(np.abs(fa - fb)))
        xa, ya, ga = a.state()
        xb, yb, gb = b.state()
        assert np.array_equal(xa, xb) and np.array_equal(ya, yb) and np.array_equal(ga, gb)
        assert b.d_g == a.d_g
    a.close()
    b.close()


def test_graph_steps_refuse_the_tree_path():
    pan = geo.circle(200, 0.5)
    c = Case(make_params(pan, tree_min_n=1000), pan)
    c.build_matrix()
    with pytest.raises(Exception):
        c.run_steps_graph(10, VINF)  # 10 x 200 births pass tree_min_n
    c.close()
