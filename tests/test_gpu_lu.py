"""FP32 LU with FP64-residual refinement for the panel system
(VX_FLAG_LU_REFINE) against a direct FP64 solve (numpy.linalg.solve on the
oracle's panel matrix) and against the FP64 Gauss-Jordan inverse path.

Bar: the solve reports convergence within the six corrections it queues, its
relative residual ||r|| / (||L|| ||x||) is at most eps64 sqrt(n) (LAPACK
dsgesv's test: an FP64 backward error), the sheet strengths agree with
numpy's FP64 solve to 1e-10 of their largest magnitude (the forward error of
two backward-stable FP64 solves of a thin plate's system), and whole time
steps equal the inverse path's to the FP64 tolerances of the step tests.
"""
import numpy as np
import pytest

from vortex_amd import geometry as geo
from vortex_amd.solver import Case, make_params, VX_FLAG_LU_REFINE, VX_PREC_F64, VX_PREC_MIXED

pytestmark = pytest.mark.gpu
VINF = (1.0, 0.0)


def profiles():
    return {
        "circle200": geo.circle(200, 0.5),
        "naca2576": geo.naca00(2576, t=0.12, chord=1.0, alpha_deg=8.0),
        "plate5200": geo.flat_plate(5200, chord=1.0, thickness=0.02, alpha_deg=15.0),
    }


@pytest.mark.parametrize("name", ["circle200", "naca2576", "plate5200"])
def test_lu_refined_sheet_matches_fp64_solve(ref, name):
    pan = profiles()[name]
    P = pan.size
    print(f"[lu] {name}: P={P}", flush=True)
    rng = np.random.default_rng(P)
    # a right-hand side with vortices near the body (not just the free stream)
    nv = 3000
    th = rng.uniform(0, 2 * np.pi, nv)
    x = 0.5 + 0.8 * np.cos(th) * rng.uniform(0.7, 1.5, nv)
    y = 0.8 * np.sin(th) * rng.uniform(0.7, 1.5, nv)
    g = rng.uniform(-0.01, 0.01, nv)
    Lr = ref.panel_matrix(pan)
    R = ref.right_part(x, y, g, pan, VINF)
    d_g = 0.0123
    sol = np.linalg.solve(Lr, np.concatenate([R, [d_g]]))[:P]
    c = Case(make_params(pan, precision=VX_PREC_MIXED, flags=VX_FLAG_LU_REFINE), pan)
    c.build_matrix()
    gs = c.birth(R, d_g)
    info = c.solve_info()
    c.close()
    print(f"[lu] {name}: {info}", flush=True)
    assert info["converged"] and 1 <= info["iters"] <= 6
    # backward error at FP64 level (the refinement's own test), and the
    # forward error against numpy's FP64 solve within the conditioning
    assert info["resid_rel"] <= 1.1e-16 * np.sqrt(P + 1)
    assert np.max(np.abs(gs - sol)) < 1e-10 * np.max(np.abs(sol))


@pytest.mark.parametrize("prec", [VX_PREC_F64, VX_PREC_MIXED])
def test_lu_time_steps_match_inverse_path(prec):
    pan = geo.naca00(2576, t=0.12, chord=1.0, alpha_deg=8.0)
    out = []
    for flags in (0, VX_FLAG_LU_REFINE):
        c = Case(make_params(pan, dt=0.001, nu=1e-4, precision=prec, tree_min_n=1, flags=flags), pan)
        c.build_matrix()
        fs = [c.time_step(VINF) for _ in range(3)]
        out.append((c.n, c.state(), fs))
        c.close()
    (n0, (x0, y0, g0), f0), (n1, (x1, y1, g1), f1) = out
    assert n0 == n1
    assert np.max(np.abs(x0 - x1)) < 1e-12 and np.max(np.abs(g0 - g1)) < 1e-12 * np.max(np.abs(g0))
    for a, b in zip(f0, f1):
        assert abs(a[1] - b[1]) < 1e-9 * max(1.0, abs(a[1]))
