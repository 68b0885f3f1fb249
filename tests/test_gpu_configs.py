"""The benchmark configurations on the GPU against the FP64 oracle.

  C0  cylinder, 200 panels, eps = panel length, births 0.5 panel length off
      the surface, dt = 0.01, FP64 direct sums: 30 full steps against the
      oracle's time loop (src/main.c:652-785, src/unit_1.cu:1077-1218), and
      the `vortex` CLI on the same run file against the oracle's force
      history and last snapshot.
  C2  flat plate, chord 1, 15 degrees, 5200 panels clustered to both edges:
      panel matrix and inverse (M L = I), the first birth sheet against
      numpy.linalg.solve, and three full steps (FP64 direct against the
      oracle; mixed precision on the tree).
  C4  symmetric 12 % airfoil (4-digit thickness polynomial), 2576 panels,
      8 degrees: the same checks.
  C1 at 2^20: the FP64 order-16 tree (the plain path's far field) and the
      tree diffusion against the oracle's direct sums on sampled targets.
  Capacity growth across an inc_step boundary, against the oracle's state.
  Step's kind order: a three-vortex case whose outcome depends on the
  opposite-sign passes running before the same-sign ones.

Tolerances (FP64 sums differ from the oracle's serial ones in order, and the
differences grow over steps): positions 1e-10 after 3 steps, 1e-9 after 30,
1e-7 after 60; circulations 1e-12; forces 1e-9 of the run's RMS drag over 30
steps, 1e-6 relative over 60; vortex counts exact.  Mixed precision: counts
within 2, forces 1e-3.  Tree as in test_gpu_tree.py.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from vortex_amd import geometry as geo
from vortex_amd.solver import Case, make_params, VX_PREC_F64, VX_PREC_MIXED
from vortex_amd.state import uniform_wake

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "vortex_amd", "vortex")
VINF = (1.0, 0.0)


def ref_params(prm):
    import vx_ref
    return vx_ref.Params(**{k: getattr(prm, k) for k, _ in vx_ref.Params._fields_})


def rel_err(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


# ---------------------------------------------------------------- C0
C0_KW = dict(dt=0.01, nu=1e-3, r_col_diff_sign=1.0, r_col_same_sign=0.3, max_ve_g=0.00628162, n_col=2,
             birth_offset_rel=0.5, precision=VX_PREC_F64)


def c0_oracle(ref, steps):
    import vx_ref
    pan = geo.circle(200, 0.5)
    prm = make_params(pan, **C0_KW)
    M = ref.invert(ref.panel_matrix(pan), 1)
    st = vx_ref.State(pan, ref_params(prm), M, capacity=8192)
    hist = []
    for _ in range(steps):
        f = st.step(VINF)
        hist.append((st.n, f.copy()))
    return pan, prm, M, st, hist


def test_c0_full_run_against_oracle(ref):
    steps = 30
    pan, prm, M, st, hist = c0_oracle(ref, steps)
    c = Case(prm, pan)
    c.build_matrix()
    assert np.max(np.abs(c.matrix() - M)) < 1e-10 * np.max(np.abs(M))
    c.set_matrix(M)
    fx_rms = math.sqrt(np.mean([f[1] ** 2 for _, f in hist]))
    for k in range(steps):
        f = c.time_step(VINF)
        n_ref, f_ref = hist[k]
        assert c.n == n_ref, (k, c.n, n_ref)
        assert np.all(np.abs(f[1:3] - f_ref[1:3]) < 1e-9 * fx_rms), (k, f, f_ref)
        assert abs(f[0] - f_ref[0]) < 1e-10, k
    x, y, g = c.state()
    assert np.max(np.abs(x - st.x[:st.n])) < 1e-9 and np.max(np.abs(y - st.y[:st.n])) < 1e-9
    assert np.max(np.abs(g - st.g[:st.n])) < 1e-9 * np.max(np.abs(st.g[:st.n]))
    c.close()


C0_CONF = """steps {steps}
saving_step {save}
dt 0.01
inc_step 8192
pr_file profile.txt
viscosity 0.001
rho 1.0
v_inf_x 1.0
v_inf_y 0.0
v_inf_incr_steps 00
n_of_points 20
x_max 10.0
x_min -2.0
y_max 10
y_min -10
ve_size {eps}
n_col 2
r_col_diff_sign 1
r_col_same_sign 0.3
max_ve_g 0.00628162
rc_x 0.0
rc_y 0.0
matrix_load 0
birth_offset 0.5
timings_file output/timings.txt
"""


def test_c0_cli_outputs_against_oracle_loop(ref, tmp_path):
    """The CLI on a C0 run file: F_p.txt (written one step late, "%lf") and
    the last snapshot against the oracle's loop."""
    steps, save = 20, 10
    pan, prm, M, st, hist = c0_oracle(ref, steps)
    tmp = str(tmp_path)
    geo.write_profile(os.path.join(tmp, "profile.txt"), pan)
    eps = float(pan["length"][0])
    assert abs(eps - prm.ve_size) < 1e-15
    with open(os.path.join(tmp, "run.conf"), "w") as f:
        f.write(C0_CONF.format(steps=steps, save=save, eps=repr(eps)))
    out = subprocess.run([CLI, "-c", "run.conf"], cwd=tmp, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    fp = np.loadtxt(os.path.join(tmp, "output", "F_p.txt"))
    assert fp.shape == (steps + 1, 4)
    ref_rows = np.array([f[1:] for _, f in hist])
    # row k + 1 holds the forces of step k (logged one step late), six decimals
    assert np.all(np.abs(fp[1:, 1:] - ref_rows) <= 1.5e-6), np.max(np.abs(fp[1:, 1:] - ref_rows))
    snap = np.loadtxt(os.path.join(tmp, "output", "kadrs", f"Kadr{steps:06d}.txt"), skiprows=1)
    n = st.n
    assert snap.shape[0] == n
    assert np.max(np.abs(snap[:, 2] - st.x[:n])) <= 1.5e-6 and np.max(np.abs(snap[:, 3] - st.y[:n])) <= 1.5e-6
    assert np.max(np.abs(snap[:, 7] - st.g[:n])) <= 1.5e-6


# ---------------------------------------------------------------- C2, C4
def big_profile(name):
    if name == "C2":
        pan = geo.flat_plate(5200, chord=1.0, thickness=0.02, alpha_deg=15.0)
        return pan, dict(nu=2e-4, dt=0.002)
    pan = geo.naca00(2576, t=0.12, chord=1.0, alpha_deg=8.0)
    return pan, dict(nu=1e-4, dt=0.001)


@pytest.mark.parametrize("name", ["C2", "C4"])
def test_large_profile_panel_system_and_first_steps(ref, name):
    import vx_ref
    pan, kw = big_profile(name)
    P = pan.size
    print(f"[configs] {name}: P={P}", flush=True)
    prm = make_params(pan, precision=VX_PREC_F64, r_col_diff_sign=1.0, r_col_same_sign=0.3, **kw)
    c = Case(prm, pan)
    L = c.panel_matrix()
    Lr = ref.panel_matrix(pan)
    # FP64 transcendentals on the GPU against the host's: near-degenerate
    # panel pairs (thin trailing edges) amplify their last-bit differences
    assert np.max(np.abs(L - Lr)) < 1e-10 * np.max(np.abs(Lr))
    c.build_matrix()
    M = c.matrix()
    rng = np.random.default_rng(5)
    for _ in range(3):
        r = rng.normal(size=P + 1)
        assert np.max(np.abs(M @ (Lr @ r) - r)) < 1e-8 * np.max(np.abs(r))
    # the first birth sheet (no vortices yet) against a direct solve
    R = c.right_part(0, VINF)
    Rr = ref.right_part(np.zeros(0), np.zeros(0), np.zeros(0), pan, VINF)
    assert np.max(np.abs(R - Rr)) < 1e-13 * np.max(np.abs(Rr))
    gs = c.birth(Rr, 0.0)
    sol = np.linalg.solve(Lr, np.concatenate([Rr, [0.0]]))
    assert np.max(np.abs(gs - sol[:P])) < 1e-8 * np.max(np.abs(sol[:P]))
    c.close()
    # three full steps: FP64 direct against the oracle loop (same inverse)
    st = vx_ref.State(pan, ref_params(prm), M, capacity=3 * P + 8192)
    c = Case(prm, pan)
    c.set_matrix(M)
    fr = []
    for k in range(3):
        f = c.time_step(VINF)
        fr.append(st.step(VINF))
        assert c.n == st.n, (k, c.n, st.n)
        assert abs(f[1] - fr[-1][1]) < 1e-9 * max(1.0, abs(fr[-1][1])), (k, f, fr[-1])
    x, y, g = c.state()
    assert np.max(np.abs(x - st.x[:st.n])) < 1e-10 and np.max(np.abs(g - st.g[:st.n])) < 1e-10 * np.max(np.abs(g))
    c.close()
    # the mixed path on the tree (tree from the first vortex on) over the same steps
    prm_m = make_params(pan, precision=VX_PREC_MIXED, tree_min_n=1, r_col_diff_sign=1.0, r_col_same_sign=0.3, **kw)
    c = Case(prm_m, pan)
    c.set_matrix(M)
    for k in range(3):
        f = c.time_step(VINF)
        assert abs(f[1] - fr[k][1]) < 1e-3 * max(1.0, abs(fr[k][1])), (k, f, fr[k])
    assert abs(c.n - st.n) <= 2
    c.close()


# ---------------------------------------------------------------- C1 at 2^20
@pytest.mark.parametrize("prec,tol_bs,tol_diff", [(VX_PREC_F64, 2e-4, 1e-9), (VX_PREC_MIXED, 3e-4, 1e-3)])
def test_c1_tree_order16_and_diffusion_at_one_million(ref, prec, tol_bs, tol_diff):
    """The plain path's order-16 tree and the tree diffusion at N = 2^20,
    on 512 sampled targets against the oracle's direct sums."""
    n = 1 << 20
    x, y, g = uniform_wake(n, seed=1)
    eps, nu = math.pi / 1000, 1e-3
    pan = geo.circle(1000, 0.5)
    c = Case(make_params(pan, precision=prec, ve_size=eps, nu=nu, order=16), pan)
    c.set_state(x, y, g)
    c.tree_biot_savart(n, (0.0, 0.0))
    u0, v0, d = c.velocity(n)
    c.tree_diffusion()
    u, v, _ = c.velocity(n)
    c.close()
    sample = 512
    ur, vr, dr = ref.biot_savart(x, y, g, eps * eps, (0.0, 0.0), n_tgt=sample)
    assert rel_err(u0[:sample], ur) < tol_bs and rel_err(v0[:sample], vr) < tol_bs
    dtol = 1e-13 if prec == VX_PREC_F64 else 1e-4
    assert np.max(np.abs(d[:sample] - dr) / dr) < dtol
    ud, vd = ref.diffusion(x, y, g, d, nu, eps, u0[:sample], v0[:sample], n_tgt=sample)
    assert rel_err(u[:sample] - u0[:sample], ud - u0[:sample]) < tol_diff
    assert rel_err(v[:sample] - v0[:sample], vd - v0[:sample]) < tol_diff


# ---------------------------------------------------------------- capacity growth
def test_capacity_grows_across_inc_step_like_the_oracle(ref):
    """inc_step 512 and a 128-panel cylinder with little collapse (the count
    passes 1400 in 20 steps): the capacity grows from 512 to 2048 in three
    quanta while vortices are live (incr_vort_quant, src/unit_1.cu:511-558).
    The growing case equals one whose capacity was reserved up front bit for
    bit over all 20 steps.  It follows the oracle, which grows its arrays
    alike, over the first 8 steps (through the first growth) to the FP64
    tolerances: this flow's vortex count is chaotic beyond that, the oracle
    itself changing its count at step 10 or 11 when 1e-13 noise is added to
    the positions after each step."""
    import vx_ref
    pan = geo.circle(128, 0.5)
    prm = make_params(pan, dt=0.01, r_col_same_sign=0.1, r_col_diff_sign=0.0, birth_offset_rel=0.5, inc_step=512,
                      precision=VX_PREC_F64)
    M = ref.invert(ref.panel_matrix(pan), 1)
    st = vx_ref.State(pan, ref_params(prm), M, capacity=512)
    c = Case(prm, pan)
    c.set_matrix(M)
    big = Case(prm, pan)
    big.set_matrix(M)
    big.reserve(64 * 1024)
    caps = {c.capacity}
    fs, fb, fr = [], [], []
    k_ref = 8
    for k in range(20):
        fs.append(c.time_step(VINF))
        fb.append(big.time_step(VINF))
        caps.add(c.capacity)
        assert c.n == big.n, (k, c.n, big.n)
        if k < k_ref:
            fr.append(st.step(VINF))
            assert c.n == st.n, (k, c.n, st.n)
            if k == k_ref - 1:
                assert len(caps) >= 2, caps
                fx_rms = math.sqrt(np.mean([f[1] ** 2 for f in fr]))
                assert max(abs(a[1] - b[1]) for a, b in zip(fs, fr)) < 1e-9 * fx_rms
                x, y, g = c.state()
                assert np.max(np.abs(x - st.x[:st.n])) < 1e-9
                assert np.max(np.abs(g - st.g[:st.n])) < 1e-12 * np.max(np.abs(g))
    assert len(caps) >= 4 and max(caps) >= 2048, caps
    assert np.array_equal(np.array(fs), np.array(fb))
    for a, b in zip(c.state(), big.state()):
        assert np.array_equal(a, b)
    c.close()
    big.close()


# ---------------------------------------------------------------- Step kind order
def test_step_runs_opposite_sign_collapse_before_same_sign(ref):
    """A(+1) at 0, B(+1) at 0.3 eps, C(-0.5) at 0.7 eps, radii 0.5 eps for
    both kinds.  Opposite sign first (src/unit_1.cu:1138-1216): B and C merge
    (0.4 eps apart), then A and the merged B (0.43 eps): one vortex of 1.5.
    Same sign first would merge A and B, leave C 0.55 eps away: two."""
    pan = geo.circle(64, 0.5)
    eps = float(pan["length"][0])
    prm = make_params(pan, dt=0.01, ve_size=eps, r_col_diff_sign=0.5, r_col_same_sign=0.5, max_ve_g=10.0,
                      n_col=1, precision=VX_PREC_F64)
    x = np.array([2.0, 2.0 + 0.3 * eps, 2.0 + 0.7 * eps])
    y = np.zeros(3)
    g = np.array([1.0, 1.0, -0.5])
    import vx_ref
    xr, yr, gr, nr, _ = vx_ref.Step(x, y, g, np.zeros(3), np.zeros(3), 3, pan, ref_params(prm))
    assert nr == 1 and abs(gr[0] - 1.5) < 1e-15
    c = Case(prm, pan)
    c.set_state(x, y, g)
    c.set_velocity(np.zeros(3), np.zeros(3), np.ones(3))
    c.step()
    X, Y, G = c.state()
    assert c.n == 1 and G[0] == gr[0] and X[0] == xr[0] and Y[0] == yr[0]
    c.close()
