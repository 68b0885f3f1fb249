"""GPU parity: every stage of libvortex_amd.so against the FP64 oracle on the
same seeded inputs.  Bars:
  * integer outcomes (absorption, deletion, compaction order and contents,
    collapse pairs and merged positions) -- bit-exact;
  * FP64 path sums -- relative 1e-12 of the largest magnitude (summation
    order differs from the oracle's serial loop);
  * mixed path (FP32 pairs, FP64 tile sums) -- velocities relative 2e-4
    (Biot-Savart, probes) and 1e-3 (diffusion) of the largest magnitude,
    nearest-3 distance relative 1e-5.  FP32 coordinate differences carry an
    error of about 6e-8 |x| / |dx| for close pairs, which bounds these.
"""
import math

import numpy as np
import pytest

from vortex_amd import geometry as geo
from vortex_amd.solver import Case, make_params, VX_PREC_F64, VX_PREC_MIXED
from vortex_amd.state import far_tail

pytestmark = pytest.mark.gpu

PRECS = [VX_PREC_F64, VX_PREC_MIXED]
TOL_V = {VX_PREC_F64: 1e-12, VX_PREC_MIXED: 2e-5}  # x10 in the asserts


def ref_params(prm):
    import vx_ref
    return vx_ref.Params(**{k: getattr(prm, k) for k, _ in vx_ref.Params._fields_})


def wake(n, seed, body_r=0.5, spread=(0.5, 1.6), gmax=0.004):
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, 2 * np.pi, n)
    r = rng.uniform(*spread, n)
    x = r * np.cos(th); y = r * np.sin(th)
    g = rng.uniform(-gmax, gmax, n)
    return x, y, g


def with_tail(x, y, g, cap):
    n = x.size
    tx, ty = far_tail(n, cap - n)
    return (np.concatenate([x, tx]), np.concatenate([y, ty]), np.concatenate([g, np.zeros(cap - n)]))


def rel_err(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def cyl():
    return geo.circle(200, 0.5)


def make_case(pan, prec, **kw):
    prm = make_params(pan, precision=prec, **kw)
    return Case(prm, pan), prm


# ---------------------------------------------------------------- A1 / A2
@pytest.mark.parametrize("maker", [lambda: geo.circle(200, 0.5), lambda: geo.naca00(256, 0.12, alpha_deg=8.0),
                                   lambda: geo.flat_plate(300, alpha_deg=15.0)])
def test_panel_matrix_and_inverse(ref, maker):
    pan = maker()
    c, _ = make_case(pan, VX_PREC_F64)
    c.build_matrix()
    L = c.panel_matrix()
    Lr = ref.panel_matrix(pan)
    assert rel_err(L, Lr) < 1e-11  # device atan/log vs glibc, amplified by cancellation
    M = c.matrix()
    Mr = ref.invert(Lr, 1)
    assert rel_err(M, Mr) < 1e-10
    assert np.abs(M @ Lr - np.eye(pan.size + 1)).max() < 1e-9


def test_bad_arguments_are_rejected(cyl):
    from vortex_amd._lib import VortexError
    c, _ = make_case(cyl, VX_PREC_F64)
    with pytest.raises(VortexError):
        c.set_matrix(np.zeros((5, 5)))
    with pytest.raises(VortexError):
        c.biot_savart(10, 5)
    with pytest.raises(VortexError):
        c.state(c.capacity + 1)
    with pytest.raises(VortexError):
        c.vort_creation((1.0, 0.0))  # no panel matrix yet


# ---------------------------------------------------------------- A3 / A4
@pytest.mark.parametrize("prec", PRECS)
def test_right_part(ref, cyl, prec):
    c, _ = make_case(cyl, prec)
    x, y, g = wake(3000, 1)
    n_src = 3008
    X, Y, G = with_tail(x, y, g, n_src)
    c.set_state(X, Y, G)
    R = c.right_part(n_src, (1.0, 0.2))
    Rr = ref.right_part(X, Y, G, cyl, (1.0, 0.2))
    assert rel_err(R, Rr) < (1e-12 if prec == VX_PREC_F64 else 5e-5)


def test_birth(ref, cyl):
    c, _ = make_case(cyl, VX_PREC_F64)
    c.build_matrix()
    M = c.matrix()
    rng = np.random.default_rng(2)
    R = rng.normal(size=cyl.size)
    x, y, g = wake(100, 3)
    c.set_state(x, y, g)
    gs = c.birth(R, 0.013)
    xo, yo, go, gsr = ref.birth(M, R, 0.013, cyl)
    assert rel_err(gs, gsr) < 1e-13
    X, Y, G = c.state()
    assert c.n == 300
    assert np.array_equal(X[100:], xo) and np.array_equal(Y[100:], yo)
    assert rel_err(G[100:], go) < 1e-13
    assert np.array_equal(X[:100], x)


# ---------------------------------------------------------------- A6 / A7 / A8 / A19
@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("n", [2, 3, 257, 4097])
def test_biot_savart_and_nearest3(ref, cyl, prec, n):
    c, prm = make_case(cyl, prec, ve_size=0.004)
    x, y, g = wake(n, 10 + n)
    if n > 10:
        x[7], y[7] = x[3], y[3]
    s = (n + 31) // 32 * 32
    X, Y, G = with_tail(x, y, g, s)
    c.set_state(X, Y, G)
    c.biot_savart(s, s, (1.0, 0.0))
    u, v, d = c.velocity(s)
    ur, vr, dr = ref.biot_savart(X, Y, G, prm.ve_size ** 2, (1.0, 0.0))
    tol = TOL_V[prec]
    assert rel_err(u - 1.0, ur - 1.0) < tol * 10 and rel_err(v, vr) < tol * 10
    assert np.max(np.abs(d - dr) / dr) < (1e-14 if prec == VX_PREC_F64 else 1e-5)


@pytest.mark.parametrize("prec", PRECS)
def test_diffusion(ref, cyl, prec):
    c, prm = make_case(cyl, prec, ve_size=0.004, nu=1e-3)
    n = 2000
    x, y, g = wake(n, 21, spread=(0.5, 0.7))
    s = 2016
    X, Y, G = with_tail(x, y, g, s)
    rng = np.random.default_rng(5)
    d = rng.uniform(0.002, 0.02, s)
    u0 = rng.normal(size=s); v0 = rng.normal(size=s)
    c.set_state(X, Y, G)
    c.set_velocity(u0, v0, d)
    c.diffusion(s, s)
    u, v, _ = c.velocity(s)
    ur, vr = ref.diffusion(X, Y, G, d, prm.nu, prm.ve_size, u0, v0)
    tol = 1e-12 if prec == VX_PREC_F64 else 1e-3
    assert rel_err(u - u0, ur - u0) < tol and rel_err(v - v0, vr - v0) < tol


@pytest.mark.parametrize("prec", PRECS)
def test_diffusion_wall(ref, prec):
    pan = geo.circle(120, 0.5)
    c, prm = make_case(pan, prec, nu=1e-3)
    rng = np.random.default_rng(8)
    n = 3000
    th = rng.uniform(0, 2 * np.pi, n)
    r = 0.5 + rng.uniform(-0.002, 0.15, n)
    x, y = r * np.cos(th), r * np.sin(th)
    x[:40] = pan["birth"][:40, 0] + 1e-5   # within 0.1 l of a birth point
    y[:40] = pan["birth"][:40, 1]
    g = rng.uniform(-0.01, 0.01, n)
    d = rng.uniform(0.002, 0.03, n)
    c.set_state(x, y, g)
    c.set_velocity(np.zeros(n), np.zeros(n), d)
    c.diffusion_wall(n)
    u, v, _ = c.velocity(n)
    ur, vr = ref.diffusion_wall(x, y, d, pan, prm.nu, prm.n_of_points)
    tol = 1e-12 if prec == VX_PREC_F64 else 5e-4
    scale = max(np.abs(ur).max(), np.abs(vr).max())
    assert np.max(np.abs(u - ur)) < tol * scale and np.max(np.abs(v - vr)) < tol * scale
    assert np.count_nonzero(ur) > 100


@pytest.mark.parametrize("prec", PRECS)
def test_velocity_probe(ref, cyl, prec):
    c, prm = make_case(cyl, prec, ve_size=0.004)
    x, y, g = wake(5000, 30)
    c.set_state(x, y, g)
    px = np.full(500, -0.15); py = 0.01 + 0.002 * np.arange(500)  # src/main.c:521-526
    u, v = c.velocity_probe(px, py, (1.0, 0.0))
    s = 5024
    X, Y, G = with_tail(x, y, g, s)
    ur, vr = ref.velocity_probe(X, Y, G, px, py, prm.ve_size ** 2, (1.0, 0.0))
    assert rel_err(u - 1, ur - 1) < TOL_V[prec] * 10 and rel_err(v, vr) < TOL_V[prec] * 10


# ---------------------------------------------------------------- A9 - A13: bit-exact
def crossing_state(pan, n, seed):
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, 2 * np.pi, n)
    r = rng.uniform(0.49, 0.6, n)
    x, y = r * np.cos(th), r * np.sin(th)
    g = rng.uniform(-0.01, 0.01, n)
    g[rng.random(n) < 0.05] = 0.0
    u = rng.normal(0, 3.0, n) - 2.0 * np.cos(th)  # many move into the body
    v = rng.normal(0, 3.0, n) - 2.0 * np.sin(th)
    x[:30] = rng.uniform(9.5, 11.0, 30)            # some leave the deletion box
    return x, y, g, u, v


@pytest.mark.parametrize("seed", [0, 1])
def test_step_move_bitexact(ref, seed):
    pan = geo.circle(200, 0.5)
    c, prm = make_case(pan, VX_PREC_F64, dt=0.02)
    n = 5000
    x, y, g, u, v = crossing_state(pan, n, seed)
    c.set_state(x, y, g)
    c.set_velocity(u, v)
    out = c.step_move()
    X, Y, G = c.state()
    xr, yr, gr, outr, _ = ref.step_move(x, y, g, u, v, n, pan, ref_params(prm))
    assert np.array_equal(X, xr) and np.array_equal(Y, yr) and np.array_equal(G, gr)
    absorbed = np.count_nonzero((g != 0) & (gr == 0) & (xr != -2e5))
    assert absorbed > 50
    assert out[0] == pytest.approx(outr[0], rel=1e-12, abs=1e-15)
    for q in (1, 2, 3):
        assert abs(out[q] - outr[q]) <= 1e-12 * max(1.0, abs(outr[q]))


@pytest.mark.parametrize("seed", [0, 1])
def test_points_move_and_relax_bitexact(ref, seed):
    """Second component: tracer points move, are absorbed and deleted as the
    oracle does; the inertial relaxation matches it bit for bit."""
    from vortex_amd.solver import points_relax
    pan = geo.circle(200, 0.5)
    c, prm = make_case(pan, VX_PREC_F64, dt=0.02)
    x, y, g, u, v = crossing_state(pan, 3000, seed)
    g = np.where(g != 0, 1.0, 0.0)
    X, Y, G = c.points_move(x, y, g, u, v)
    xr, yr, gr = ref.points_move(x, y, g, u, v, pan, ref_params(prm))
    assert np.array_equal(X, xr) and np.array_equal(Y, yr) and np.array_equal(G, gr)
    assert np.count_nonzero((g != 0) & (gr == 0) & (xr != -2e5)) > 20   # some absorbed
    assert np.count_nonzero(xr == -2e5) > 5                              # some deleted
    rng = np.random.default_rng(seed)
    ue, ve = rng.normal(0, 1, x.size), rng.normal(0, 1, x.size)
    U, V = points_relax(G, ue, ve, u, v, 0.1, 0.02)
    ur, vr = ref.points_relax(gr, ue, ve, u, v, 0.1, 0.02)
    assert np.array_equal(U, ur) and np.array_equal(V, vr)


@pytest.mark.parametrize("seed", range(4))
def test_compact_bitexact(ref, cyl, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 20000))
    x = rng.normal(size=n); y = rng.normal(size=n); g = rng.normal(size=n)
    dead = rng.random(n) < rng.uniform(0.0, 0.9)
    g[dead] = rng.choice([0.0, 5e-13], size=dead.sum())
    c, _ = make_case(cyl, VX_PREC_F64)
    c.set_state(x, y, g)
    c.compact()
    X, Y, G = c.state(n)
    xr, yr, gr, nr = ref.compact(x, y, g)
    assert c.n == nr
    assert np.array_equal(X, xr) and np.array_equal(Y, yr) and np.array_equal(G, gr)


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("seed", range(3))
def test_setka_collapse_bitexact(ref, cyl, kind, seed):
    rng = np.random.default_rng(40 + seed)
    n = 30000
    x = rng.uniform(-1, 3, n); y = rng.uniform(-1, 1, n)
    g = rng.uniform(-0.004, 0.004, n)
    c, prm = make_case(cyl, VX_PREC_F64, ve_size=0.004, r_col_diff_sign=1.0, r_col_same_sign=0.6,
                       max_ve_g=0.006)
    c.set_state(x, y, g)
    col = c.setka(kind)
    r = prm.r_col_diff_sign if kind == 1 else prm.r_col_same_sign
    r2 = r * r * (prm.ve_size * prm.ve_size)
    colr = ref.setka(x, y, g, kind, r2, prm.max_ve_g)
    assert np.array_equal(col, colr)
    assert (colr >= 0).sum() > 100
    c.collapse(kind)
    X, Y, G = c.state(n)
    xr, yr, gr = ref.collapse(x, y, g, colr, kind)
    assert np.array_equal(X, xr) and np.array_equal(Y, yr) and np.array_equal(G, gr)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("r_same,share", [(0.6, 0), (0.6, 1), (1.0, 1)])
def test_Step_bitexact_from_identical_velocities(ref, prec, r_same, share):
    """Step (move, compaction, both kinds' collapse passes) bit for bit against
    the oracle: with a grid built for every collapse pass (the default below
    2^17 vortices) and, forced by col_share_min, with grids kept from pass to
    pass, one per kind (radii 1.0 and 0.6) or one for both (equal radii)."""
    pan = geo.circle(200, 0.5)
    c, prm = make_case(pan, prec, dt=0.02, ve_size=0.004, r_col_diff_sign=1.0, r_col_same_sign=r_same,
                       max_ve_g=0.006)
    if share:
        c.set_tuning(col_share_min=1)
    n = 8000
    x, y, g, u, v = crossing_state(pan, n, 7)
    c.set_state(x, y, g)
    c.set_velocity(u, v)
    out = c.step()
    print(f"[Step] r_same={r_same} share={share}: {n} -> {c.n}", flush=True)
    X, Y, G = c.state(n)
    xr, yr, gr, nr, outr = ref.Step(x, y, g, u, v, n, pan, ref_params(prm))
    assert c.n == nr
    assert np.array_equal(X, xr) and np.array_equal(Y, yr) and np.array_equal(G, gr)
    assert abs(out[1] - outr[1]) <= 1e-12 * max(1.0, abs(outr[1]))


# ---------------------------------------------------------------- full step
@pytest.mark.parametrize("prec", PRECS)
def test_time_step_from_oracle_state(ref, prec):
    """Run the oracle 6 steps on a 64-panel cylinder, hand its state to the
    GPU, and compare one full time step (birth, Speed, Step) of both."""
    import vx_ref
    pan = geo.circle(64, 0.5)
    prm = make_params(pan, dt=0.01, r_col_same_sign=0.3, r_col_diff_sign=0.3, birth_offset_rel=0.5,
                      precision=prec)
    M = ref.invert(ref.panel_matrix(pan), 1)
    st = vx_ref.State(pan, ref_params(prm), M, capacity=8192)
    for _ in range(6):
        st.step((1.0, 0.0))
    c = Case(prm, pan)
    c.set_matrix(M)
    n0 = st.n
    s = (n0 + 64 + 31) // 32 * 32
    c.set_state(st.x[:s], st.y[:s], st.g[:s])
    c.n = n0
    c.d_g = st.d_g
    f = c.time_step((1.0, 0.0))
    fr = st.step((1.0, 0.0))
    X, Y, G = c.state()
    if prec == VX_PREC_F64:
        assert c.n == st.n
        assert np.max(np.abs(X - st.x[:st.n])) < 1e-12 and np.max(np.abs(G - st.g[:st.n])) < 1e-14
        assert abs(f[1] - fr[1]) < 1e-9 * max(1, abs(fr[1]))
    else:
        assert abs(c.n - st.n) <= 2
        assert abs(f[1] - fr[1]) < 1e-3 * max(1, abs(fr[1]))


# ---------------------------------------------------------------- full size properties
def test_one_million_invariants_and_mixed_vs_f64():
    """N = 2^20: the antisymmetry invariants of the smoothed kernel hold for
    the F64 path, and the mixed path agrees with it."""
    from vortex_amd.state import uniform_wake
    pan = geo.circle(1000, 0.5)
    n = 1 << 20
    x, y, g = uniform_wake(n, seed=1)
    out = {}
    for prec in PRECS:
        c, prm = make_case(pan, prec, ve_size=math.pi / 1000)
        c.set_state(x, y, g)
        c.biot_savart(n, n, (0.0, 0.0))
        out[prec] = c.velocity(n)
        c.close()
    u, v, d = out[VX_PREC_F64]
    scale = np.sum(np.abs(g)) * np.abs(g).max() / (math.pi / 1000)
    assert abs(np.sum(g * u)) < 1e-10 * scale and abs(np.sum(g * v)) < 1e-10 * scale
    um, vm, dm = out[VX_PREC_MIXED]
    assert rel_err(um, u) < 3e-4 and rel_err(vm, v) < 3e-4
    assert np.max(np.abs(dm - d) / d) < 1e-3  # direct mixed path: FP32 absolute coordinates
