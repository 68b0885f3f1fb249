"""Pins of the oracle's particle stages (A6-A13, A19) against closed forms,
invariants and brute force.  CPU only."""
import math

import numpy as np
import pytest
from scipy import integrate

from vortex_amd import geometry as geo
import parallel_forms as ro


def cloud(n, seed=0, spread=1.0, gmax=0.01):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-spread, spread, n)
    y = rng.uniform(-spread, spread, n)
    g = rng.uniform(-gmax, gmax, n)
    return x, y, g


# ---------------------------------------------------------------- A6 / A19
def test_biot_savart_antisymmetry_invariants(ref):
    x, y, g = cloud(300, 1, spread=0.05)  # many pairs inside the smoothing radius
    eps = 0.01
    u, v, _ = ref.biot_savart(x, y, g, eps * eps)
    scale = np.sum(np.abs(g)) * np.max(np.abs(g)) / eps
    assert abs(np.sum(g * u)) < 1e-13 * scale * len(x)
    assert abs(np.sum(g * v)) < 1e-13 * scale * len(x)
    assert abs(np.sum(g * (x * u + y * v))) < 1e-13 * scale * len(x)


@pytest.mark.parametrize("G,a", [(0.3, 0.2), (1.0, 0.05)])
def test_corotating_and_counterrotating_pairs(ref, G, a):
    x = np.array([-a, a]); y = np.zeros(2)
    u, v, _ = ref.biot_savart(x, y, np.array([G, G]), 1e-6)
    om = G / (4 * math.pi * a * a)
    assert np.allclose(v, [-om * a, om * a], rtol=1e-14) and np.allclose(u, 0, atol=1e-16)
    u, v, _ = ref.biot_savart(x, y, np.array([G, -G]), 1e-6)
    U = G / (4 * math.pi * a)
    assert np.allclose(v, [U, U], rtol=1e-14)  # the pair translates along +y


@pytest.mark.parametrize("N", [3, 7, 16])
def test_thomson_ngon(ref, N):
    R, G = 0.4, 0.02
    th = 2 * math.pi * np.arange(N) / N
    x, y = R * np.cos(th), R * np.sin(th)
    u, v, _ = ref.biot_savart(x, y, np.full(N, G), 1e-8)
    vt = -u * np.sin(th) + v * np.cos(th)
    vr = u * np.cos(th) + v * np.sin(th)
    assert np.allclose(vt, G * (N - 1) / (4 * math.pi * R), rtol=1e-12)
    assert np.max(np.abs(vr)) < 1e-14


def test_smoothing_caps_close_pairs(ref):
    eps = 0.1
    x = np.array([0.0, 0.01]); y = np.zeros(2)
    u, v, _ = ref.biot_savart(x, y, np.array([1.0, 1.0]), eps * eps)
    assert np.allclose(v, [-0.01 / (2 * math.pi * eps * eps), 0.01 / (2 * math.pi * eps * eps)], rtol=1e-14)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 50])
def test_nearest3_brute_force(ref, n):
    x, y, g = cloud(n, 5)
    if n >= 4:
        x[3], y[3] = x[2], y[2]  # a coincident pair counts as distance 0
    _, _, d = ref.biot_savart(x, y, g, 1e-4)
    for i in range(n):
        r2 = np.sort(np.array([(x[i] - x[j]) ** 2 + (y[i] - y[j]) ** 2 for j in range(n) if j != i]
                              + [1e5, 1e5, 1e5]))[:3]
        assert d[i] == pytest.approx(np.sqrt(r2).sum() / 3, rel=1e-15)


def test_probe_equals_biot_savart_at_offset_points(ref):
    x, y, g = cloud(40, 2)
    px = np.array([5.0, -7.0]); py = np.array([1.0, 3.0])
    u, v = ref.velocity_probe(x, y, g, px, py, 1e-4, (1.0, 0.5))
    X = np.concatenate([px, x]); Y = np.concatenate([py, y]); Gm = np.concatenate([[0.0, 0.0], g])
    ub, vb, _ = ref.biot_savart(X, Y, Gm, 1e-4, (1.0, 0.5))
    assert np.allclose(u, ub[:2], rtol=1e-14) and np.allclose(v, vb[:2], rtol=1e-14)


# ---------------------------------------------------------------- A7
def test_diffusion_pair_closed_form(ref):
    """Two vortices: V1 = nu/d1 * (r1 - r2)/|r1 - r2|, independent of g."""
    nu, d = 1e-3, np.array([0.02, 0.03])
    x = np.array([0.0, 0.03]); y = np.array([0.0, 0.04])
    for g in ([0.5, 0.5], [0.5, -0.2], [1e-3, 2.0]):
        u, v = ref.diffusion(x, y, np.array(g), d, nu, 0.01)
        r = math.hypot(0.03, 0.04)
        assert np.allclose([u[0], v[0]], [nu / d[0] * (-0.03 / r), nu / d[0] * (-0.04 / r)], rtol=1e-13)
        assert np.allclose([u[1], v[1]], [nu / d[1] * (0.03 / r), nu / d[1] * (0.04 / r)], rtol=1e-13)


def test_diffusion_ring_is_radial_and_scale_free(ref):
    N, R = 12, 0.05
    th = 2 * math.pi * np.arange(N) / N
    x, y = R * np.cos(th), R * np.sin(th)
    d = np.full(N, 0.02)
    u, v = ref.diffusion(x, y, np.full(N, 0.01), d, 1e-3, 0.001)
    vr = u * np.cos(th) + v * np.sin(th)
    vt = -u * np.sin(th) + v * np.cos(th)
    assert np.all(vr > 0) and np.allclose(vr, vr[0], rtol=1e-12) and np.max(np.abs(vt)) < 1e-15
    u2, v2 = ref.diffusion(x, y, np.full(N, 7.0), d, 1e-3, 0.001)
    assert np.allclose(u2, u, rtol=1e-12) and np.allclose(v2, v, rtol=1e-12)


def test_diffusion_threshold_and_small_I1(ref):
    x = np.array([0.0, 1e-9, 0.5]); y = np.zeros(3)
    # the coincident partner (dist < 0.001 ve_size) is skipped; the far one is
    # below the |I1| > 1e-5 cut, so no diffusive velocity at all
    u, v = ref.diffusion(x, y, np.array([1.0, 1.0, 1.0]), np.full(3, 0.01), 1e-3, 0.01)
    assert u[0] == 0 and v[0] == 0


# ---------------------------------------------------------------- A8
def test_wall_far_and_near_cases(ref):
    pan = geo.circle(100, 0.5)
    nu = 1e-3
    # far from the body: no panel within 5 lengths -> unchanged
    u, v = ref.diffusion_wall(np.array([3.0]), np.array([0.0]), np.array([0.01]), pan, nu, 20)
    assert u[0] == 0 and v[0] == 0
    # on top of panel 7's birth point: closed form of the first near panel
    b = pan["birth"][7]; dd = 0.004; L = pan["length"][7]; n = pan["norm"][7]
    u, v = ref.diffusion_wall(np.array([b[0]]), np.array([b[1]]), np.array([dd]), pan, nu, 20)
    ex = 2 * dd * (1 - math.exp(-L / (2 * dd))) / (math.pi * dd * dd) * nu
    assert np.allclose([u[0], v[0]], ex * n, rtol=1e-13)


def test_wall_midrange_vs_adaptive_quadrature(ref):
    """With many midpoints the quadrature of I_0/I_3 converges to the integral."""
    pan = geo.circle(60, 0.5)
    nu, dd = 1e-3, 0.02
    L = pan["length"][0]
    pt = pan["contr"][0] + 1.0 * L * pan["norm"][0]
    u, v = ref.diffusion_wall(np.array([pt[0]]), np.array([pt[1]]), np.array([dd]), pan, nu, 4000)
    II0 = 2 * math.pi * dd * dd
    II3 = np.zeros(2)
    for f in range(pan.size):
        r = math.hypot(pt[0] - pan["birth"][f][0], pt[1] - pan["birth"][f][1])
        lf = pan["length"][f]
        if not (0.1 * lf < r < 5 * lf):
            assert r > 0.1 * lf
            continue
        a0, a1, nn = pan["left"][f], pan["right"][f], -pan["norm"][f]

        def eta(s):
            q = a0 + s * (a1 - a0)
            return (pt - q) / dd

        def f0(s):
            e = eta(s); me = math.hypot(*e)
            return (e @ nn) / (e @ e) * (me + 1) * math.exp(-me) * lf

        def f3(s, k):
            e = eta(s)
            return nn[k] * math.exp(-math.hypot(*e)) * lf

        II0 += -dd * integrate.quad(f0, 0, 1, epsabs=1e-14)[0]
        II3 -= np.array([integrate.quad(f3, 0, 1, args=(k,), epsabs=1e-14)[0] for k in (0, 1)])
    ex = nu * II3 / II0
    assert np.allclose([u[0], v[0]], ex, rtol=2e-6, atol=1e-12)


# ---------------------------------------------------------------- A9 / A10
def square_body():
    vx = np.array([-1.0, -1.0, 1.0, 1.0]); vy = np.array([-1.0, 1.0, 1.0, -1.0])
    return geo.panels_from_vertices(vx, vy)


def params(pan, **kw):
    import vx_ref
    bx = geo.profile_bbox(pan)
    p = vx_ref.Params(dt=0.1, ve_size=0.01, nu=1e-3, rho=1.0, vinf_x=1.0, vinf_y=0.0,
                      x_min=-5, x_max=5, y_min=-5, y_max=5,
                      px_min=bx[0], px_max=bx[1], py_min=bx[2], py_max=bx[3],
                      r_col_diff_sign=1.0, r_col_same_sign=0.25, max_ve_g=0.1, rc_x=0.0, rc_y=0.0,
                      birth_offset_rel=0.0, n_of_points=20, n_col=2)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_hitting_integer_cases(ref):
    pan = square_body()
    prm = params(pan)
    # top panel runs (-1,1) -> (1,1) = panel 1
    assert ref.hitting(pan, prm, (0.0, 0.5), (0.0, 2.0)) == (True, 1)
    # straight through: the crossing nearest the old position wins
    assert ref.hitting(pan, prm, (0.0, -2.0), (0.0, 2.0)) == (True, 1)
    assert ref.hitting(pan, prm, (0.0, 2.0), (0.0, -2.0)) == (True, 3)
    # misses: outside the profile box, and a chord inside the box but outside the body
    assert ref.hitting(pan, prm, (3.0, 3.0), (2.0, 2.0))[0] is False
    assert ref.hitting(pan, prm, (0.5, 1.5), (0.9, 1.2))[0] is False
    # start exactly on a panel (within the 1e-12 snap) counts as a crossing
    assert ref.hitting(pan, prm, (0.0, 0.0), (0.0, 1.0 + 1e-13))[0] is True


def test_step_move_absorb_delete_and_impulse(ref):
    pan = square_body()
    prm = params(pan, dt=1.0)
    x = np.array([0.0, 4.5, 3.0, 2.0, 0.0]); y = np.array([2.0, 0.0, 3.0, 2.0, -3.0])
    g = np.array([0.5, 0.2, 0.0, -0.3, 0.7])
    u = np.array([0.0, 1.0, 0.0, 0.1, 0.0]); v = np.array([-1.5, 0.0, 0.0, 0.0, 0.1])
    xo, yo, go, out, (dg, fx, fy, m) = ref.step_move(x, y, g, u, v, 5, pan, prm)
    assert go[0] == 0 and dg[0] == 0.5 and (xo[0], yo[0]) == (0.0, 0.5)       # absorbed
    assert go[1] == 0 and (xo[1], yo[1]) == (-2e5, -2e5)                        # left the box
    assert go[2] == 0 and (xo[2], yo[2]) == (3.0, 3.0)                          # dead, moved only
    assert go[3] == -0.3 and xo[3] == 2.1                                       # free
    # entries 1..4 are the P = 4 newborns: impulse g(-y, x) and g |r - rc|^2;
    # the absorbed entry 0 subtracts g(-y_c, x_c) at the hit panel's control point
    c = pan["contr"][1]
    assert (fx[0], fy[0], m[0]) == (-(0.5 * -c[1]), -(0.5 * c[0]), -(0.5 * (c[0] ** 2 + c[1] ** 2)))
    assert (fx[4], fy[4], m[4]) == (0.7 * 3.0, 0.0, 0.7 * 9.0)
    assert (fx[1], fy[1]) == (0.2 * -0.0, 0.2 * 4.5)
    assert out[0] == 0.5
    assert out[1] == pytest.approx(fx.sum() / prm.dt, rel=1e-15)
    assert out[3] == pytest.approx(m.sum() / (2 * prm.dt), rel=1e-15)


# ---------------------------------------------------------------- second component
def test_points_move_absorb_delete_no_impulse(ref):
    """Tracer points move like vortices but carry no impulse terms; the same
    point set as the vortex case gives the same positions and flags."""
    pan = square_body()
    prm = params(pan, dt=1.0)
    x = np.array([0.0, 4.5, 3.0, 2.0, 0.0]); y = np.array([2.0, 0.0, 3.0, 2.0, -3.0])
    g = np.array([1.0, 1.0, 0.0, 1.0, 1.0])
    u = np.array([0.0, 1.0, 0.0, 0.1, 0.0]); v = np.array([-1.5, 0.0, 0.0, 0.0, 0.1])
    xo, yo, go = ref.points_move(x, y, g, u, v, pan, prm)
    assert go[0] == 0 and (xo[0], yo[0]) == (0.0, 0.5)       # absorbed at the top panel
    assert go[1] == 0 and (xo[1], yo[1]) == (-2e5, -2e5)      # left the box
    assert go[2] == 0 and (xo[2], yo[2]) == (3.0, 3.0)        # dead, moved only
    assert go[3] == 1.0 and xo[3] == 2.1                      # free
    assert go[4] == 1.0 and (xo[4], yo[4]) == (0.0, -2.9)
    vx_, vy_, vg, _, _ = ref.step_move(x, y, g, u, v, 5, pan, prm)
    assert np.array_equal(xo, vx_) and np.array_equal(yo, vy_) and np.array_equal(go, vg)


def test_points_relax_known_answers(ref):
    """v = (v_env - v_prev)/rel_t*dt + v_prev with v_prev read as FP32; 0 where
    g = 0 (src/kernel.cu:1478-1489)."""
    g = np.array([1.0, 1.0, 0.0, 2.0])
    ue = np.array([1.5, 0.25, 9.0, 0.1]); ve = np.array([-0.5, 0.25, 9.0, 0.3])
    u = np.array([0.5, 0.25, 4.0, 0.1]); v = np.array([0.5, 0.25, 4.0, 0.3])
    uo, vo = ref.points_relax(g, ue, ve, u, v, 0.5, 0.25)
    assert (uo[0], vo[0]) == (1.0, 0.0)                 # exact binary fractions
    assert (uo[1], vo[1]) == (0.25, 0.25)               # at rest in the flow
    assert (uo[2], vo[2]) == (0.0, 0.0)                 # absorbed point stops
    # a previous velocity that FP32 cannot hold is rounded before use
    p = float(np.float32(0.1))
    assert p != 0.1 and uo[3] == (0.1 - p) / 0.5 * 0.25 + p


# ---------------------------------------------------------------- A11
@pytest.mark.parametrize("seed", range(12))
def test_compact_matches_closed_form(ref, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(0, 200))
    x = rng.normal(size=n); y = rng.normal(size=n)
    g = rng.normal(size=n)
    dead = rng.random(n) < rng.uniform(0, 1)
    g[dead] = rng.choice([0.0, 1e-13, -5e-13], size=dead.sum())
    xo, yo, go, nn = ref.compact(x, y, g)
    xc, yc, gc, nc = ro.compact_closed_form(x, y, g, n)
    assert nn == nc
    assert np.array_equal(xo, xc) and np.array_equal(yo, yc) and np.array_equal(go, gc)


def test_compact_edge_cases(ref):
    for g in ([], [0.0], [1.0], [0.0, 0.0, 0.0], [1.0, 2.0], [0.0, 1.0], [1.0, 0.0]):
        g = np.array(g, dtype=float); n = g.size
        x = np.arange(n, dtype=float); y = -x
        xo, yo, go, nn = ref.compact(x, y, g)
        xc, yc, gc, nc = ro.compact_closed_form(x, y, g, n)
        assert nn == nc and np.array_equal(xo, xc) and np.array_equal(go, gc)


# ---------------------------------------------------------------- A12 / A13
def brute_col(x, y, g, kind, r2, maxg):
    n = x.size
    col = np.full(n, -1, dtype=np.int32)
    for i in range(n):
        j = np.arange(i + 1, n)
        d2 = (x[i] - x[j]) * (x[i] - x[j]) + (y[i] - y[j]) * (y[i] - y[j])
        if kind == 1:
            ok = (g[i] * g[j] < 0) & (d2 < r2)
        else:
            ok = (g[i] * g[j] > 0) & (d2 < r2) & (np.abs(g[i] + g[j]) < maxg)
        if ok.any():
            col[i] = j[ok][0]
    return col


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("seed", range(6))
def test_setka_and_greedy_collapse(ref, kind, seed):
    rng = np.random.default_rng(100 + seed)
    n = 400
    x = rng.uniform(0, 0.2, n); y = rng.uniform(0, 0.2, n)
    g = rng.uniform(-0.01, 0.01, n)
    g[rng.random(n) < 0.05] = 0.0
    r2, maxg = 0.012 ** 2, 0.012
    col = ref.setka(x, y, g, kind, r2, maxg)
    assert np.array_equal(col, brute_col(x, y, g, kind, r2, maxg))
    assert (col >= 0).sum() > 10
    xs, ys, gs = ref.collapse(x, y, g, col, kind)
    xp, yp, gp = ro.collapse_parallel(x, y, g, col, kind)
    assert np.array_equal(xs, xp) and np.array_equal(ys, yp) and np.array_equal(gs, gp)
    assert abs(gs.sum() - g.sum()) < 1e-15
    if kind == 2:
        live = gs != 0
        assert np.allclose((gs * xs)[live].sum(), (g * x).sum(), rtol=1e-12)


def test_collapse_chains_and_shared_targets(ref):
    """Hand-made index graphs: chains (i -> j -> k) and several i sharing one j."""
    x = np.arange(8, dtype=float) * 1e-3; y = np.zeros(8)
    g = np.array([1, 2, 3, 4, 5, 6, 7, 8], dtype=float) * 1e-3
    for col in ([1, 2, 3, -1, 5, 6, 7, -1], [4, 4, 4, 5, 5, -1, -1, -1], [7, 7, 7, 7, 7, 7, 7, -1],
                [1, 3, 3, 7, 6, 6, 7, -1]):
        col = np.array(col, dtype=np.int32)
        for kind in (1, 2):
            a = ref.collapse(x, y, g, col, kind)
            b = ro.collapse_parallel(x, y, g, col, kind)
            for p, q in zip(a, b):
                assert np.array_equal(p, q), (col, kind)


# ---------------------------------------------------------------- loop
def test_time_loop_circulation_bookkeeping(ref):
    """Sum of live g plus the pending absorbed d_g stays zero when nothing leaves the box."""
    import vx_ref
    pan = geo.circle(64, 0.5)
    prm = params(pan, dt=0.01, ve_size=pan["length"][0], x_min=-1e3, x_max=1e3, y_min=-1e3, y_max=1e3,
                 max_ve_g=1.0, r_col_same_sign=0.3, r_col_diff_sign=0.3, birth_offset_rel=0.5)
    M = ref.invert(ref.panel_matrix(pan), 1)
    st = vx_ref.State(pan, prm, M, capacity=4096)
    for k in range(15):
        st.step((1.0, 0.0))
        tot = st.g[:st.n].sum() + st.d_g
        assert abs(tot) < 1e-12 * max(1.0, np.abs(st.g[:st.n]).sum()), (k, tot)
    assert st.n > 64


def test_step_kind_order_known_answer(ref):
    """Step runs the opposite-sign collapse passes before the same-sign ones
    (src/unit_1.cu:1138-1216).  A(+1) at 0, B(+1) at 0.3 eps, C(-0.5) at
    0.7 eps, both radii 0.5 eps: B and C merge first (0.4 eps apart; |g|-
    weighted centroid 0.65/1.5 eps, g = 0.5), then A and the merged vortex
    (0.433 eps apart; g-weighted centroid at 0.2167/1.5 eps, g = 1.5): one
    vortex.  Same-sign first would merge A and B and leave C 0.55 eps away."""
    pan = geo.circle(64, 0.5)
    eps = float(pan["length"][0])
    prm = params(pan, ve_size=eps, r_col_diff_sign=0.5, r_col_same_sign=0.5, max_ve_g=10.0, n_col=1)
    x = np.array([2.0, 2.0 + 0.3 * eps, 2.0 + 0.7 * eps])
    y = np.zeros(3)
    g = np.array([1.0, 1.0, -0.5])
    xr, yr, gr, nr, _ = ref.Step(x, y, g, np.zeros(3), np.zeros(3), 3, pan, prm)
    assert nr == 1
    assert gr[0] == 1.5 and yr[0] == 0.0
    xb = (x[1] * 1.0 + x[2] * 0.5) / 1.5
    assert abs(xr[0] - (x[0] * 1.0 + xb * 0.5) / 1.5) < 1e-14
