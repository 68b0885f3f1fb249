"""GPU parity of the Morton-tree velocity stages against the FP64 oracle's
direct sums (the tree has no counterpart in the reference: its own tree is
inactive in every shipped run file and is not the parity target).

Stated tolerances (relative to the largest magnitude of the compared field):
  * Biot-Savart, FP64, 32 multipole terms, theta 0.5: 1e-8
  * Biot-Savart, FP64, 16 terms: 2e-4;  mixed, 16 terms: 3e-4
  * nearest-3 distance: FP64 1e-13 (exact up to rounding), mixed 1e-4
    (bucket-relative FP32 coordinates)
  * diffusion (cut-off 30 d FP64, 18 d mixed; the isolated vortices take the
    domain-scale path): FP64 1e-9, mixed 1e-3 (FP32 terms, amplified where
    I1 = sum g e^{-r/d} cancels)
"""
import math

import numpy as np
import pytest

from vortex_amd import geometry as geo
from vortex_amd.solver import Case, make_params, VX_PREC_F64, VX_PREC_MIXED
from vortex_amd.state import far_tail, uniform_wake

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def wake_cloud(n, seed):
    """A clustered wake: dense shear layers near a cylinder, sparse far field,
    a few isolated vortices (exercise the exact nearest-3 fallback)."""
    rng = np.random.default_rng(seed)
    k = n // 2
    th = rng.uniform(0, 2 * np.pi, k)
    r = 0.5 + rng.exponential(0.02, k)
    x1, y1 = r * np.cos(th), r * np.sin(th)
    x2 = rng.uniform(0.5, 4.0, n - k - 8)
    y2 = 0.3 * np.sin(2 * x2) + rng.normal(0, 0.05, n - k - 8)
    x3 = np.array([6.0, 7.5, -1.5, 3.0, 8.0, 9.5, 5.0, -1.9])
    y3 = np.array([3.0, -4.0, 1.0, 6.0, 0.0, 2.0, -8.0, -1.0])
    x = np.concatenate([x1, x2, x3]); y = np.concatenate([y1, y2, y3])
    g = rng.uniform(-0.004, 0.004, n)
    return x, y, g


def tail(x, y, g, s):
    n = x.size
    tx, ty = far_tail(n, s - n)
    return np.concatenate([x, tx]), np.concatenate([y, ty]), np.concatenate([g, np.zeros(s - n)])


@pytest.fixture(scope="module")
def cloud40k(ref):
    n = 40000
    x, y, g = wake_cloud(n, 1)
    s = (n + 31) // 32 * 32
    X, Y, G = tail(x, y, g, s)
    eps = 0.004
    u, v, d = ref.biot_savart(X, Y, G, eps * eps, (1.0, 0.0))
    return X, Y, G, s, eps, u, v, d


@pytest.mark.parametrize("prec,order,tol", [(VX_PREC_F64, 32, 1e-10), (VX_PREC_F64, 20, 1e-7),
                                            (VX_PREC_F64, 16, 2e-6), (VX_PREC_MIXED, 16, 2e-6),
                                            (VX_PREC_MIXED, 20, 2e-6), (VX_PREC_MIXED, 32, 2e-6)])
def test_tree_biot_savart_vs_direct_oracle(cloud40k, prec, order, tol):
    """Orders 16, 20 and 32 (truncation about theta^order, theta = 0.5;
    FP32 far fields through the matrix-core M2L) against the direct sums.
    Measured on this cloud: FP64 2e-7, 1e-8 and 1.5e-12 of the largest
    velocity at orders 16, 20 and 32; mixed 2.7e-7 at every order (the FP32
    floor); the bars keep about a decade of margin."""
    X, Y, G, s, eps, ur, vr, dr = cloud40k
    pan = geo.circle(200, 0.5)
    c = Case(make_params(pan, precision=prec, ve_size=eps, order=order), pan)
    c.set_state(X, Y, G)
    c.tree_biot_savart(s, (1.0, 0.0))
    u, v, d = c.velocity(s)
    info = c.tree_stats()
    print(f"[tree] prec={prec} order={order}: rel err u {rel_err(u - 1.0, ur - 1.0):.3g} v {rel_err(v, vr):.3g}",
          flush=True)
    assert info["order"] == order
    assert info["m"] == s  # every slot is near the box here
    assert info["m2p_evals"] > 0 and info["bs_pairs"] < s * s / 10
    assert rel_err(u - 1.0, ur - 1.0) < tol and rel_err(v, vr) < tol
    dtol = 1e-13 if prec == VX_PREC_F64 else 1e-4
    assert np.max(np.abs(d - dr) / dr) < dtol


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 200, 1000, 5000])
@pytest.mark.parametrize("prec", [VX_PREC_F64, VX_PREC_MIXED])
@pytest.mark.parametrize("exact", [0, 1])
def test_tree_small_sizes_vs_direct_oracle(ref, n, prec, exact):
    """Trees of one bucket, of a few buckets with no node above the cut
    nodes, and with partial batches of far nodes: velocities and nearest-3
    distances against the FP64 direct sums, with and without the exact
    nearest-3 search forced for every target."""
    rng = np.random.default_rng(100 + n)
    if n >= 1000:
        x, y, g = wake_cloud(n, 5)
    else:
        x, y = rng.uniform(-1.0, 1.0, n), rng.uniform(-1.0, 1.0, n)
        g = rng.uniform(-0.004, 0.004, n)
    s = (n + 31) // 32 * 32
    X, Y, G = tail(x, y, g, s)
    eps = 0.004
    ur, vr, dr = ref.biot_savart(X, Y, G, eps * eps, (1.0, 0.0))
    pan = geo.circle(200, 0.5)
    c = Case(make_params(pan, precision=prec, ve_size=eps, flags=exact), pan)
    c.set_state(X, Y, G)
    c.tree_biot_savart(s, (1.0, 0.0))
    u, v, d = c.velocity(s)
    tol = 2e-4 if prec == VX_PREC_F64 else 3e-4
    scale = max(np.max(np.abs(ur[:n] - 1.0)), np.max(np.abs(vr[:n])), 1e-12)
    assert np.max(np.abs(u[:n] - ur[:n])) <= tol * scale + 1e-12
    assert np.max(np.abs(v[:n] - vr[:n])) <= tol * scale + 1e-12
    if n >= 4:  # the three nearest lie among the vortices, not the far tail
        dtol = 1e-13 if (prec == VX_PREC_F64 or exact) else 1e-4
        assert np.max(np.abs(d[:n] - dr[:n]) / dr[:n]) < dtol


@pytest.mark.parametrize("prec", [VX_PREC_F64, VX_PREC_MIXED])
def test_exact_nearest_search_for_every_target(cloud40k, prec):
    """The exact nearest-3 search (otherwise run only where the near field
    cannot guarantee the answer) forced on for every target."""
    from vortex_amd._lib import VX_PREC_F64 as F64
    X, Y, G, s, eps, ur, vr, dr = cloud40k
    pan = geo.circle(200, 0.5)
    c = Case(make_params(pan, precision=prec, ve_size=eps, flags=1), pan)
    c.set_state(X, Y, G)
    c.tree_biot_savart(s, (1.0, 0.0))
    _, _, d = c.velocity(s)
    assert c.tree_stats()["fallback"] == s
    assert np.max(np.abs(d - dr) / dr) < 1e-13  # the exact search runs in FP64 for both paths
    _ = F64


@pytest.mark.parametrize("prec,tol", [(VX_PREC_F64, 1e-9), (VX_PREC_MIXED, 1e-3)])
def test_tree_diffusion_vs_direct_oracle(ref, prec, tol):
    n = 20000
    x, y, g = wake_cloud(n, 2)
    s = (n + 31) // 32 * 32
    X, Y, G = tail(x, y, g, s)
    eps, nu = 0.004, 1e-3
    pan = geo.circle(200, 0.5)
    c = Case(make_params(pan, precision=prec, ve_size=eps, nu=nu, order=32), pan)
    c.set_state(X, Y, G)
    c.tree_biot_savart(s, (0.0, 0.0))
    u0, v0, d = c.velocity(s)
    c.tree_diffusion()
    u, v, _ = c.velocity(s)
    ur, vr = ref.diffusion(X, Y, G, d, nu, eps, u0, v0)
    assert rel_err(u - u0, ur - u0) < tol and rel_err(v - v0, vr - v0) < tol
    info = c.tree_stats()
    assert 0 < info["diff_pairs"] < s * s / 4  # the cut-off prunes
    assert info["heavy"] > 0  # isolated vortices go through the split direct kernel


@pytest.mark.parametrize("prec,tol", [(VX_PREC_F64, 1e-9), (VX_PREC_MIXED, 1e-3)])
def test_tree_diffusion_after_the_caller_sets_d(ref, prec, tol):
    """The Biot-Savart pass leaves a copy of d in the tree's sorted order for
    the diffusion's far-target selection; a d handed in by the caller
    afterwards (vx_case_set_velocity) must replace that copy, so the
    diffusion sees the caller's d everywhere, including which targets reach
    the domain scale."""
    n = 20000
    x, y, g = wake_cloud(n, 4)
    s = (n + 31) // 32 * 32
    X, Y, G = tail(x, y, g, s)
    eps, nu = 0.004, 1e-3
    pan = geo.circle(200, 0.5)
    c = Case(make_params(pan, precision=prec, ve_size=eps, nu=nu, order=32), pan)
    c.set_state(X, Y, G)
    c.tree_biot_savart(s, (0.0, 0.0))
    u0, v0, d = c.velocity(s)
    d2 = d * np.random.default_rng(5).uniform(0.5, 3.0, s)
    c.set_velocity(u0, v0, d2)
    c.tree_diffusion()
    u, v, _ = c.velocity(s)
    ur, vr = ref.diffusion(X, Y, G, d2, nu, eps, u0, v0)
    assert rel_err(u - u0, ur - u0) < tol and rel_err(v - v0, vr - v0) < tol
    c.close()


@pytest.mark.parametrize("prec", [VX_PREC_F64, VX_PREC_MIXED])
def test_tree_diffusion_census_counts_in_radius_pairs(ref, prec):
    """The census (vx_tree_census) leaves the velocities bit for bit and
    counts, of the evaluated pairs, those within each target's cut-off
    radius; on a strip dense enough that no target's radius reaches 5 % of
    the domain (none goes to the domain-scale kernel), that count equals the
    brute-force count of (target, source) pairs with r < cut d."""
    n = 8000
    rng = np.random.default_rng(12)
    x, y = rng.uniform(0.0, 40.0, n), rng.uniform(0.0, 0.05, n)
    g = rng.uniform(-0.004, 0.004, n)
    s = (n + 31) // 32 * 32
    X, Y, G = tail(x, y, g, s)
    eps, nu = 0.004, 1e-3
    pan = geo.circle(64, 0.05)
    out = []
    for census in (False, True):
        c = Case(make_params(pan, precision=prec, ve_size=eps, nu=nu), pan)
        c.set_state(X, Y, G)
        c.tree_census(census)
        c.tree_biot_savart(s, (0.0, 0.0))
        _, _, d = c.velocity(s)
        c.tree_diffusion()
        u, v, _ = c.velocity(s)
        out.append((u, v, d, c.tree_stats()))
        c.close()
    (u0, v0, d, i0), (u1, v1, _, i1) = out
    assert np.array_equal(u0, u1) and np.array_equal(v0, v1)
    assert i0["diff_in_radius"] == -1 and i1["heavy"] == 0
    cut = 18.0 if prec == VX_PREC_MIXED else 30.0
    m = i1["m"]
    within = 0
    for a in range(0, n, 1000):
        dx = x[a:a + 1000, None] - x[None, :]
        dy = y[a:a + 1000, None] - y[None, :]
        within += int((dx * dx + dy * dy < (cut * d[a:a + 1000, None]) ** 2).sum())
    print(f"[census] prec={prec}: evaluated {i1['diff_pairs']} in radius {i1['diff_in_radius']} "
          f"brute force {within} (m={m})", flush=True)
    # FP32 distances may move pairs right at the radius across it
    assert abs(i1["diff_in_radius"] - within) <= 1e-4 * within
    assert within <= i1["diff_pairs"]


@pytest.mark.parametrize("prec,tol", [(VX_PREC_F64, 1e-9), (VX_PREC_MIXED, 1e-3)])
def test_tree_diffusion_far_list_grows_past_its_first_capacity(ref, prec, tol):
    """A sparse cloud in which most targets reach beyond 5 % of the domain:
    the first diffusion lists 4096 of them and leaves the rest in their
    groups' waves; the next one sizes its list from that count and lists
    more.  Both match the direct sums."""
    n = 20000
    rng = np.random.default_rng(11)
    x, y = rng.uniform(-1.5, 9.5, n), rng.uniform(-9.0, 9.0, n)
    g = rng.uniform(0.001, 0.004, n)  # one sign: no cancellation in the weight sums
    s = (n + 31) // 32 * 32
    X, Y, G = tail(x, y, g, s)
    eps, nu = 0.004, 1e-3
    pan = geo.circle(200, 0.5)
    c = Case(make_params(pan, precision=prec, ve_size=eps, nu=nu, order=32), pan)
    c.set_state(X, Y, G)
    heavy = []
    for _ in range(2):
        c.tree_biot_savart(s, (0.0, 0.0))
        u0, v0, d = c.velocity(s)
        c.tree_diffusion()
        u, v, _ = c.velocity(s)
        ur, vr = ref.diffusion(X, Y, G, d, nu, eps, u0, v0)
        assert rel_err(u - u0, ur - u0) < tol and rel_err(v - v0, vr - v0) < tol
        heavy.append(c.tree_stats()["heavy"])
    assert heavy[0] == 4096 and heavy[1] > 4096


@pytest.mark.parametrize("prec", [VX_PREC_F64, VX_PREC_MIXED])
def test_tree_diffusion_overflowing_far_list_is_reproducible(prec):
    """The first diffusion of a sparse cloud lists only the first 4096 of its
    domain-scale targets (in sorted order) and leaves the rest in their
    groups' waves; which targets take which path, and so every bit of u and
    v, must not depend on scheduling.  Three fresh cases on the same state
    agree bit for bit."""
    n = 20000
    rng = np.random.default_rng(11)
    x, y = rng.uniform(-1.5, 9.5, n), rng.uniform(-9.0, 9.0, n)
    g = rng.uniform(0.001, 0.004, n)
    s = (n + 31) // 32 * 32
    X, Y, G = tail(x, y, g, s)
    pan = geo.circle(200, 0.5)
    outs = []
    for _ in range(3):
        c = Case(make_params(pan, precision=prec, ve_size=0.004, nu=1e-3), pan)
        c.set_state(X, Y, G)
        c.tree_biot_savart(s, (0.0, 0.0))
        c.tree_diffusion()
        assert c.tree_stats()["heavy"] == 4096  # the list overflowed
        u, v, d = c.velocity(s)
        outs.append((u.copy(), v.copy()))
        c.close()
    for u, v in outs[1:]:
        assert np.array_equal(u, outs[0][0]) and np.array_equal(v, outs[0][1])


def test_tree_and_direct_speed_agree_inside_a_time_step(ref):
    """Whole time steps with the tree switched on from 2000 vortices against
    the direct path on the same start state (FP64)."""
    pan = geo.circle(200, 0.5)
    out = {}
    for tmin in (0, 2000):
        prm = make_params(pan, dt=0.01, r_col_same_sign=0.3, r_col_diff_sign=0.3, birth_offset_rel=0.5,
                          precision=VX_PREC_F64, tree_min_n=tmin, order=32)
        c = Case(prm, pan)
        c.build_matrix()
        x, y, g = wake_cloud(6000, 3)
        c.set_state(x, y, g)
        fs = [c.time_step((1.0, 0.0)) for _ in range(2)]
        out[tmin] = (c.n, c.state(), fs)
    (n0, (x0, y0, g0), f0), (n1, (x1, y1, g1), f1) = out[0], out[2000]
    assert abs(n0 - n1) <= 2
    k = min(n0, n1)
    assert np.median(np.abs(x0[:k] - x1[:k])) < 1e-9
    assert abs(f0[-1][1] - f1[-1][1]) < 1e-5 * max(1.0, abs(f0[-1][1]))


@pytest.mark.parametrize("prec,order", [(VX_PREC_MIXED, 16), (VX_PREC_F64, 32)])
def test_fused_time_step_matches_the_stage_calls(prec, order):
    """vx_time_step builds the tree on a second stream while the birth stage
    runs (the newborns' positions taken from the case's birth points); the
    three stage calls run one after the other.  Both must give the same state
    and forces, bit for bit, over several steps."""
    pan = geo.circle(300, 0.5)
    prm = make_params(pan, dt=0.01, r_col_same_sign=0.3, r_col_diff_sign=0.3, birth_offset_rel=0.5,
                      precision=prec, tree_min_n=2000, order=order)
    x, y, g = wake_cloud(20000, 5)
    res = []
    for fused in (True, False):
        c = Case(prm, pan)
        c.build_matrix()
        c.set_state(x, y, g)
        fs = []
        for _ in range(3):
            if fused:
                fs.append(c.time_step((1.0, 0.0)))
            else:
                c.vort_creation((1.0, 0.0))
                c.speed((1.0, 0.0))
                fs.append(c.step())
        res.append((c.n, c.state(), np.array(fs)))
        c.close()
    (n0, s0, f0), (n1, s1, f1) = res
    assert n0 == n1
    for a, b in zip(s0, s1):
        assert np.array_equal(a, b)
    assert np.array_equal(f0, f1)


@pytest.mark.parametrize("prec,r_same", [(VX_PREC_MIXED, 0.6), (VX_PREC_F64, 0.6), (VX_PREC_MIXED, 1.0)])
def test_step_collapse_kept_grid_matches_rebuilt_grids(prec, r_same):
    """Whole time steps (tree velocities) of a cloud with thousands of pairs
    within the collapse radii, three passes per kind, each merging and filling
    holes: with the collapse grid kept from pass to pass (col_share_min = 1;
    the survivors and the vortices moved into holes take pool nodes), one per
    kind (radii 1.0 and 0.6) or one for both (equal radii), the state ends bit
    for bit as with a grid built for every pass."""
    pan = geo.circle(300, 0.5)
    eps = 0.004
    prm = make_params(pan, dt=0.01, ve_size=eps, r_col_diff_sign=1.0, r_col_same_sign=r_same, max_ve_g=0.006,
                      birth_offset_rel=0.5, n_col=3, precision=prec, tree_min_n=2000)
    rng = np.random.default_rng(21)
    n = 30000
    x, y = rng.uniform(0.6, 3.0, n), rng.uniform(-0.5, 0.5, n)
    g = rng.uniform(-0.004, 0.004, n)
    res = []
    for share in (1, 0):
        c = Case(prm, pan)
        c.build_matrix()
        c.set_tuning(col_share_min=share)
        c.set_state(x, y, g)
        merged = 0
        for _ in range(2):
            before = c.n + pan.size  # the births come first
            c.time_step((1.0, 0.0))
            merged += before - c.n
        res.append((c.n, c.state(), merged))
        c.close()
    (n0, s0, m0), (n1, s1, m1) = res
    print(f"[collapse grid] prec={prec} r_same={r_same}: n {n0} {n1}, removed {m0} {m1}", flush=True)
    assert m0 > 2000  # absorbed, deleted and merged; mostly merged here
    assert n0 == n1
    for a, b in zip(s0, s1):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("prec,order", [(VX_PREC_MIXED, 16), (VX_PREC_F64, 32)])
def test_speed_stage_matches_the_separate_kernels(prec, order):
    """vx_speed runs the wall term and the domain-scale diffusion targets on
    a second stream beside the groups' diffusion; the separate entry points
    run everything in order on one stream.  Same u, v, d bit for bit."""
    pan = geo.circle(300, 0.5)
    prm = make_params(pan, dt=0.01, birth_offset_rel=0.5, precision=prec, tree_min_n=2000, order=order)
    x, y, g = wake_cloud(20000, 7)
    s = (x.size + 31) // 32 * 32
    out = []
    for split in (False, True):
        c = Case(prm, pan)
        c.set_state(x, y, g)
        if split:
            c.tree_biot_savart(s, (1.0, 0.0))
            c.tree_diffusion()
            c.diffusion_wall(s)
        else:
            c.speed((1.0, 0.0))
        assert c.tree_stats()["fallback"] > 0  # the isolated vortices take the exact search
        out.append(c.velocity(x.size))
        c.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_tree_one_million(ref):
    """N = 2^20 uniform wake: mixed tree against the FP64 tree with 32 terms,
    and the FP64 tree against the oracle's direct sum on a target sample."""
    n = 1 << 20
    x, y, g = uniform_wake(n, seed=1)
    eps = math.pi / 1000
    pan = geo.circle(1000, 0.5)
    res = {}
    for prec, order in ((VX_PREC_F64, 32), (VX_PREC_MIXED, 16)):
        c = Case(make_params(pan, precision=prec, ve_size=eps, order=order), pan)
        c.set_state(x, y, g)
        c.tree_biot_savart(n, (0.0, 0.0))
        res[prec] = c.velocity(n)
        c.close()
    u, v, d = res[VX_PREC_F64]
    um, vm, dm = res[VX_PREC_MIXED]
    assert rel_err(um, u) < 3e-4 and rel_err(vm, v) < 3e-4
    assert np.max(np.abs(dm - d) / d) < 1e-4
    sample = 512
    ur, vr, dr = ref.biot_savart(x, y, g, eps * eps, (0.0, 0.0), n_tgt=sample)
    assert rel_err(u[:sample], ur) < 1e-8 and rel_err(v[:sample], vr) < 1e-8
    assert np.max(np.abs(d[:sample] - dr) / dr) < 1e-13


def test_step_probes_over_the_tree_match_direct_probes():
    """A time step's probe velocities (the reference's velocity_control, 500
    points as src/main.c:521-526 places them, plus a line through the wake)
    are evaluated over the step's tree when the step uses it.  The birth does
    not depend on the velocity path, so every path sees the same state at the
    probes: the tree (FP64 and mixed) and the mixed direct sums (FP64
    separations, FP32 kernel) agree with the FP64 direct sums to 2e-6 of the
    largest probe velocity (measured: FP64 tree 1.5e-8)."""
    from vortex_amd.state import uniform_wake
    pan = geo.circle(1000, 0.5)
    eps = float(pan["length"][0])
    x, y, g = uniform_wake(200000, seed=7)
    px = np.concatenate([np.full(500, -0.15), np.linspace(0.6, 9.5, 300)])
    py = np.concatenate([0.01 + 0.002 * np.arange(500), np.linspace(-1.5, 1.5, 300)])
    out = {}
    for prec in (VX_PREC_F64, VX_PREC_MIXED):
        for tmin in (65536, 0):
            c = Case(make_params(pan, dt=0.005, ve_size=eps, precision=prec, tree_min_n=tmin), pan)
            c.build_matrix()
            c.set_state(x, y, g)
            c.set_probes(px, py)
            c.time_step((1.0, 0.0))
            out[prec, tmin] = c.probes()
            c.close()
    ud, vd = out[VX_PREC_F64, 0]
    scale = max(np.max(np.abs(ud)), np.max(np.abs(vd)))
    for key in ((VX_PREC_F64, 65536), (VX_PREC_MIXED, 65536), (VX_PREC_MIXED, 0)):
        ut, vt = out[key]
        err = max(np.max(np.abs(ut - ud)), np.max(np.abs(vt - vd))) / scale
        print(f"[tree] probes prec={key[0]} tree={key[1] > 0}: rel err {err:.3g}", flush=True)
        assert err < 2e-6, key
