"""Pins of the oracle's panel system (A1 matrix, A2 inverse, A3 RHS, A4 birth)
against quadrature, closed forms and numpy.  CPU only."""
import math
import os

import numpy as np
import pytest
from scipy import integrate

from vortex_amd import geometry as geo

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def point_vortex_vel(px, py, sx, sy):
    """Velocity at (px,py) of a unit point vortex at (sx,sy), counter-clockwise."""
    dx, dy = px - sx, py - sy
    r2 = dx * dx + dy * dy
    return -dy / (2 * math.pi * r2), dx / (2 * math.pi * r2)


def sheet_avg_quadrature(pan, i, j, npts=48):
    """tau_i . (1/l_i) int_i int_j K ds dt for a unit-strength sheet on panel j."""
    xg, wg = np.polynomial.legendre.leggauss(npts)
    s = 0.5 * (xg + 1.0); w = 0.5 * wg
    a0, a1 = pan["left"][i], pan["right"][i]
    b0, b1 = pan["left"][j], pan["right"][j]
    P = a0[None, :] + s[:, None] * (a1 - a0)[None, :]
    Q = b0[None, :] + s[:, None] * (b1 - b0)[None, :]
    dx = P[:, None, 0] - Q[None, :, 0]
    dy = P[:, None, 1] - Q[None, :, 1]
    r2 = dx * dx + dy * dy
    ww = w[:, None] * w[None, :]
    u = np.sum(ww * (-dy) / (2 * math.pi * r2)) * pan["length"][j]
    v = np.sum(ww * dx / (2 * math.pi * r2)) * pan["length"][j]
    return u * pan["tang"][i][0] + v * pan["tang"][i][1]


def test_panel_layout_matches_reference_tpanel(ref):
    assert ref.lib().vxr_sizeof_panel() == 120 == geo.PANEL_DTYPE.itemsize


@pytest.mark.parametrize("maker", [lambda: geo.circle(40, 0.5), lambda: geo.naca00(40, 0.12, alpha_deg=8.0),
                                   lambda: geo.flat_plate(48, thickness=0.1, alpha_deg=15.0)])
def test_matrix_far_entries_vs_gauss_legendre(ref, maker):
    pan = maker()
    P = pan.size
    L = ref.panel_matrix(pan)
    rng = np.random.default_rng(0)
    checked = 0
    for _ in range(60):
        i, j = rng.integers(0, P, 2)
        if min((i - j) % P, (j - i) % P) < 3:
            continue
        q = sheet_avg_quadrature(pan, i, j)
        assert abs(L[i, j] - q) <= 1e-9 + 1e-7 * abs(q), (i, j, L[i, j], q)
        checked += 1
    assert checked > 20


def test_matrix_adjacent_entries_vs_adaptive_quadrature(ref):
    pan = geo.circle(24, 0.5)
    L = ref.panel_matrix(pan)
    for (i, j) in [(3, 4), (4, 3), (0, 23), (23, 0), (10, 11)]:
        a0, a1 = pan["left"][i], pan["right"][i]
        b0, b1 = pan["left"][j], pan["right"][j]
        tau = pan["tang"][i]

        def f(t, s):
            p = a0 + s * (a1 - a0); q = b0 + t * (b1 - b0)
            u, v = point_vortex_vel(p[0], p[1], q[0], q[1])
            return (u * tau[0] + v * tau[1]) * pan["length"][j]

        q, _ = integrate.dblquad(f, 0.0, 1.0, 0.0, 1.0, epsabs=1e-12, epsrel=1e-10)
        assert abs(L[i, j] - q) < 1e-7, (i, j, L[i, j], q)


def test_matrix_structure(ref):
    pan = geo.circle(16)
    L = ref.panel_matrix(pan)
    P = 16
    assert np.all(np.diag(L)[:P] == -0.5)
    assert np.allclose(L[P, :P], pan["length"]) and np.all(L[:P, P] == 1.0) and L[P, P] == 0.0


@pytest.mark.parametrize("mode", [0, 1])
def test_invert_matches_numpy_and_is_inverse(ref, mode):
    pan = geo.circle(300, 0.5)
    L = ref.panel_matrix(pan)
    M = ref.invert(L, mode)
    assert np.allclose(M @ L, np.eye(L.shape[0]), atol=1e-11)
    Mn = np.linalg.inv(L)
    assert np.max(np.abs(M - Mn)) <= 1e-10 * np.max(np.abs(Mn))


def test_reference_pivot_rule_breaks_down_on_closed_airfoil(ref):
    """The look-ahead pivot of src/unit_1.cu:182-201 picks the pivot of column
    k+1 before eliminating column k; on a closed body the P x P block is
    singular and the rule hits a pivot below 1e-12 (reference exits), or
    returns an inaccurate inverse.  Standard partial pivoting is accurate."""
    L = ref.panel_matrix(geo.naca00(120, 0.12, alpha_deg=8.0))
    with pytest.raises(np.linalg.LinAlgError):
        ref.invert(L, 0)
    M = ref.invert(L, 1)
    assert np.allclose(M @ L, np.eye(L.shape[0]), atol=1e-10)
    Lp = ref.panel_matrix(geo.flat_plate(1000, alpha_deg=15.0))
    bad = ref.invert(Lp, 0)
    assert np.abs(bad @ Lp - np.eye(1001)).max() > 1e-3
    good = ref.invert(Lp, 1)
    assert np.abs(good @ Lp - np.eye(1001)).max() < 1e-9


def test_invert_singular_matrix_reports_error(ref):
    A = np.ones((5, 5))
    with pytest.raises(np.linalg.LinAlgError):
        ref.invert(A)


def test_invert_needs_first_pivot_swap(ref):
    A = np.array([[0.0, 2.0, 1.0], [3.0, 1.0, 0.0], [1.0, 0.0, 4.0]])
    M = ref.invert(A)
    assert np.allclose(M, np.linalg.inv(A), atol=1e-14)


def test_circle_sheet_closed_form_and_constraint(ref):
    """Uniform flow U past a circle: |gamma(theta)| = 2U|sin theta|, sum gamma l = d_g."""
    P, U = 200, 1.0
    pan = geo.circle(P, 0.5)
    M = ref.invert(ref.panel_matrix(pan))
    Rp = ref.right_part(np.zeros(0), np.zeros(0), np.zeros(0), pan, (U, 0.0))
    th = np.arctan2(pan["contr"][:, 1], pan["contr"][:, 0])
    _, _, gnew, gs = ref.birth(M, Rp, 0.0, pan)
    assert np.max(np.abs(np.abs(gs) - 2 * U * np.abs(np.sin(th)))) < 0.02
    # opposite signs on the upper and lower halves
    assert np.all(np.sign(gs[np.sin(th) > 0.1]) == -np.sign(gs[np.sin(th) < -0.1]).mean())
    for dg in (0.0, 0.37, -1.5):
        _, _, gnew, gs = ref.birth(M, Rp, dg, pan)
        assert abs(np.sum(gs * pan["length"]) - dg) < 1e-12
        assert abs(np.sum(gnew) - dg) < 1e-12  # newborns carry exactly d_g


def test_circle_solution_golden(ref):
    """Sheet strengths against the committed numpy.linalg.solve fixture."""
    data = np.load(os.path.join(GOLD, "circle200_sheet.npz"))
    pan = geo.circle(200, 0.5)
    M = ref.invert(ref.panel_matrix(pan))
    Rp = ref.right_part(np.zeros(0), np.zeros(0), np.zeros(0), pan, (1.0, 0.0))
    _, _, _, gs = ref.birth(M, Rp, 0.0, pan)
    assert np.max(np.abs(gs - data["g_sheet"])) < 1e-10 * np.max(np.abs(data["g_sheet"]))


def test_right_part_vs_adaptive_quadrature(ref):
    pan = geo.circle(30, 0.5)
    rng = np.random.default_rng(3)
    n = 7
    ang = rng.uniform(0, 2 * np.pi, n); rad = rng.uniform(0.55, 2.0, n)
    x, y, g = rad * np.cos(ang), rad * np.sin(ang), rng.uniform(-0.01, 0.01, n)
    vinf = (0.8, -0.3)
    R = ref.right_part(x, y, g, pan, vinf)
    for i in (0, 7, 19):
        a0, a1 = pan["left"][i], pan["right"][i]
        uu = vv = 0.0
        for k in range(n):
            fu = lambda s: point_vortex_vel(*(a0 + s * (a1 - a0)), x[k], y[k])[0]
            fv = lambda s: point_vortex_vel(*(a0 + s * (a1 - a0)), x[k], y[k])[1]
            uu += g[k] * integrate.quad(fu, 0, 1, epsabs=1e-13)[0]
            vv += g[k] * integrate.quad(fv, 0, 1, epsabs=1e-13)[0]
        t = pan["tang"][i]
        expect = -((uu + vinf[0]) * t[0] + (vv + vinf[1]) * t[1])
        assert abs(R[i] - expect) < 1e-11, (i, R[i], expect)


def test_birth_positions_and_pairing(ref):
    pan = geo.circle(12, 0.5)
    M = ref.invert(ref.panel_matrix(pan))
    Rp = np.linspace(-1, 1, 12)
    xo, yo, go, gs = ref.birth(M, Rp, 0.1, pan)
    mid = 0.5 * (pan["left"] + pan["right"])
    assert np.allclose(xo, mid[:, 0] + 1e-6 * pan["norm"][:, 0], atol=0, rtol=1e-15)
    nxt = pan["n_of_rpanel"].astype(int)
    assert np.allclose(go, 0.5 * (gs * pan["length"] + gs[nxt] * pan["length"][nxt]), rtol=1e-14)
    xo2, _, _, _ = ref.birth(M, Rp, 0.1, pan, birth_offset_rel=0.5)
    assert np.allclose(xo2, mid[:, 0] + 0.5 * pan["length"] * pan["norm"][:, 0], rtol=1e-15)
