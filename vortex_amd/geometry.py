"""Panel geometries in the reference's profile format.

A profile file (read by ``load_profile``, src/main.c:220-256) has two header
lines, the panel count P, then P records of 16 columns::

    n  left.x left.y  right.x right.y  contr.x contr.y  birth.x birth.y
       norm.x norm.y  tang.x tang.y  length  n_of_lpanel n_of_rpanel

Panels run clockwise with outward normals.  The tangent column is used
verbatim by the panel matrix and the birth right-hand side, so it is neither
renormalised nor reoriented here.  ``PANEL_DTYPE`` is byte-identical to the
reference ``tPanel`` (include/definitions.h:92-113, 120 bytes), so arrays of
it can be handed to the C-ABI as ``tPanel*``.

The generators build the bodies of the benchmark cases: a circle, a thin
ellipse with cosine clustering ("flat plate"), and the symmetric 4-digit
thickness polynomial airfoil; ``rotate`` follows turn.py:13,33-41
(counter-clockwise rotation of every vector column about the origin).
"""
from __future__ import annotations

import math
import numpy as np

PANEL_DTYPE = np.dtype({
    "names": ["n", "left", "right", "contr", "birth", "norm", "tang", "length",
              "n_of_lpanel", "n_of_rpanel"],
    "formats": ["<u4", ("<f8", 2), ("<f8", 2), ("<f8", 2), ("<f8", 2), ("<f8", 2),
                ("<f8", 2), "<f8", "<u4", "<u4"],
    "offsets": [0, 8, 24, 40, 56, 72, 88, 104, 112, 116],
    "itemsize": 120,
})

BIRTH_FILE_OFFSET = 1e-4  # birth column of the shipped circle profile: left end + 1e-4 normal


def panels_from_vertices(vx: np.ndarray, vy: np.ndarray) -> np.ndarray:
    """Closed clockwise polygon (vertex k -> k+1) to a panel array."""
    vx = np.asarray(vx, dtype=np.float64)
    vy = np.asarray(vy, dtype=np.float64)
    P = vx.size
    pan = np.zeros(P, dtype=PANEL_DTYPE)
    lx, ly = vx, vy
    rx, ry = np.roll(vx, -1), np.roll(vy, -1)
    dx, dy = rx - lx, ry - ly
    ln = np.hypot(dx, dy)
    tx, ty = dx / ln, dy / ln
    nx, ny = -ty, tx  # clockwise traversal: outward normal is the +90 deg turn
    pan["n"] = np.arange(P, dtype=np.uint32)
    pan["left"][:, 0], pan["left"][:, 1] = lx, ly
    pan["right"][:, 0], pan["right"][:, 1] = rx, ry
    pan["contr"][:, 0], pan["contr"][:, 1] = (lx + rx) / 2, (ly + ry) / 2
    pan["birth"][:, 0] = lx + BIRTH_FILE_OFFSET * nx
    pan["birth"][:, 1] = ly + BIRTH_FILE_OFFSET * ny
    pan["norm"][:, 0], pan["norm"][:, 1] = nx, ny
    pan["tang"][:, 0], pan["tang"][:, 1] = -tx, -ty
    pan["length"] = ln
    idx = np.arange(P)
    pan["n_of_lpanel"] = (idx - 1) % P
    pan["n_of_rpanel"] = (idx + 1) % P
    return pan


def _orient_clockwise(vx, vy):
    area = 0.5 * np.sum(vx * np.roll(vy, -1) - np.roll(vx, -1) * vy)
    if area > 0:
        return vx[::-1].copy(), vy[::-1].copy()
    return vx, vy


def circle(P: int, radius: float = 0.5, cx: float = 0.0, cy: float = 0.0) -> np.ndarray:
    """P equal straight panels on a circle, starting at the leftmost point."""
    th = math.pi - 2.0 * math.pi * np.arange(P) / P
    return panels_from_vertices(cx + radius * np.cos(th), cy + radius * np.sin(th))


def flat_plate(P: int, chord: float = 1.0, thickness: float = 0.02,
               alpha_deg: float = 0.0) -> np.ndarray:
    """Thin elliptic plate; uniform ellipse angle = cosine clustering to both edges."""
    th = math.pi - 2.0 * math.pi * np.arange(P) / P
    vx = 0.5 * chord * np.cos(th)
    vy = 0.5 * thickness * np.sin(th)
    pan = panels_from_vertices(vx, vy)
    return rotate(pan, -alpha_deg) if alpha_deg else pan


def naca00(P: int, t: float = 0.12, chord: float = 1.0, alpha_deg: float = 0.0) -> np.ndarray:
    """Symmetric 4-digit thickness polynomial, closed trailing edge, cosine spacing.

    P must be even.  The body spans x in [0, chord] before rotation and is
    rotated about the quarter chord."""
    if P % 2:
        raise ValueError("naca00 needs an even panel count")
    h = P // 2
    beta = np.linspace(0.0, math.pi, h + 1)
    xc = 0.5 * (1.0 - np.cos(beta))
    yt = 5.0 * t * (0.2969 * np.sqrt(xc) - 0.1260 * xc - 0.3516 * xc ** 2
                    + 0.2843 * xc ** 3 - 0.1036 * xc ** 4)
    # clockwise: leading edge -> upper surface -> trailing edge -> lower surface
    ux, uy = xc, yt
    lx, ly = xc[::-1][1:-1], -yt[::-1][1:-1]
    vx = np.concatenate([ux, lx]) * chord
    vy = np.concatenate([uy, ly]) * chord
    vx, vy = _orient_clockwise(vx, vy)
    pan = panels_from_vertices(vx - 0.25 * chord, vy)
    return rotate(pan, -alpha_deg) if alpha_deg else pan


def rotate(pan: np.ndarray, ang_deg: float) -> np.ndarray:
    """turn.py:13,33-41: rotate every vector column counter-clockwise by ang_deg."""
    a = math.radians(ang_deg)
    R = np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]])
    out = pan.copy()
    for f in ("left", "right", "contr", "birth", "norm", "tang"):
        out[f] = pan[f] @ R.T
    return out


def profile_bbox(pan: np.ndarray):
    """Profile box from the left end points (src/main.c:235-254)."""
    lx, ly = pan["left"][:, 0], pan["left"][:, 1]
    return float(lx.min()), float(lx.max()), float(ly.min()), float(ly.max())


def write_profile(path: str, pan: np.ndarray) -> None:
    """Write the 16-column profile format (two header lines, count, records)."""
    with open(path, "w", encoding="cp1251") as f:
        f.write("--profile-for-vortex--\n--number-of-panels--\n")
        f.write(f"{pan.size}\n")
        for p in pan:
            f.write("%u %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %u %u\n" % (
                p["n"], *p["left"], *p["right"], *p["contr"], *p["birth"], *p["norm"],
                *p["tang"], p["length"], p["n_of_lpanel"], p["n_of_rpanel"]))


def read_profile(path: str) -> np.ndarray:
    """Read a profile file; trailer lines after the P records are ignored."""
    with open(path, "rb") as f:
        lines = f.read().decode("cp1251", errors="replace").splitlines()
    P = int(lines[2].split()[0])
    pan = np.zeros(P, dtype=PANEL_DTYPE)
    rows = []
    k = 3
    while len(rows) < P:
        tok = lines[k].split()
        k += 1
        if not tok:
            continue
        rows.append(tok)
    for i, tok in enumerate(rows):
        v = [float(t) for t in tok[:16]]
        pan[i]["n"] = int(v[0])
        pan[i]["left"] = v[1:3]
        pan[i]["right"] = v[3:5]
        pan[i]["contr"] = v[5:7]
        pan[i]["birth"] = v[7:9]
        pan[i]["norm"] = v[9:11]
        pan[i]["tang"] = v[11:13]
        pan[i]["length"] = v[13]
        pan[i]["n_of_lpanel"] = int(v[14])
        pan[i]["n_of_rpanel"] = int(v[15])
    return pan
