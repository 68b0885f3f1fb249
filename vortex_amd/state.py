"""Host-side vortex state helpers: the far tail of unused capacity and the
synthetic benchmark states.

Unused capacity slots hold g = 0 at far-away positions, as the reference's
``zero_Kernel`` does (src/kernel.cu:767-777: r = 2e5 + 2e5 * rand()).  Here the
"random" factor is a fixed 31-bit hash of the slot index (splitmix64), so the
tail is reproducible and identical on the host and on the device
(``vx_fill_tail`` in vortex_amd/csrc/vx_state.hip uses the same hash).
"""
from __future__ import annotations

import numpy as np

TAIL_SEED = 12345
_M64 = (1 << 64) - 1


def splitmix64(z: np.ndarray) -> np.ndarray:
    z = z.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def far_tail(start: int, count: int, seed: int = TAIL_SEED):
    """Positions of capacity slots [start, start+count): x = y = 2e5 + 2e5*h."""
    k = np.arange(start, start + count, dtype=np.uint64)
    h = (splitmix64(k ^ np.uint64(seed)) >> np.uint64(33)).astype(np.float64)
    x = 2.0e5 * h + 2.0e5
    return x, x.copy()


def uniform_wake(n: int, seed: int = 1, max_ve_g: float = 0.00628162,
                 box=(-0.6, 10.0, -2.0, 2.0), body_radius: float = 0.5):
    """S1M-uniform synthetic state: positions uniform in ``box`` minus the disk
    r < body_radius, circulations +-U(0, max_ve_g) with alternating sign."""
    rng = np.random.default_rng(seed)
    xs, ys = [], []
    got = 0
    while got < n:
        m = int((n - got) * 1.1) + 64
        x = rng.uniform(box[0], box[1], m)
        y = rng.uniform(box[2], box[3], m)
        keep = x * x + y * y >= body_radius * body_radius
        xs.append(x[keep]); ys.append(y[keep])
        got += int(keep.sum())
    x = np.concatenate(xs)[:n]
    y = np.concatenate(ys)[:n]
    mag = rng.uniform(0.0, max_ve_g, n)
    sign = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    return x, y, mag * sign


def jittered_wake(n: int, seed: int = 1, max_ve_g: float = 0.00628162, min_sep: float = 0.002,
                  box=(-0.6, 10.0, -2.0, 2.0), body_radius: float = 0.5, layers: bool = False):
    """A wake of exactly n vortices that holds its size: one vortex per cell
    of a square grid over ``box`` minus the disk r < body_radius + min_sep,
    jittered inside its cell so that any two stay at least ``min_sep`` apart
    (above the collapse radii, so the first steps do not merge thousands of
    pairs as a uniform random cloud does).  Circulations: U(0, max_ve_g) in
    magnitude, alternating in sign over a random order of the cells; with
    ``layers`` instead U(max_ve_g / 3, max_ve_g), negative above y = 0 and
    positive below, as in the two shear layers behind a body in a flow along
    x (random signs make the weights of the diffusive velocity cancel, and the
    ratio of its sums then throws vortices out of the box)."""
    rng = np.random.default_rng(seed)
    area = (box[1] - box[0]) * (box[3] - box[2]) - np.pi * (body_radius + min_sep) ** 2
    h = np.sqrt(area / (n * 1.02))
    if h <= min_sep:
        raise ValueError("n too large for the minimum separation")
    nx = int(np.floor((box[1] - box[0]) / h))
    ny = int(np.floor((box[3] - box[2]) / h))
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    cx = box[0] + (ix.ravel() + 0.5) * h
    cy = box[2] + (iy.ravel() + 0.5) * h
    amp = 0.5 * (h - min_sep)
    x = cx + rng.uniform(-amp, amp, cx.size)
    y = cy + rng.uniform(-amp, amp, cy.size)
    keep = np.hypot(x, y) >= body_radius + min_sep
    x, y = x[keep], y[keep]
    if x.size < n:
        raise ValueError("grid holds fewer than n cells outside the body")
    pick = rng.permutation(x.size)[:n]
    x, y = x[pick], y[pick]
    if layers:
        mag = rng.uniform(max_ve_g / 3.0, max_ve_g, n)
        return x, y, np.where(y > 0.0, -mag, mag)
    mag = rng.uniform(0.0, max_ve_g, n)
    sign = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    return x, y, mag * sign
