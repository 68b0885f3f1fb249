"""Oracle Morton keys and stable order, pinned by bit-interleave known answers
and brute force.

The tree's keys are a build-new structure (the reference's own tree,
src/kernel.cu:79-763, is a bisection kd-tree and inactive), so the known
answers come from the definition: bit b of the x cell index goes to key bit
2b, bit b of the y cell index to key bit 2b + 1.
"""
import numpy as np
import pytest


def interleave_bruteforce(ix, iy):
    k = 0
    for b in range(16):
        k |= ((ix >> b) & 1) << (2 * b)
        k |= ((iy >> b) & 1) << (2 * b + 1)
    return k


@pytest.mark.parametrize("v,want", [
    (0, 0), (1, 1), (2, 4), (3, 5), (0x8000, 0x40000000), (0xFFFF, 0x55555555),
    (0xAAAA, 0x44444444), (0x5555, 0x11111111), (0x1234, 0x01040510),
])
def test_spread16_known_answers(ref, v, want):
    assert ref.lib().vxr_spread16(v) == want


def test_interleave_known_answers(ref):
    sp = ref.lib().vxr_spread16
    cases = [((1, 0), 1), ((0, 1), 2), ((1, 1), 3), ((3, 5), 0b100111), ((65534, 65534), 0xFFFFFFFC),
             ((0, 65534), 0xAAAAAAA8), ((65534, 0), 0x55555554)]
    for (ix, iy), want in cases:
        assert (sp(ix) | (sp(iy) << 1)) & 0xFFFFFFFF == want, (ix, iy)
        assert interleave_bruteforce(ix, iy) == want


def test_morton_keys_against_bruteforce(ref):
    rng = np.random.default_rng(7)
    n = 3000
    x = rng.uniform(-3, 12, n)
    y = rng.uniform(-5, 5, n)
    g = rng.normal(size=n)
    g[rng.random(n) < 0.2] = 0.0
    # dead slots far away (excluded) and near the box (included)
    x[:50] = 2e5 + np.arange(50)
    g[:50] = 0.0
    x[50:60] = 11.5
    g[50:60] = 0.0
    x[60:70] = np.nan
    box = (-2.0, 10.0, -10.0, 10.0)
    keys, m, ob = ref.morton_keys(x, y, g, box)
    inc = np.isfinite(x) & np.isfinite(y) & (
        (g != 0) | ((np.maximum(0, np.maximum(box[0] - x, x - box[1])) ** 2 +
                     np.maximum(0, np.maximum(box[2] - y, y - box[3])) ** 2) <= 400.0 ** 2))
    assert m == inc.sum()
    x0, x1, y0, y1 = x[inc].min(), x[inc].max(), y[inc].min(), y[inc].max()
    assert tuple(ob) == (x0, x1, y0, y1)
    for i in range(n):
        if not inc[i]:
            assert keys[i] == 0xFFFFFFFF
            continue
        fx = min(max((x[i] - x0) / (x1 - x0) * 65534.0, 0.0), 65534.0)
        fy = min(max((y[i] - y0) / (y1 - y0) * 65534.0, 0.0), 65534.0)
        assert keys[i] == interleave_bruteforce(int(fx), int(fy)), i
    # the box corners map to the extreme keys
    lo = np.argmin(np.where(inc, x, np.inf))
    assert (keys[lo] & 0x55555555) == 0


def test_morton_keys_degenerate_boxes(ref):
    # one included vortex: zero-width box maps to cell 0; nothing included: unit box
    x = np.array([1.0, 3e5]); y = np.array([2.0, 0.0]); g = np.array([0.5, 0.0])
    keys, m, ob = ref.morton_keys(x, y, g, (-2.0, 10.0, -10.0, 10.0))
    assert m == 1 and keys[0] == 0 and keys[1] == 0xFFFFFFFF
    keys, m, ob = ref.morton_keys(x[1:], y[1:], g[1:], (-2.0, 10.0, -10.0, 10.0))
    assert m == 0 and tuple(ob) == (0.0, 1.0, 0.0, 1.0)


@pytest.mark.parametrize("n", [0, 1, 7, 5000])
def test_stable_order_matches_numpy(ref, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 64, n).astype(np.uint32) * np.uint32(0x01010101)  # many ties
    keys[rng.random(n) < 0.1] = 0xFFFFFFFF
    perm = ref.stable_order(keys)
    assert np.array_equal(perm, np.argsort(keys, kind="stable"))
