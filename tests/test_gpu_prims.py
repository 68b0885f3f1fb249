"""GPU parity of the tree's Morton keys and sort order: bit-exact against the
oracle's keys (tests/test_oracle_morton.py pins those by bit-interleave known
answers) and its stable order, through the library's own ballot-ranked LSD
radix sort.  Sizes cover one partial tile, many tiles with a ragged end,
heavy key ties (duplicate positions) and excluded slots that sort last.

The scans and ordered selections behind compaction, hitting, collapse, the
wall list and the right-hand side's near sources are covered bit for bit by
the step, compaction and collapse parity tests (test_gpu_parity.py).
"""
import numpy as np
import pytest

from vortex_amd import geometry as geo
from vortex_amd.solver import Case, make_params, VX_PREC_F64, VX_PREC_MIXED
from vortex_amd.state import far_tail

pytestmark = pytest.mark.gpu

BOX = (-2.0, 10.0, -10.0, 10.0)


def cloud(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.5, 9.5, n)
    y = rng.uniform(-3, 3, n)
    g = rng.uniform(-0.01, 0.01, n)
    # duplicated positions (equal keys: order by slot) and a block of copies
    k = n // 10
    if k:
        src = rng.integers(0, n, k)
        dst = rng.integers(0, n, k)
        x[dst], y[dst] = x[src], y[src]
        x[: min(n, 300)] = 1.25
        y[: min(n, 300)] = -0.75
    # dead slots: some near the box (kept), the far tail (excluded)
    dead = rng.random(n) < 0.05
    g[dead] = 0.0
    return x, y, g


@pytest.mark.parametrize("n,prec", [(37, VX_PREC_F64), (4096, VX_PREC_MIXED), (70001, VX_PREC_F64),
                                    (300003, VX_PREC_MIXED)])
def test_tree_keys_and_order_bitexact(ref, n, prec):
    print(f"[prims] n={n} prec={prec}", flush=True)
    x, y, g = cloud(n, n)
    s = (n + 31) // 32 * 32 + 64          # padded count plus far-tail slots
    tx, ty = far_tail(n, s - n)
    X = np.concatenate([x, tx]); Y = np.concatenate([y, ty]); G = np.concatenate([g, np.zeros(s - n)])
    pan = geo.circle(64, 0.5)
    prm = make_params(pan, precision=prec, ve_size=0.004, box=BOX)
    with Case(prm, pan) as c:
        c.set_state(X, Y, G)
        c.tree_biot_savart(s, (1.0, 0.0))
        keys, perm = c.tree_order(s)
        m = c.tree_stats()["m"]
    kref, mref, _ = ref.morton_keys(X, Y, G, BOX)
    pref = ref.stable_order(kref)
    assert m == mref
    assert np.array_equal(perm, pref)
    assert np.array_equal(keys, kref[pref])
    assert np.all(keys[mref:] == 0xFFFFFFFF) and np.all(keys[:mref] != 0xFFFFFFFF)
