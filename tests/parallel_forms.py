"""Order-independent (parallel) formulations of the reference's sequential
compaction and collapse, as the HIP kernels compute them.  Used by the tests
to cross-check the oracle's sequential loops (src/kernel.cu:1159-1178,
1237-1271) against a second, differently structured algorithm."""
import numpy as np

DELT = 1e-12


def compact_closed_form(x, y, g, n):
    """swap-with-last removal, closed form.

    n1 = number of live entries.  The j-th hole of [0, n1) (ascending)
    receives the j-th live entry of [n1, n) (descending).  Every position e of
    [n1, n) ends dead (g = 0) holding the position of src(e): e+1 when e+1 < n
    is dead, otherwise the start of e's chain -- the hole h_j for the chain
    that ends at the j-th live entry, or n1 itself below the last one."""
    x = np.array(x, dtype=float); y = np.array(y, dtype=float); g = np.array(g, dtype=float)
    xo, yo, go = x.copy(), y.copy(), g.copy()
    live = np.abs(g[:n]) >= DELT
    n1 = int(live.sum())
    holes = [i for i in range(n1) if not live[i]]
    donors = [i for i in range(n - 1, n1 - 1, -1) if live[i]]
    assert len(holes) == len(donors)
    for h, dn in zip(holes, donors):
        xo[h], yo[h], go[h] = x[dn], y[dn], g[dn]
    k = len(donors)
    for e in range(n1, n):
        above = sum(1 for dn in donors if dn > e)
        j = above + 1
        start = holes[j - 1] if j <= k else n1
        src = e + 1 if (e + 1 < n and not live[e + 1]) else start
        xo[e], yo[e], go[e] = x[src], y[src], 0.0
    return xo, yo, go, n1


def collapse_parallel(x, y, g, COL, kind):
    """Greedy merge resolved without a sequential sweep.

    A candidate i (COL[i] >= 0) is active when no active candidate i' < i has
    COL[i'] = i (it would have been killed first).  Among the active
    candidates sharing one target j, the smallest merges with j as it was;
    the others merge with the already killed j (g = 0 at 1e10)."""
    x = np.array(x, dtype=float); y = np.array(y, dtype=float); g = np.array(g, dtype=float)
    n = len(COL)
    cand = [i for i in range(n) if COL[i] >= 0]
    children = {i: [] for i in range(n)}
    for i in cand:
        children[int(COL[i])].append(i)
    state = {}
    while len(state) < len(cand):
        for i in cand:
            if i in state:
                continue
            ch = children[i]
            if all(c in state for c in ch):
                state[i] = not any(state[c] for c in ch)
    active = [i for i in cand if state[i]]
    first = {}
    for i in active:
        j = int(COL[i])
        first[j] = min(first.get(j, i), i)
    xo, yo, go = x.copy(), y.copy(), g.copy()
    for i in active:
        j = int(COL[i])
        if first[j] == i:
            xj, yj, gj = x[j], y[j], g[j]
        else:
            xj, yj, gj = 1e10, 1e10, 0.0
        gi = g[i]
        if kind == 1:
            mn = abs(gi) + abs(gj)
            xo[i] = (x[i] * abs(gi) + xj * abs(gj)) / abs(mn)
            yo[i] = (y[i] * abs(gi) + yj * abs(gj)) / abs(mn)
            go[i] = gi + gj
        else:
            ng = gi + gj
            xo[i] = (x[i] * gi + xj * gj) / ng
            yo[i] = (y[i] * gi + yj * gj) / ng
            go[i] = ng
    for i in active:
        j = int(COL[i])
        xo[j], yo[j], go[j] = 1e10, 1e10, 0.0
    return xo, yo, go
