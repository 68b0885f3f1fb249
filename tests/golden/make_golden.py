"""Writes the small golden fixtures under tests/golden/.

circle200_sheet.npz: sheet strengths of uniform flow U=(1,0) past a circle of
radius 0.5 with 200 panels, from numpy.linalg.solve of the (P+1)^2 panel
system (influence matrix from the oracle, pinned separately against
quadrature) with right-hand side [-U.tau; 0].
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

from vortex_amd import geometry as geo  # noqa: E402
import vx_ref  # noqa: E402


def main():
    pan = geo.circle(200, 0.5)
    L = vx_ref.panel_matrix(pan)
    rhs = np.zeros(201)
    rhs[:200] = -(1.0 * pan["tang"][:, 0] + 0.0 * pan["tang"][:, 1])
    g = np.linalg.solve(L, rhs)
    np.savez(os.path.join(HERE, "circle200_sheet.npz"), g_sheet=g[:200])


if __name__ == "__main__":
    main()
