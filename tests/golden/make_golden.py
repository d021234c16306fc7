"""Writes the golden fixtures used by the oracle tests.

Periodic Delaunay rings of seeded random monolayers, computed independently of
the oracle with scipy.spatial.Delaunay (Qhull) on a 9-sheet periodic copy — the
same construction the reference uses for non-square boxes
(DelaunayCGAL.cpp:79-154, PeriodicTriangulationNineSheeted). Positions come from
std::mt19937(seed) + uniform_real_distribution<double>, which the oracle's
random_init reproduces (reference Simple2DCell.cpp:85-100); the fixture stores
the positions themselves so the rings do not depend on that generator.

Run:  python tests/golden/make_golden.py   (writes tests/golden/delaunay_rings.npz)
"""
import os
import sys

import numpy as np
from scipy.spatial import Delaunay

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def nine_sheet_rings(x, y, L):
    N = len(x)
    pts = np.stack([x, y], 1)
    allp, idx = [], []
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            allp.append(pts + np.array([dx * L, dy * L]))
            idx.append(np.arange(N))
    allp = np.concatenate(allp)
    idx = np.concatenate(idx)
    tri = Delaunay(allp)
    ip, ind = tri.vertex_neighbor_vertices
    center = 4 * N  # (dx, dy) = (0, 0) block
    rings = []
    for i in range(N):
        nb = ind[ip[center + i]: ip[center + i + 1]]
        # counter-clockwise order around i, canonical rotation: smallest index first
        d = allp[nb] - pts[i]
        ang = np.arctan2(d[:, 1], d[:, 0])
        order = nb[np.argsort(ang)]
        ids = list(idx[order])
        k0 = int(np.argmin(ids))
        rings.append(ids[k0:] + ids[:k0])
    return rings


def main():
    from oracle import oracle as O  # positions only (mt19937 draws)
    out = {}
    for N, seed in [(64, 13377), (256, 1), (1000, 13377), (2000, 42)]:
        x, y, th, L = O.random_init(N, seed)
        rings = nine_sheet_rings(x, y, L)
        nn = np.array([len(r) for r in rings], np.int32)
        flat = np.array([v for r in rings for v in r], np.int32)
        out["N%d_s%d_x" % (N, seed)] = x
        out["N%d_s%d_y" % (N, seed)] = y
        out["N%d_s%d_L" % (N, seed)] = np.array([L])
        out["N%d_s%d_nn" % (N, seed)] = nn
        out["N%d_s%d_rings" % (N, seed)] = flat
    np.savez_compressed(os.path.join(HERE, "delaunay_rings.npz"), **out)
    print("wrote", os.path.join(HERE, "delaunay_rings.npz"))


if __name__ == "__main__":
    main()
