"""Writes tests/golden/winding_golden.npz: small meshes, points and their FP64 solid angles.

The values come from the FP64 brute force of the oracle and are cross-checked against the
independent numpy implementation before they are written.  Meshes: the single octant triangle
and the unit cube of the reference's pytest fixtures (modules/conftest.py:29-37, 108-144), a
level-2 icosphere, an open displaced sphere and a torus.  Points: uniform in each box x1.4 plus
near-surface offsets (1e-6 .. 1e-2 x diagonal).

Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from lagrange_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402

CUBE_V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
CUBE_F = np.array([[0, 3, 2], [2, 1, 0], [4, 5, 6], [6, 7, 4], [1, 2, 6], [6, 5, 1], [4, 7, 3], [3, 0, 4],
                   [2, 3, 7], [7, 6, 2], [0, 1, 5], [5, 4, 0]])


def main():
    meshes = {
        "single_triangle": W.Mesh(np.eye(3), np.array([[0, 1, 2]])),
        "cube": W.Mesh(CUBE_V, CUBE_F),
        "icosphere_nu4": W.icosphere(4),
        "open_displaced_nu8": W.displaced_sphere(8, 17, delete_fraction=0.1),
        "torus_24x8": W.torus(24, 8, seed=5, amplitude=0.2),
    }
    out = {}
    for i, (name, m) in enumerate(meshes.items()):
        V, F = m.f32()
        lo, hi = V.min(0), V.max(0)
        q = [W.uniform_queries(lo, hi, 160, 100 + i, 1.4)]
        if F.shape[0] > 1:
            q.append(W.near_surface_queries(W.Mesh(V.astype(np.float64), F.astype(np.int64)), 40, 200 + i, -6, -2))
        Q = np.concatenate(q)
        om = O.bruteforce(V, F, Q)
        om2 = O.bruteforce_numpy(V, F, Q)
        assert np.abs(om - om2).max() < 1e-11, name
        out[f"{name}/V"] = V
        out[f"{name}/F"] = F
        out[f"{name}/Q"] = Q
        out[f"{name}/omega"] = om
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "winding_golden.npz")
    np.savez_compressed(path, **out)
    print(path, sum(v.nbytes for v in out.values()), "bytes")


if __name__ == "__main__":
    main()
