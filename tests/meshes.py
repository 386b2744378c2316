"""Meshes used by the tests: the unit cube (meshCuboid, 6 tetrahedra) and the
8-tetrahedra regular octahedron of HyTeG's P1GMG3DConvergenceTest (fixture)."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def octahedron():
    d = np.load(os.path.join(HERE, "golden", "octahedron_8el.npz"))
    return d["coords"], d["cells"]


def match(coords_a, coords_b, tol=1e-9):
    """permutation perm with coords_b[perm[i]] == coords_a[i]"""
    ka = np.round(np.asarray(coords_a) / tol).astype(np.int64)
    kb = np.round(np.asarray(coords_b) / tol).astype(np.int64)
    lut = {tuple(k): i for i, k in enumerate(kb)}
    return np.array([lut[tuple(k)] for k in ka])


def relmax(a, b):
    s = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / s)
