"""Independent brute-force FEM assembly (numpy), written from the textbook
definition of P1 finite elements rather than from the reference's code.

Given macro vertices and macro cells (local vertex order), refine every macro
tetrahedron uniformly (Bey's 8-subtet rule, expressed as the 6 lattice
micro-cell shapes of a structured tetrahedral lattice), identify micro vertices
by their exact rational lattice coordinates, assemble the global stiffness
matrix from P1 element matrices K = |T| G G^T, and return it with the micro
vertex coordinates.  Comparisons against the oracle / product go through the
coordinates, so this code shares no numbering with them.
"""
from fractions import Fraction

import numpy as np
import scipy.sparse as sp

# micro-tets of the standard lattice refinement, as vertex offsets
SUBTETS = [
    [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)],
    [(1, 0, 0), (1, 1, 0), (0, 1, 0), (1, 0, 1)],
    [(1, 0, 0), (0, 1, 0), (1, 0, 1), (0, 0, 1)],
    [(1, 1, 0), (1, 1, 1), (0, 1, 1), (1, 0, 1)],
    [(1, 0, 1), (0, 1, 1), (0, 0, 1), (0, 1, 0)],
    [(0, 1, 0), (1, 1, 0), (1, 0, 1), (0, 1, 1)],
]
SHIFT = [0, 1, 1, 2, 1, 1]


def p1_element_stiffness(P):
    """K_ij = vol * grad(l_i) . grad(l_j) for barycentric coordinates l_i."""
    A = np.ones((4, 4))
    A[:, 1:] = P
    Ainv = np.linalg.inv(A)          # rows 1..3 of Ainv^T give gradients
    G = Ainv[1:, :].T                # G[i] = grad l_i
    vol = abs(np.linalg.det(A)) / 6.0
    return vol * G @ G.T


def assemble(coords, cells, level):
    coords = np.asarray(coords, dtype=float)
    m = 2 ** level
    key_to_id = {}
    pts = []

    def vid(c, q):
        # exact barycentric key: sorted (global vertex id, weight) pairs
        v = cells[c]
        t = (m - q[0] - q[1] - q[2], q[0], q[1], q[2])
        key = tuple(sorted((int(v[k]), t[k]) for k in range(4) if t[k] != 0))
        if key not in key_to_id:
            key_to_id[key] = len(pts)
            pts.append(sum(coords[g] * Fraction(w, m).__float__() for g, w in key))
        return key_to_id[key]

    rows, cols, vals = [], [], []
    for c in range(len(cells)):
        V = coords[np.asarray(cells[c])]
        E = np.stack([V[1] - V[0], V[2] - V[0], V[3] - V[0]], axis=1) / m
        for t, offs in enumerate(SUBTETS):
            wt = m - SHIFT[t]
            P = np.array([V[0] + E @ np.array(o, dtype=float) for o in offs])
            K = p1_element_stiffness(P)
            for z in range(wt):
                for y in range(wt - z):
                    for x in range(wt - z - y):
                        ids = [vid(c, (x + o[0], y + o[1], z + o[2])) for o in offs]
                        for i in range(4):
                            for j in range(4):
                                rows.append(ids[i])
                                cols.append(ids[j])
                                vals.append(K[i, j])
    n = len(pts)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    A.sum_duplicates()
    return A, np.array(pts)


def match(coords_a, coords_b, tol=1e-9):
    """permutation perm with coords_b[perm[i]] == coords_a[i]"""
    ka = np.round(np.asarray(coords_a) / tol).astype(np.int64)
    kb = np.round(np.asarray(coords_b) / tol).astype(np.int64)
    lut = {tuple(k): i for i, k in enumerate(kb)}
    return np.array([lut[tuple(k)] for k in ka])
