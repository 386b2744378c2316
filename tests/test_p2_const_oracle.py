"""Pins of the oracle's P2 Gauss-Seidel sweep and of its P2 GMG with Gauss-Seidel
smoothing (the set-up of the reference's P2GMG3DConvergenceTest)."""
import numpy as np
import scipy.sparse as sp

from meshes import octahedron
from oracle import pyoracle as po


def test_gs_sweep_is_the_textbook_sweep():
    """the triangular-solve form equals the row-by-row definition
    x_i <- (b_i - sum_{j != i} a_ij x_j) / a_ii in canonical order, rows outside the
    mask untouched"""
    rng = np.random.default_rng(11)
    n = 40
    R = sp.random(n, n, density=0.15, random_state=3)
    A = (R + R.T + 8 * sp.eye(n)).tocsr()
    b, x0 = rng.normal(size=n), rng.normal(size=n)
    m = rng.uniform(size=n) < 0.7
    x = x0.copy()
    po.gs_sweep(A, x, b, m)
    y = x0.copy()
    D = A.toarray()
    for i in range(n):
        if m[i]:
            y[i] = (b[i] - D[i] @ y + D[i, i] * y[i]) / D[i, i]
    assert np.abs(x - y).max() < 1e-12


def test_p2_gs_gmg_meets_the_reference_rate():
    """P2GMG3DConvergenceTest.cpp:60-140: octahedron, levels 0-3, V(2,2) Gauss-Seidel,
    CG on level 0, u random inside and sin(x) sinh(y) z on the boundary, f = 0: the
    squared Inner residual drops below 3e-2 of its previous value every cycle"""
    X, cells = octahedron()
    m = po.Mesh.from_arrays(X, cells)
    g = po.P2GSGMG(m, 0, 3)
    L = 3
    P = m.coords(L + 1)
    fl = m.flags(L + 1)
    u = np.where(fl == po.INNER, np.random.default_rng(0).uniform(0, 1, len(P)),
                 np.sin(P[:, 0]) * np.sinh(P[:, 1]) * P[:, 2])
    f = np.zeros_like(u)
    inner = fl == po.INNER
    last = None
    for _ in range(4):
        if last is None:
            r = g.residual(L, u, f)
            last = r[inner] @ r[inner]
        g.cycle(L, u, f)
        r = g.residual(L, u, f)
        cur = r[inner] @ r[inner]
        assert cur / last < 3e-2, cur / last
        last = cur
