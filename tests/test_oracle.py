"""Pins of the CPU oracle (oracle/hg_oracle.c) -- runs without a GPU.

Every check here is independent of the oracle's own formulas: a brute-force
numpy assembly written from the textbook P1 definition (golden fixture), the
reference's own test thresholds, closed-form identities of P1 elements, or a
direct sparse solve.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

import bruteforce as bf
from meshes import match, octahedron, relmax
from oracle import pyoracle as po

GOLD = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "p1_bruteforce.npz"))


def mesh(name):
    if name == "cube":
        return po.Mesh.cuboid()
    X, cells = octahedron()
    return po.Mesh.from_arrays(X, cells)


def test_refinement_tiles_the_macro_cell():
    # the 6 micro-cell shapes must tile a macro tetrahedron conformingly:
    # volumes add up and every interior micro-face is shared by two micro-cells
    V = np.array([[0, 0, 0], [1.0, 0, 0], [0.2, 1.1, 0], [0.1, 0.3, 0.9]])
    for L in (1, 2, 3):
        m = 2 ** L
        vol, faces = 0.0, {}
        for t, offs in enumerate(bf.SUBTETS):
            wt = m - bf.SHIFT[t]
            for z in range(wt):
                for y in range(wt - z):
                    for x in range(wt - z - y):
                        q = [(x + o[0], y + o[1], z + o[2]) for o in offs]
                        P = np.array([V[0] + (a / m) * (V[1] - V[0]) + (b / m) * (V[2] - V[0]) + (c / m) * (V[3] - V[0])
                                      for a, b, c in q])
                        vol += abs(np.linalg.det(np.c_[np.ones(4), P])) / 6
                        for skip in range(4):
                            f = tuple(sorted(q[:skip] + q[skip + 1:]))
                            faces[f] = faces.get(f, 0) + 1
        assert count_cells(m) == m ** 3
        assert abs(vol - abs(np.linalg.det(np.c_[np.ones(4), V])) / 6) < 1e-14
        on_bnd = lambda f: any(all(p[k] == 0 for p in f) for k in range(3)) or all(sum(p) == m for p in f)  # noqa
        for f, n in faces.items():
            assert n == (1 if on_bnd(f) else 2), (L, f, n)


def count_cells(m):
    return sum(sum(1 for z in range(m - s) for y in range(m - s - z) for x in range(m - s - z - y)) for s in bf.SHIFT)


@pytest.mark.parametrize("name", ["cube", "octa"])
@pytest.mark.parametrize("L", [1, 2, 3])
def test_apply_matches_bruteforce_fixture(name, L):
    m = mesh(name)
    op = po.P1Op(m, L)
    P = GOLD[f"{name}_L{L}_coords"]
    perm = match(m.coords(L), P)
    x = GOLD[f"{name}_L{L}_x"][perm]
    y = np.zeros_like(x)
    op.apply(x, y, po.ALL)
    assert relmax(y, GOLD[f"{name}_L{L}_Ax"][perm]) < 1e-13
    assert relmax(op.csr().diagonal(), GOLD[f"{name}_L{L}_diag"][perm]) < 1e-13


@pytest.mark.parametrize("name", ["cube", "octa"])
@pytest.mark.parametrize("L", [2, 3])
def test_linear_functions_in_kernel(name, L):
    # P1LaplaceOperator3DTest.cpp:48-118: A u = 0 on inner DoFs for u = 0, 1,
    # 42x, 42x + y + 1337z; discrete L2 error < 2.8e-13
    m = mesh(name)
    op = po.P1Op(m, L)
    X, fl = m.coords(L), m.flags(L)
    npts = (fl == po.INNER).sum()
    for coef in ([0, 0, 0, 0], [0, 0, 0, 1], [42, 0, 0, 0], [42, 1, 1337, 0]):
        u = X @ np.array(coef[:3], float) + coef[3]
        r = np.zeros_like(u)
        op.apply(u, r, po.INNER)
        err = np.sqrt(po.dot(r, r, fl, po.INNER) / npts)
        assert err < 2.8e-13 * max(1.0, np.abs(coef).max() / 42.0)


@pytest.mark.parametrize("name,vol", [("cube", 1.0), ("octa", 1.0 / 3.0 * 2)])
def test_energy_of_linear_function(name, vol):
    # u = a.x is represented exactly, so u^T A u = |a|^2 |Omega|
    m = mesh(name)
    op = po.P1Op(m, 3)
    a = np.array([1.0, -2.0, 0.5])
    u = m.coords(3) @ a
    Au = np.zeros_like(u)
    op.apply(u, Au, po.ALL)
    assert abs(u @ Au - (a @ a) * vol) < 1e-12


def test_symmetric_positive_definite_on_inner():
    m = mesh("octa")
    A = po.P1Op(m, 2).csr()
    assert abs(A - A.T).max() < 1e-15
    fl = m.flags(2)
    inner = np.where(fl == po.INNER)[0]
    ev = np.linalg.eigvalsh(A[inner][:, inner].toarray())
    assert ev.min() > 0


@pytest.mark.parametrize("name", ["cube", "octa"])
def test_restriction_is_transpose_of_prolongation(name):
    m = mesh(name)
    rng = np.random.default_rng(3)
    for Lf in (1, 2, 3):
        f = rng.standard_normal(m.ndofs(Lf))
        c = rng.standard_normal(m.ndofs(Lf - 1))
        Rf = np.zeros(m.ndofs(Lf - 1))
        m.restrict(Lf, f, Rf, po.ALL)
        Pc = np.zeros(m.ndofs(Lf))
        m.prolongate(Lf - 1, c, Pc, po.ALL, False)
        assert abs(Rf @ c - f @ Pc) < 1e-12 * (abs(Rf @ c) + 1)


def test_prolongation_reproduces_linear_functions():
    # VertexDoFLinearProlongation3DTest.cpp:107-130
    m = mesh("octa")
    a = np.array([0.3, -1.2, 2.0])
    for Lc in (0, 1, 2):
        c = m.coords(Lc) @ a + 0.7
        f = np.zeros(m.ndofs(Lc + 1))
        m.prolongate(Lc, c, f, po.ALL, False)
        assert np.abs(f - (m.coords(Lc + 1) @ a + 0.7)).max() < 1e-13


def test_restriction_of_one_on_cell_interiors():
    # centre 1 + 14 neighbours x 1/2 = 8 on coarse points inside macro-cells
    m = mesh("cube")
    Lf = 3
    f = np.ones(m.ndofs(Lf))
    c = np.zeros(m.ndofs(Lf - 1))
    m.restrict(Lf, f, c, po.ALL)
    nb = m.nbnd(Lf - 1)
    assert np.allclose(c[nb:], 8.0)


def test_flag_semantics_of_transfer():
    # non-matching owners are left untouched (communicateAdditively excludes them)
    m = mesh("cube")
    fl = m.flags(2)
    c = np.full(m.ndofs(1), 5.0)
    m.restrict(2, np.ones(m.ndofs(2)), c, po.INNER)
    assert np.all(c[m.flags(1) == po.DIRICHLET] == 5.0)
    f = np.full(m.ndofs(2), 3.0)
    m.prolongate(1, np.ones(m.ndofs(1)), f, po.INNER, True)
    assert np.all(f[fl == po.DIRICHLET] == 3.0)
    assert np.allclose(f[fl == po.INNER], 4.0)


def test_cg_matches_direct_solve():
    m = mesh("octa")
    L = 2
    op = po.P1Op(m, L)
    A = op.csr()
    fl = m.flags(L)
    I = np.where(fl == po.INNER)[0]
    rng = np.random.default_rng(5)
    b = np.zeros(m.ndofs(L))
    b[I] = rng.uniform(-1, 1, len(I))
    x = np.zeros_like(b)
    op.cg(x, b, po.GMG_FLAG, 1000, 1e-14, 1e-30)
    xs = sla.spsolve(A[I][:, I].tocsc(), b[I])
    assert relmax(x[I], xs) < 1e-10


def test_gmg_octahedron_rate_bound():
    # P1GMG3DConvergenceTest.cpp:60-146: octahedron, L0..3, V(3,3) GS, CG coarse,
    # random inner initial guess, 4 cycles, ratio of squared residuals < 3.2e-2
    m = mesh("octa")
    L = 3
    P, fl = m.coords(L), m.flags(L)
    ex = np.sin(P[:, 0]) * np.sinh(P[:, 1]) * P[:, 2]
    x = np.where(fl == po.INNER, np.random.default_rng(1).uniform(0, 1, len(ex)), ex)
    k, hist = m.gmg(0, L, x, np.zeros_like(x), 3, 3, cycles=4)
    assert k == 4
    assert np.all((hist[1:] / hist[:-1]) ** 2 < 3.2e-2)


def test_gmg_converges_to_direct_solution():
    m = mesh("cube")
    L = 3
    fl = m.flags(L)
    I = np.where(fl == po.INNER)[0]
    b = np.zeros(m.ndofs(L))
    b[I] = np.random.default_rng(0).uniform(-1, 1, len(I))
    x = np.zeros_like(b)
    k, hist = m.gmg(0, L, x, b, 3, 3, cycles=40, tol=1e-13)
    assert hist[-1] <= 1e-13 * hist[0]
    A = po.P1Op(m, L).csr()
    xs = sla.spsolve(A[I][:, I].tocsc(), b[I])
    assert relmax(x[I], xs) < 1e-10


# ---- Gauss-Seidel orderings -------------------------------------------------------


def _closure(m, L, p, q):
    """q lies in the closure of p's owner primitive (its owner's vertices are a
    subset of p's owner's vertices)"""
    return set(m.owner_vertices(L, q)) <= set(m.owner_vertices(L, p))


@pytest.mark.parametrize("name,L", [("cube", 3), ("octa", 2)])
@pytest.mark.parametrize("relax", [1.0, 1.3])
def test_lexicographic_sor_row_by_row(name, L, relax):
    """or_p1_sor equals the row-by-row definition on the assembled matrix:
    x_p <- (1 - w) x_p + w (b_p - sum_{q != p} a_pq x_q) / a_pp in canonical order
    (macro vertices, edges, faces, then cell interiors), where a macro-vertex/
    edge/face DoF reads the neighbours owned by other primitives outside its
    owner's closure from the values at the sweep's start (HyTeG's ghost layers
    are synchronised once before the sweep, P1Operator.hpp:373-412) and all
    other neighbours at their latest values"""
    m = mesh(name)
    op = po.P1Op(m, L)
    A = op.csr()
    n = A.shape[0]
    nb = m.nbnd(L)
    fl = m.flags(L)
    rng = np.random.default_rng(31 + L)
    x0, b = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    # the owner-vertex sets locate each DoF: it is a positive combination of them
    P = m.coords(L)
    V = m.coords(0)
    for p in range(0, n, max(1, n // 50)):
        vs = list(m.owner_vertices(L, p))
        M = np.vstack([V[vs].T, np.ones(len(vs))])
        lam, *_ = np.linalg.lstsq(M, np.r_[P[p], 1.0], rcond=None)
        assert np.abs(M @ lam - np.r_[P[p], 1.0]).max() < 1e-12 and (lam > 1e-12).all()
    for flag in (po.GMG_FLAG, po.ALL):
        x = x0.copy()
        op.sor(x, b, flag, relax)
        y = x0.copy()
        start = x0.copy()
        for p in range(n):
            if not (fl[p] & flag):
                continue
            row = A.getrow(p)
            s, diag = b[p], 0.0
            for q, a in zip(row.indices, row.data):
                if q == p:
                    diag += a
                    continue
                v = start[q] if (p < nb and q < nb and not _closure(m, L, p, q)) else y[q]
                s -= a * v
            y[p] = (1 - relax) * y[p] + relax * s / diag
        assert np.abs(x - y).max() < 1e-13 * max(1.0, np.abs(y).max()), (flag, relax)


def _simultaneous_colour_sweep(A, colour, mask, x, b, relax, backward):
    """every colour's rows updated at once from the values before that colour"""
    D = A.diagonal()
    order = sorted(set(colour[mask].tolist()), reverse=backward)
    for c in order:
        rows = np.where(mask & (colour == c))[0]
        r = b[rows] - A[rows] @ x + D[rows] * x[rows]
        x[rows] = (1 - relax) * x[rows] + relax * r / D[rows]
    return x


def _greedy(A, rows):
    """greedy colouring of `rows` in their order over the coupling graph of A"""
    col = {}
    A = A.tocsr()
    for p in rows:
        used = {col[q] for q in A.indices[A.indptr[p]:A.indptr[p + 1]] if q in col and q != p}
        c = 0
        while c in used:
            c += 1
        col[p] = c
    return col


@pytest.mark.parametrize("name,L", [("cube", 3), ("octa", 3)])
def test_multicolour_sweep_and_conflict_checker(name, L):
    """the colour-order sweep of the oracle equals updating each colour's rows
    simultaneously exactly when the colour-conflict checker finds no coupled pair
    inside a colour: with the parity colours of the cell interiors and a greedy
    colouring of the macro vertex/edge/face DoFs both hold, with two colours merged
    the checker reports the coupled pairs and the sweeps differ"""
    m = mesh(name)
    A = po.P1Op(m, L).csr()
    n = A.shape[0]
    nb = m.nbnd(L)
    g = _greedy(A, range(nb))
    nbc = max(g.values()) + 1
    colour = np.zeros(n, dtype=np.int32)
    for p, c in g.items():
        colour[p] = c
    # cell-interior parity colours from the lattice coordinates of each cell
    mm = 2 ** L
    for c in range(m.counts()[3]):
        for z in range(1, mm):
            for y in range(1, mm - z):
                for xx in range(1, mm - z - y):
                    colour[m.locate(L, c, xx, y, z)] = nbc + (xx & 1) + 2 * (y & 1) + 4 * (z & 1)
    assert po.colour_conflicts(A, colour) == []
    # brute force: the checker's pairs are all coupled same-colour pairs
    Ad = A.toarray()
    rng = np.random.default_rng(7)
    x0, b = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    mask = (m.flags(L) & po.GMG_FLAG) != 0
    for relax in (1.0, 1.3):
        for backward in (False, True):
            x = po.sor_multicolour(A, colour, mask, x0.copy(), b, relax, backward)
            y = _simultaneous_colour_sweep(A, colour, mask, x0.copy(), b, relax, backward)
            assert np.abs(x - y).max() < 1e-13, (relax, backward)
            assert np.array_equal(x[~mask], x0[~mask])
    bad = colour.copy()
    bad[bad == nbc + 1] = nbc  # merge two interior parity colours
    pairs = po.colour_conflicts(A, bad)
    brute = sorted({(min(p, q), max(p, q)) for p in range(n) for q in np.nonzero(Ad[p])[0]
                    if p != q and bad[p] == bad[q]})
    assert pairs == brute and len(pairs) > 0
    x = po.sor_multicolour(A, bad, mask, x0.copy(), b, 1.0)
    y = _simultaneous_colour_sweep(A, bad, mask, x0.copy(), b, 1.0, False)
    assert np.abs(x - y).max() > 1e-6
