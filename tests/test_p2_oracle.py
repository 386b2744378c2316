"""Pins of the P2 / shell part of the CPU oracle -- runs without a GPU.

Independent checks only: the mesh counts the reference documents, geometric
properties that fix the shell map, exactness of the quadrature, closed-form
P2 element matrices, and integral identities of the assembled operator.
"""
from math import factorial

import numpy as np
import pytest
from scipy import integrate

from hyteg_amd import problems
from oracle import pyoracle as po


@pytest.fixture(scope="module")
def shell():
    return po.Mesh.shell(**problems.SHELL)


def test_shell_counts_and_radii(shell):
    # MeshGenSphericalShell.cpp:699: (ntan-1)^2 (nrad-1) 60 cells; 486 vertices,
    # 2724 edges, 4160 faces for ntan=5, nrad=3
    assert shell.counts() == (486, 2724, 4160, 1920)
    X = shell.coords(0)
    r = np.linalg.norm(X, axis=1)
    assert np.allclose(np.sort(np.unique(np.round(r, 12))), [0.55, 0.775, 1.0])
    # Dirichlet exactly on the two spheres
    fl = shell.flags(0)
    assert np.all((fl == po.DIRICHLET) == (np.abs(r - 0.775) > 1e-9))


def test_shell_cells_have_positive_volume_and_tile_the_shell(shell):
    X = shell.coords(0)
    C = shell.cell_vertices()
    P = X[C]
    vol = np.einsum("ij,ij->i", P[:, 1] - P[:, 0], np.cross(P[:, 2] - P[:, 0], P[:, 3] - P[:, 0])) / 6
    assert np.all(np.abs(vol) > 1e-6)
    # the flat-faced shell lies inside the curved one and fills most of it
    bnd = np.abs(np.linalg.norm(X, axis=1) - 1.0) < 1e-12
    assert np.abs(vol).sum() < 4 / 3 * np.pi * (1 - 0.55 ** 3)
    assert np.abs(vol).sum() > 0.9 * 4 / 3 * np.pi * (1 - 0.55 ** 3)
    assert bnd.sum() == 162


def test_shell_map_properties(shell):
    # the map is radial, fixes the macro vertices and maps the inner/outer
    # triangle of every prism onto its sphere; DF is the derivative of F
    X = shell.coords(0)
    C = shell.cell_vertices()
    B = shell.blend()
    rng = np.random.default_rng(0)
    for c in rng.choice(len(C), 40, replace=False):
        V = X[C[c]]
        for v in V:
            F, _ = po.shell_map(B[c], v)
            assert np.abs(F - v).max() < 1e-14
        lam = rng.dirichlet(np.ones(4))
        x = lam @ V
        F, DF = po.shell_map(B[c], x)
        assert np.linalg.norm(np.cross(F, x)) < 1e-14
        h = 1e-6
        fd = np.array([(po.shell_map(B[c], x + h * e)[0] - po.shell_map(B[c], x - h * e)[0]) / (2 * h)
                       for e in np.eye(3)]).T
        assert np.abs(fd - DF).max() < 1e-8
        # points on a face with all three vertices on one sphere stay on it
        r = np.linalg.norm(V, axis=1)
        for rs in (0.55, 0.775, 1.0):
            on = np.abs(r - rs) < 1e-12
            if on.sum() == 3:
                w = rng.dirichlet(np.ones(3))
                F, _ = po.shell_map(B[c], w @ V[on])
                assert abs(np.linalg.norm(F) - rs) < 1e-14


def test_quadrature_is_degree_4_exact():
    pts, w = po.quadrature()
    assert np.all(w > 0) and abs(w.sum() - 1 / 6) < 1e-15
    assert np.allclose(pts.sum(axis=1), 1.0, atol=1e-15)
    for a in range(5):
        for b in range(5 - a):
            for c in range(5 - a - b):
                exact = factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 3)
                got = (w * pts[:, 1] ** a * pts[:, 2] ** b * pts[:, 3] ** c).sum()
                assert abs(got - exact) < 1e-13 * exact


def p2_stiffness_exact(P):
    """closed-form P2 stiffness of an affine tetrahedron: integrals of products of
    barycentric monomials, int l^a = |T| 3! a! / (|a| + 3)!"""
    A = np.ones((4, 4))
    A[:, 1:] = P
    G = np.linalg.inv(A)[1:, :].T  # grad l_i
    vol = abs(np.linalg.det(A)) / 6

    def integral(pows):
        num = np.prod([factorial(k) for k in pows])
        return vol * 6 * num / factorial(sum(pows) + 3)

    # basis: vertices l_i(2 l_i - 1), edges 4 l_a l_b; gradients as sums of c * l_m * grad l_n
    edges = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    terms = []
    for i in range(4):
        terms.append([(4.0, i, i), (-1.0, None, i)])
    for a, b in edges:
        terms.append([(4.0, a, b), (4.0, b, a)])
    K = np.zeros((10, 10))
    for i in range(10):
        for j in range(10):
            s = 0.0
            for ci, mi, ni in terms[i]:
                for cj, mj, nj in terms[j]:
                    pows = [0, 0, 0, 0]
                    for m in (mi, mj):
                        if m is not None:
                            pows[m] += 1
                    s += ci * cj * (G[ni] @ G[nj]) * integral(pows)
            K[i, j] = s
    return K


def test_element_matrix_closed_form_affine():
    rng = np.random.default_rng(1)
    for _ in range(5):
        P = rng.standard_normal((4, 3))
        A = po.p2_element_matrix(P)
        assert np.abs(A - p2_stiffness_exact(P)).max() < 1e-12 * np.abs(A).max()


def test_element_matrix_rows_sum_to_zero_and_symmetric(shell):
    a, om, ph = problems.shell_coefficient(0)
    k = po.KCoeff(a, om, ph)
    X = shell.coords(0)
    C = shell.cell_vertices()
    B = shell.blend()
    for c in (0, 17, 640, 1919):
        P = X[C[c]]
        # a micro-cell of level 2 inside this macro-cell
        sub = P[0] + 0.25 * np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]) @ (P[1:] - P[0])
        A = po.p2_element_matrix(sub, B[c], k)
        assert np.abs(A.sum(axis=1)).max() < 1e-13 * np.abs(A).max()
        assert np.abs(A - A.T).max() < 1e-14 * np.abs(A).max()
        assert np.linalg.eigvalsh(A).min() > -1e-13 * np.abs(A).max()


def test_operator_energy_of_quadratic_on_cube():
    # u = x^2 + y z is a P2 function: u^T A u = int |grad u|^2 = 4/3 + 1/3 + 1/3
    m = po.Mesh.cuboid()
    for L in (1, 2):
        op = po.P2Op(m, L)
        X = m.coords(L + 1)
        u = X[:, 0] ** 2 + X[:, 1] * X[:, 2]
        Au = np.zeros_like(u)
        op.apply(u, Au, po.ALL)
        assert abs(u @ Au - 2.0) < 1e-12


def test_operator_with_variable_k_on_cube():
    # u = x is exact, u^T A u = int k; the 11-point rule on 8^L micro-cells is
    # compared with adaptive quadrature of k over the cube
    a, om, ph = problems.shell_coefficient(3, modes=4, frequency=1.0)
    k = po.KCoeff(a, om, ph)
    m = po.Mesh.cuboid()
    op = po.P2Op(m, 3, k)
    X = m.coords(4)
    u = X[:, 0].copy()
    Au = np.zeros_like(u)
    op.apply(u, Au, po.ALL)
    ref, _ = integrate.nquad(lambda z, y, x: k(np.array([x, y, z]))[0], [[0, 1]] * 3, opts={"epsabs": 1e-12})
    assert abs(u @ Au - ref) < 1e-7 * ref
    # constants are in the kernel of every row
    one = np.ones_like(u)
    op.apply(one, Au, po.ALL)
    assert np.abs(Au).max() < 1e-12


def test_operator_symmetric_on_shell(shell):
    a, om, ph = problems.shell_coefficient(0)
    op = po.P2Op(shell, 0, po.KCoeff(a, om, ph))
    A = op.matrix()
    assert abs(A - A.T).max() < 1e-13 * abs(A).max()
    fl = shell.flags(1)
    inner = np.where(fl == po.INNER)[0]
    assert np.linalg.eigvalsh(A[inner][:, inner].toarray()).min() > 0
    d = op.inv_diag()
    assert np.allclose(d, 1.0 / A.diagonal())


# ---- the reduced element form of the low-precision applies ------------------------


def test_reduced_form_equals_the_reference_form_on_affine_cells():
    """affine tetrahedron, k = 1: the 4-point degree-2 rule integrates the P2
    stiffness exactly, so the reduced form equals the reference form (whose
    closed form is pinned above); with a constant k the matrix scales by k rounded
    to binary16"""
    p = np.array([[0.1, 0.0, 0.2], [0.9, 0.1, 0.0], [0.2, 1.1, 0.1], [0.1, 0.3, 0.8]])
    A = po.p2_element_matrix(p)
    R = po.p2_element_matrix(p, reduced=True)
    assert np.abs(R - A).max() < 1e-13 * np.abs(A).max()
    k = po.KCoeff([0.0], [[0.0, 0.0, 0.0]], [0.0])  # exp(0) = 1
    assert np.abs(po.p2_element_matrix(p, k=k, reduced=True) - A).max() < 1e-13 * np.abs(A).max()
    c = np.exp(0.7 * np.sin(0.4))  # k = exp(0.7 sin(0 . x + 0.4)), constant
    k = po.KCoeff([0.7], [[0.0, 0.0, 0.0]], [0.4])
    c16 = float(np.float16(np.float32(c)))
    assert np.abs(po.p2_element_matrix(p, k=k, reduced=True) - c16 * A).max() < 1e-13 * c * np.abs(A).max()


def test_reduced_form_converges_to_the_reference_form_on_the_shell():
    """on the blended shell the reduced form is symmetric, has zero row sums
    (constants in its kernel) and differs from the reference form by O(h): the
    relative difference of a micro-cell's matrix halves from level to level"""
    m = po.Mesh.shell(**problems.SHELL)
    B = m.blend()
    k = po.KCoeff(*problems.shell_coefficient(0))
    X, cv = m.coords(0), m.cell_vertices()
    for c in (0, 700, 1919):
        P = X[cv[c]]
        diffs = []
        for L in range(2, 8):
            h = 2.0 ** -L
            pp = np.array([P[0], P[0] + h * (P[1] - P[0]), P[0] + h * (P[2] - P[0]), P[0] + h * (P[3] - P[0])])
            A = po.p2_element_matrix(pp, B[c], k)
            R = po.p2_element_matrix(pp, B[c], k, reduced=True)
            assert np.abs(R - R.T).max() <= 1e-15 * np.abs(R).max()
            assert np.abs(R.sum(axis=1)).max() < 1e-13 * np.abs(R).max()
            diffs.append(np.abs(R - A).max() / np.abs(A).max())
        ratios = np.array(diffs[:-1]) / np.array(diffs[1:])
        assert diffs[-1] < 2e-3 and np.all((ratios > 1.6) & (ratios < 2.5)), (c, diffs)
