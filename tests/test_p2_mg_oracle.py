"""Pins of the oracle's P2 multigrid components (CPU only).

* quadratic prolongation reproduces quadratics exactly
  (P2QuadraticProlongation3DTest.cpp:195-263);
* restriction of the constant 1 on the single-tetrahedron mesh gives the
  weights P2QuadraticRestriction3DTest.cpp:46-103 lists for cell-interior DoFs
  (this also pins the micro-cell decomposition: 6/4/6/6/4/6/4 neighbour
  elements of the X/Y/Z/XY/XZ/YZ/XYZ edge DoFs);
* restriction is the transpose of prolongation;
* Chebyshev coefficients agree with the closed forms of
  ChebyshevSmoother.hpp:340-380 and with 1 - x p(x) = T_k((theta-x)/delta)/T_k(theta/delta);
* power iteration, PCG and the V-cycle behave as their definitions require.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from hyteg_amd import problems
from meshes import octahedron
from oracle import pyoracle as po

TET = (np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]), np.array([[0, 1, 2, 3]]))


def _quadratic(X):
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    return 1.0 + 2 * x - 3 * y + 0.5 * z + x * x - 2 * x * y + 3 * y * z - z * z + 0.7 * x * z + 1.3 * y * y


@pytest.mark.parametrize("mesh", ["cube", "octa", "tet"])
def test_prolongation_reproduces_quadratics(mesh):
    m = {"cube": po.Mesh.cuboid, "octa": lambda: po.Mesh.from_arrays(*octahedron()),
         "tet": lambda: po.Mesh.from_arrays(*TET)}[mesh]()
    for Lc in (0, 1, 2):
        cx, fx = m.coords(Lc + 1), m.coords(Lc + 2)
        coarse = _quadratic(cx)
        fine = np.full(len(fx), 7.0)
        m.p2_prolongate(Lc, coarse, fine, po.ALL)
        assert np.abs(fine - _quadratic(fx)).max() < 1e-13
        # prolongateAndAdd adds; Dirichlet DoFs are left alone without the flag
        fl = m.flags(Lc + 2)
        fine2 = np.ones(len(fx))
        m.p2_prolongate(Lc, coarse, fine2, po.INNER, add=True)
        ref = np.where(fl == po.INNER, 1.0 + _quadratic(fx), 1.0)
        assert np.abs(fine2 - ref).max() < 1e-13


@pytest.mark.parametrize("lower", [3, 4])
def test_restriction_of_one_on_single_tet(lower):
    m = po.Mesh.from_arrays(*TET)
    nf, nc = m.ndofs(lower + 2), m.ndofs(lower + 1)
    coarse = np.zeros(nc)
    m.p2_restrict(lower + 1, np.ones(nf), coarse, po.ALL)
    w = 2 ** (lower + 1) + 1
    expected_edge = {(1, 0, 0): 6, (0, 1, 0): 4, (0, 0, 1): 6, (1, 1, 0): 6, (1, 0, 1): 4, (0, 1, 1): 6,
                     (1, 1, 1): 4}
    seen = set()
    for z in range(1, w):
        for y in range(1, w - z):
            for x in range(1, w - y - z - 1):
                if x + y + z >= w - 1:
                    continue
                i = m.locate(lower + 1, 0, x, y, z)
                par = (x & 1, y & 1, z & 1)
                if par == (0, 0, 0):
                    exp = 1.0 + 14 * 3 / 8 - 14 / 8 - 3 * 24 / 8 - 24 / 8
                else:
                    n = expected_edge[par]
                    exp = 1.0 + 2 * 3 / 4 + n * 2 * 0.5 + n * 0.25 + n * 0.25
                assert abs(coarse[i] - exp) < 1e-12, (x, y, z, coarse[i], exp)
                seen.add(par)
    assert len(seen) == 8


def test_restriction_is_transpose_of_prolongation():
    m = po.Mesh.shell(**problems.SHELL)
    rng = np.random.default_rng(3)
    for Lc in (0, 1):
        nc, nf = m.ndofs(Lc + 1), m.ndofs(Lc + 2)
        c = rng.standard_normal(nc)
        f = rng.standard_normal(nf)
        Pc = np.zeros(nf)
        m.p2_prolongate(Lc, c, Pc, po.ALL)
        Rf = np.zeros(nc)
        m.p2_restrict(Lc + 1, f, Rf, po.ALL)
        assert abs(Rf @ c - f @ Pc) < 1e-11 * abs(f @ Pc) + 1e-11


def _reference_closed_forms(order, lower, upper):
    """ChebyshevSmoother.hpp:340-380, orders 1-4"""
    theta, delta = 0.5 * (upper + lower), 0.5 * (upper - lower)
    if order == 1:
        return [1.0 / theta]
    if order == 2:
        t0 = 1.0 / (delta ** 2 - 2 * theta ** 2)
        return [-4 * theta * t0, 2 * t0]
    if order == 3:
        t0, t1 = 3 * delta ** 2, theta ** 2
        t2 = 1.0 / (-4 * theta ** 3 + theta * t0)
        return [t2 * (t0 - 12 * t1), 12 / (t0 - 4 * t1), -4 * t2]
    t0, t1 = delta ** 2, theta ** 2
    t2 = 8 * t0
    t3 = 1.0 / (delta ** 4 + 8 * theta ** 4 - t1 * t2)
    return [t3 * (32 * theta ** 3 - 16 * theta * t0), t3 * (-48 * t1 + t2), 32 * theta * t3, -8 * t3]


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_chebyshev_coefficients(order):
    lower, upper = 0.3 * 1.7, 1.2 * 1.7
    c = po.chebyshev_coefficients(order, lower, upper)
    ref = np.array(_reference_closed_forms(order, lower, upper))
    assert np.abs(c - ref).max() < 1e-13 * np.abs(ref).max()
    theta, delta = 0.5 * (upper + lower), 0.5 * (upper - lower)
    lam = np.linspace(0.0, upper, 17)
    T = np.polynomial.chebyshev.Chebyshev.basis(order)
    lhs = 1.0 - lam * np.polyval(c[::-1], lam)
    assert np.abs(lhs - T((theta - lam) / delta) / T(theta / delta)).max() < 1e-12
    for k in range(5, 9):
        c = po.chebyshev_coefficients(k, lower, upper)
        T = np.polynomial.chebyshev.Chebyshev.basis(k)
        lhs = 1.0 - lam * np.polyval(c[::-1], lam)
        assert np.abs(lhs - T((theta - lam) / delta) / T(theta / delta)).max() < 1e-10


def _restricted_matrix(op, flags, flag):
    A = op.matrix()
    idx = np.flatnonzero(flags & flag)
    return A[idx][:, idx].tocsc(), idx


def test_power_iteration_and_pcg_on_shell():
    m = po.Mesh.shell(**problems.SHELL)
    k = po.KCoeff(*problems.shell_coefficient(0))
    op = po.P2Op(m, 0, k)
    fl = m.flags(1)
    inv = op.inv_diag()
    x = problems.canonical_uniform(op.n, 11)
    x0 = x.copy()
    rho = op.power_iteration(inv, x, 60)
    A = op.matrix()
    B = inv[:, None] * A.toarray()
    # the iteration of SpectrumEstimation.hpp:56-90 restated with the assembled matrix
    v = x0 / np.linalg.norm(x0)
    aux = B @ v
    for _ in range(60):
        v = aux / np.linalg.norm(aux)
        aux = B @ v
    assert abs(rho - v @ aux) < 1e-10 * rho
    lam = np.abs(np.linalg.eigvals(B)).max()
    assert 0.97 * lam < rho <= lam * (1 + 1e-12)
    b = np.where(fl == po.INNER, problems.canonical_uniform(op.n, 4), 0.0)
    x = np.zeros(op.n)
    it = op.cg(x, b, rel_tol=1e-14)
    Ai, idx = _restricted_matrix(op, fl, po.GMG_FLAG)
    xs = spla.spsolve(Ai, b[idx])
    assert np.abs(x[idx] - xs).max() < 1e-10 * np.abs(xs).max()
    assert it < 400


def test_p2_vcycle_converges_on_shell():
    m = po.Mesh.shell(**problems.SHELL)
    k = po.KCoeff(*problems.shell_coefficient(0))
    L = 2
    fl = m.flags(L + 1)
    b = np.where(fl == po.INNER, problems.canonical_uniform(len(fl), 0), 0.0)
    op = po.P2Op(m, L, k)
    rho = op.power_iteration(op.inv_diag(), problems.canonical_uniform(len(fl), 1), 20)
    x = np.zeros(len(fl))
    kk, hist = m.p2_gmg(0, L, k, x, b, 2, 2, 3, rho, cycles=12, tol=1e-8)
    factors = hist[1:] / hist[:-1]
    assert hist[-1] <= 1e-8 * hist[0]
    assert factors.max() < 0.35, factors


@pytest.mark.parametrize("order", [1, 2, 3, 5])
def test_chebyshev_step_is_its_error_polynomial(order):
    """one smoother step maps the error e = x - x* to P(D^-1 A) e with
    P(t) = T_p((theta - t) / delta) / T_p(theta / delta), theta = (hi + lo) / 2,
    delta = (hi - lo) / 2, on the interval [0.3 rho, 1.2 rho] the smoother is set up
    with (ChebyshevSmoother.hpp:112-232, 246-330; Adams et al. 2003): the step is
    checked on the assembled matrix of the shell's level-0 P2 operator through an
    eigen-decomposition of D^-1/2 A D^-1/2, independently of the coefficients"""
    m = po.Mesh.shell(**problems.SHELL)
    k = po.KCoeff(*problems.shell_coefficient(0))
    op = po.P2Op(m, 0, k)
    fl = m.flags(1)
    inv = op.inv_diag()
    Ai, idx = _restricted_matrix(op, fl, po.GMG_FLAG)
    d = inv[idx]
    S = np.sqrt(d)[:, None] * Ai.toarray() * np.sqrt(d)[None, :]
    lam, V = np.linalg.eigh(S)
    rho = lam.max()
    lo, hi = 0.3 * rho, 1.2 * rho
    theta, delta = 0.5 * (hi + lo), 0.5 * (hi - lo)

    def cheb(p, t):
        return np.cos(p * np.arccos(t)) if abs(t) <= 1 else np.cosh(p * np.arccosh(abs(t))) * np.sign(t) ** p

    P = np.array([cheb(order, (theta - t) / delta) for t in lam]) / cheb(order, theta / delta)
    rng = np.random.default_rng(60 + order)
    xs = np.zeros(op.n)
    xs[idx] = rng.uniform(-1, 1, len(idx))
    b = np.zeros(op.n)
    b[idx] = Ai @ xs[idx]
    x = np.zeros(op.n)
    x[idx] = rng.uniform(-1, 1, len(idx))
    e0 = (x - xs)[idx]
    op.chebyshev_smooth(inv, po.chebyshev_coefficients(order, lo, hi), x, b)
    # D^-1 A = D^-1/2 S D^1/2, so P(D^-1 A) e = D^1/2 V P(Lambda) V^T D^-1/2 e
    e1 = np.sqrt(d) * (V @ (P * (V.T @ (e0 / np.sqrt(d)))))
    assert np.abs((x - xs)[idx] - e1).max() < 1e-10 * np.abs(e0).max()
    assert np.array_equal(x[~np.isin(np.arange(op.n), idx)], xs[~np.isin(np.arange(op.n), idx)])
