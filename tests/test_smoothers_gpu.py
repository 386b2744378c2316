"""Smoother and operator interfaces of the P1 and P2 paths against the oracle,
through the C ABI:

* the multicolour Gauss-Seidel / SOR sweeps (forward and backward, relax 1 and
  1.3) equal the oracle's sweep in the same colour order (<= 1e-12), and the
  colour-conflict checker finds no coupled pair inside a colour of the GPU's
  colouring (so updating a colour's rows in parallel is the sequential sweep);
* inverse diagonal, smooth_jac and gemv / applyScaled equal their definitions on
  the oracle's assembled matrix or apply;
* the P1 and P2 Gauss-Seidel multigrid cycles reproduce the residual factors of
  an oracle V-cycle that sweeps in the GPU's colour order, every cycle.
"""
import numpy as np
import pytest

import hyteg_amd as hg
from hyteg_amd import problems
from meshes import octahedron, relmax
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-12


def p1_meshes(ctx, name):
    if name == "cube":
        return hg.PrimitiveStorage.cuboid(ctx), po.Mesh.cuboid()
    if name == "octa":
        X, cells = octahedron()
    else:  # the shell's macro mesh, affine (the constant-stencil operator ignores blending)
        sh = po.Mesh.shell(**problems.SHELL)
        X, cells = sh.coords(0), sh.cell_vertices()
    return hg.PrimitiveStorage.from_cells(ctx, X, cells), po.Mesh.from_arrays(X, cells)


P1_CASES = [("cube", 3), ("cube", 4), ("octa", 3), ("shell", 2)]


@pytest.mark.parametrize("name,L", P1_CASES)
def test_p1_colouring_has_no_conflicts(ctx, name, L):
    st, om = p1_meshes(ctx, name)
    A = hg.P1ConstantLaplaceOperator(st, L, L)
    colour, nbc = A.sor_colours(L)
    Acsr = po.P1Op(om, L).csr()
    print(f"P1 colouring {name} L{L}: {nbc} boundary colours", flush=True)
    assert colour.min() == 0 and colour.max() == nbc + 7
    assert (colour[: om.nbnd(L)] < nbc).all() and (colour[om.nbnd(L):] >= nbc).all()
    assert po.colour_conflicts(Acsr, colour) == []


@pytest.mark.parametrize("name,L", P1_CASES)
def test_p1_sor_sweep_parity(ctx, name, L):
    st, om = p1_meshes(ctx, name)
    A = hg.P1ConstantLaplaceOperator(st, L, L)
    colour, _ = A.sor_colours(L)
    Acsr = po.P1Op(om, L).csr()
    n = st.ndofs(L)
    fl = om.flags(L)
    rng = np.random.default_rng(200 + L)
    x0, b = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    x, fb = hg.P1Function("x", st, L, L), hg.P1Function("b", st, L, L)
    fb.upload(L, b)
    for flag in (hg.GMG_FLAG, hg.ALL):
        for relax in (1.0, 1.3):
            for backward in (False, True):
                print(f"P1 SOR {name} L{L} flag {flag} relax {relax} backward {backward}", flush=True)
                x.upload(L, x0)
                if backward:
                    A.smooth_sor_backwards(x, fb, relax, L, flag)
                else:
                    A.smooth_sor(x, fb, relax, L, flag)
                ref = po.sor_multicolour(Acsr, colour, (fl & flag) != 0, x0.copy(), b, relax, backward)
                assert relmax(x.download(L), ref) < TOL, (flag, relax, backward)
    x.destroy()
    fb.destroy()


@pytest.mark.parametrize("name,L", [("cube", 4), ("octa", 3)])
def test_p1_inverse_diagonal_jacobi_and_gemv(ctx, name, L):
    st, om = p1_meshes(ctx, name)
    A = hg.P1ConstantLaplaceOperator(st, L, L)
    op = po.P1Op(om, L)
    Acsr = op.csr()
    n = st.ndofs(L)
    fl = om.flags(L)
    rng = np.random.default_rng(300 + L)
    u, r, y0 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    fu, fr, fd, fi = (hg.P1Function(s, st, L, L) for s in "urdi")
    fu.upload(L, u)
    fr.upload(L, r)
    A.inverse_diagonal(L, fi)
    inv = 1.0 / Acsr.diagonal()
    assert relmax(fi.download(L), inv) < 1e-14
    Au = Acsr @ u
    for flag in (hg.GMG_FLAG, hg.ALL, hg.INNER):
        m = (fl & flag) != 0
        print(f"P1 smooth_jac / gemv {name} L{L} flag {flag}", flush=True)
        fd.upload(L, y0)
        A.smooth_jac(fd, fr, fu, 0.8, L, flag)
        ref = np.where(m, u + 0.8 * inv * (r - Au), y0)
        assert relmax(fd.download(L), ref) < TOL
        # the composition the reference writes (apply, assign, multElementwise, assign)
        # gives the same numbers bit for bit
        fd.upload(L, y0)
        A.apply(fu, fd, L, flag)
        fd.assign([1.0, -1.0], [fr, fd], L, flag)
        fd.mult(fi, fd, L, flag)
        fd.assign([1.0, 0.8], [fu, fd], L, flag)
        got = fd.download(L)
        fd.upload(L, y0)
        A.smooth_jac(fd, fr, fu, 0.8, L, flag)
        assert np.array_equal(fd.download(L), got)
        for alpha, beta in ((0.7, 0.0), (0.7, 1.0), (-1.3, -0.3), (1.0, 2.0)):
            fd.upload(L, y0)
            A.gemv(alpha, fu, beta, fd, L, flag)
            ref = np.where(m, alpha * Au + beta * y0, y0)
            assert relmax(fd.download(L), ref) < TOL, (alpha, beta)
        fd.upload(L, y0)
        A.apply_scaled(2.5, fu, fd, L, flag, hg.ADD)
        assert relmax(fd.download(L), np.where(m, 2.5 * Au + y0, y0)) < TOL
    for f in (fu, fr, fd, fi):
        f.destroy()


class _MulticolourP1GMG:
    """V(pre, post) cycles of GeometricMultigridSolver with Gauss-Seidel smoothing
    (GeometricMultigridSolver.hpp:180-308), built from the oracle's apply,
    restriction, prolongation and CG, sweeping in the given colour orders"""

    def __init__(self, mesh, minL, maxL, colours, pre=3, post=3, coarse_tol=1e-16):
        self.m, self.minL, self.pre, self.post, self.coarse_tol = mesh, minL, pre, post, coarse_tol
        self.ops = {L: po.P1Op(mesh, L) for L in range(minL, maxL + 1)}
        self.A = {L: self.ops[L].csr() for L in self.ops}
        self.mask = {L: (mesh.flags(L) & po.GMG_FLAG) != 0 for L in self.ops}
        self.colours = colours

    def residual(self, L, x, b):
        r = np.zeros_like(x)
        m = self.mask[L]
        r[m] = b[m] - (self.A[L] @ x)[m]
        return r

    def cycle(self, L, x, b):
        if L == self.minL:
            self.ops[L].cg(x, b, po.GMG_FLAG, 10000, self.coarse_tol, self.coarse_tol)
            return
        for _ in range(self.pre):
            po.sor_multicolour(self.A[L], self.colours[L], self.mask[L], x, b)
        r = self.residual(L, x, b)
        cb = np.zeros(self.ops[L - 1].n)
        self.m.restrict(L, r, cb, po.GMG_FLAG)
        cx = np.zeros_like(cb)
        self.cycle(L - 1, cx, cb)
        self.m.prolongate(L - 1, cx, x, po.GMG_FLAG, True)
        for _ in range(self.post):
            po.sor_multicolour(self.A[L], self.colours[L], self.mask[L], x, b)


def test_p1_gmg_factors_equal_the_multicolour_oracle_every_cycle(ctx):
    """cube, level 5, V(3,3), CG on level 0, RHS U[-1,1] on the Inner DoFs
    (seed 0): the GPU cycle's residual factors equal those of the oracle V-cycle
    sweeping in the GPU's colour order, from the first cycle on (the reference's
    lexicographic ordering is compared separately, within 5 % from the second
    cycle, test_p1_gpu.py)"""
    st, om = p1_meshes(ctx, "cube")
    L = 5
    A = hg.P1ConstantLaplaceOperator(st, 0, L)
    colours = {lv: A.sor_colours(lv)[0] for lv in range(0, L + 1)}
    fl = om.flags(L)
    b = np.where(fl == hg.INNER, np.random.default_rng(0).uniform(-1, 1, len(fl)), 0.0)
    x, fb = hg.P1Function("x", st, 0, L), hg.P1Function("b", st, 0, L)
    fb.upload(L, b)
    print("P1 GMG cube L5, multicolour oracle", flush=True)
    k, hist = hg.gmg_solve(A, x, fb, 0, L, 3, 3, cycles=10)
    g = _MulticolourP1GMG(om, 0, L, colours)
    xo, ho = np.zeros_like(b), []
    for _ in range(10):
        ho.append(np.linalg.norm(g.residual(L, xo, b)[fl == hg.INNER]))
        g.cycle(L, xo, b)
    ho.append(np.linalg.norm(g.residual(L, xo, b)[fl == hg.INNER]))
    fg, fo = hist[1:] / hist[:-1], np.array(ho[1:]) / np.array(ho[:-1])
    print("gpu", np.round(fg, 6).tolist(), "oracle", np.round(fo, 6).tolist(), flush=True)
    assert np.abs(fg / fo - 1).max() < 1e-6
    assert relmax(x.download(L), xo) < 1e-9


def test_p1_gmg_coarse_tolerance_with_and_without_graphs(ctx):
    """with a coarse level above 0 (cube, levels 2-5), the V-cycles replayed from a
    HIP graph stop their coarse CG at coarseTol like the eager cycles (the stopping
    test runs on the device): both give the same residuals and iterate bit for bit,
    and the factors equal the multicolour oracle's, every cycle"""
    st, om = p1_meshes(ctx, "cube")
    minL, L = 2, 5
    A = hg.P1ConstantLaplaceOperator(st, minL, L)
    fl = om.flags(L)
    b = np.where(fl == hg.INNER, np.random.default_rng(3).uniform(-1, 1, len(fl)), 0.0)
    out = []
    for graphs in (True, False):
        print(f"P1 GMG levels {minL}-{L}, graphs {graphs}", flush=True)
        ctx.set_graphs(graphs)
        x, fb = hg.P1Function("x", st, minL, L), hg.P1Function("b", st, minL, L)
        fb.upload(L, b)
        k, hist = hg.gmg_solve(A, x, fb, minL, L, 3, 3, cycles=8, coarse_max_iter=10000, coarse_tol=1e-12)
        out.append((hist, x.download(L)))
        x.destroy()
        fb.destroy()
    ctx.set_graphs(True)
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    colours = {lv: A.sor_colours(lv)[0] for lv in range(minL, L + 1)}
    g = _MulticolourP1GMG(om, minL, L, colours, coarse_tol=1e-12)
    xo, ho = np.zeros_like(b), []
    for _ in range(8):
        ho.append(np.linalg.norm(g.residual(L, xo, b)[fl == hg.INNER]))
        g.cycle(L, xo, b)
    ho.append(np.linalg.norm(g.residual(L, xo, b)[fl == hg.INNER]))
    hist = out[0][0]
    fg, fo = hist[1:] / hist[:-1], np.array(ho[1:]) / np.array(ho[:-1])
    print("gpu", np.round(fg, 6).tolist(), "oracle", np.round(fo, 6).tolist(), flush=True)
    assert np.abs(fg / fo - 1).max() < 1e-6


# ---- P2 -------------------------------------------------------------------------------


def p2c_meshes(ctx, name):
    if name == "cube":
        return hg.PrimitiveStorage.cuboid(ctx), po.Mesh.cuboid()
    X, cells = octahedron()
    return hg.PrimitiveStorage.from_cells(ctx, X, cells), po.Mesh.from_arrays(X, cells)


@pytest.mark.parametrize("name,L", [("cube", 2), ("cube", 3), ("octa", 2)])
def test_p2_constant_colouring_and_sweep_parity(ctx, name, L):
    st, om = p2c_meshes(ctx, name)
    A = hg.P2ConstantLaplaceOperator(st, L, L)
    colour, nbc = A.sor_colours(L)
    M = po.P2Op(om, L).matrix()
    print(f"P2 constant colouring {name} L{L}: {nbc} boundary colours", flush=True)
    assert po.colour_conflicts(M, colour) == []
    n = M.shape[0]
    fl = om.flags(L + 1)
    rng = np.random.default_rng(400 + L)
    x0, b = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    x, fb = hg.P2Function("x", st, L, L), hg.P2Function("b", st, L, L)
    fb.upload(L, b)
    for flag in (hg.GMG_FLAG, hg.ALL):
        for relax in (1.0, 1.3):
            for backward in (False, True):
                x.upload(L, x0)
                if backward:
                    A.smooth_sor_backwards(x, fb, L, flag, relax)
                else:
                    A.smooth_sor(x, fb, L, flag, relax)
                ref = po.sor_multicolour(M, colour, (fl & flag) != 0, x0.copy(), b, relax, backward)
                assert relmax(x.download(L), ref) < TOL, (flag, relax, backward)
    x.destroy()
    fb.destroy()


def test_p2_constant_gmg_factors_equal_the_multicolour_oracle(ctx):
    """P2GMG3DConvergenceTest's set-up (octahedron, levels 0-3, V(2,2), CG on level 0,
    f = 0, random interior start, sin(x) sinh(y) z on the boundary): GPU factors equal
    those of the oracle cycle sweeping in the GPU's colour order, every cycle"""
    st, om = p2c_meshes(ctx, "octa")
    minL, maxL = 0, 3
    A = hg.P2ConstantLaplaceOperator(st, minL, maxL)
    colours = {lv: A.sor_colours(lv)[0] for lv in range(minL, maxL + 1)}
    P = om.coords(maxL + 1)
    fl = om.flags(maxL + 1)
    u0 = np.where(fl == po.INNER, np.random.default_rng(0).uniform(0, 1, len(P)),
                  np.sin(P[:, 0]) * np.sinh(P[:, 1]) * P[:, 2])
    X, B = hg.P2Function("x", st, minL, maxL), hg.P2Function("b", st, minL, maxL)
    X.upload(maxL, u0)
    B.upload(maxL, np.zeros(len(u0)))
    k, hist = hg.p2_const_gmg_solve(A, X, B, minL, maxL, 2, 2, cycles=4)
    g = po.P2GSGMG(om, minL, maxL, colours=colours)
    u, inner, ho = u0.copy(), fl == po.INNER, []
    for _ in range(4):
        ho.append(np.linalg.norm(g.residual(maxL, u, np.zeros_like(u))[inner]))
        g.cycle(maxL, u, np.zeros_like(u))
    ho.append(np.linalg.norm(g.residual(maxL, u, np.zeros_like(u))[inner]))
    fg, fo = hist[1:] / hist[:-1], np.array(ho[1:]) / np.array(ho[:-1])
    print("gpu", fg, "oracle", fo, flush=True)
    assert np.abs(fg / fo - 1).max() < 1e-6


def test_p2_gemv_and_jacobi_on_the_shell(ctx):
    st = hg.PrimitiveStorage.shell(ctx, **problems.SHELL)
    om = po.Mesh.shell(**problems.SHELL)
    kc = problems.shell_coefficient(0)
    L = 2
    A = hg.P2DivKGradOperator(st, L, L, *kc)
    op = po.P2Op(om, L, po.KCoeff(*kc))
    n = op.n
    fl = om.flags(L + 1)
    rng = np.random.default_rng(500)
    u, r, y0 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    fu, fr, fd = (hg.P2Function(s, st, L, L) for s in "urd")
    fu.upload(L, u)
    fr.upload(L, r)
    Au = np.zeros(n)
    op.apply(u, Au, po.ALL)
    inv = op.inv_diag()
    for flag in (hg.GMG_FLAG, hg.ALL):
        m = (fl & flag) != 0
        for alpha, beta in ((0.7, 0.0), (0.7, 1.0), (-1.3, -0.3)):
            print(f"P2 gemv flag {flag} alpha {alpha} beta {beta}", flush=True)
            fd.upload(L, y0)
            A.gemv(alpha, fu, beta, fd, L, flag)
            ref = np.where(m, alpha * Au + beta * y0, y0)
            assert relmax(fd.download(L), ref) < TOL, (flag, alpha, beta)
        fd.upload(L, y0)
        A.smooth_jac(fd, fr, fu, 0.6, L, flag)
        ref = np.where(m, u + 0.6 * inv * (r - Au), y0)
        assert relmax(fd.download(L), ref) < TOL
    for f in (fu, fr, fd):
        f.destroy()
