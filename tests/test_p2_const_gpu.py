"""P2ConstantLaplaceOperator on the GPU against the oracle, through the C ABI:
apply and inverse diagonal (FP64, <= 1e-12 relative max-norm, against the P2
Laplacian assembled by the oracle on affine meshes), the multicoloured
Gauss-Seidel sweep (fixed point, energy decrease) and the P2 GMG of
P2GMG3DConvergenceTest (its own rate bound)."""
import numpy as np
import pytest
from scipy.sparse.linalg import spsolve

import hyteg_amd as hg
from meshes import octahedron, relmax
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def meshes(ctx):
    X, cells = octahedron()
    return {"cube": (hg.PrimitiveStorage.cuboid(ctx), po.Mesh.cuboid()),
            "octahedron": (hg.PrimitiveStorage.from_cells(ctx, X, cells), po.Mesh.from_arrays(X, cells))}


@pytest.mark.parametrize("name,L", [("cube", 2), ("cube", 3), ("octahedron", 2)])
def test_apply_and_diagonal_parity(ctx, meshes, name, L):
    st, om = meshes[name]
    A = hg.P2ConstantLaplaceOperator(st, L, L)
    op = po.P2Op(om, L)
    n = op.n
    rng = np.random.default_rng(90 + L)
    u, y0 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    src, dst = hg.P2Function("s", st, L, L), hg.P2Function("d", st, L, L)
    src.upload(L, u)
    for flag in (hg.ALL, hg.INNER, hg.DIRICHLET):
        for update in (hg.REPLACE, hg.ADD):
            print(f"P2 constant apply {name} L{L} flag {flag} update {update}", flush=True)
            dst.upload(L, y0)
            A.apply(src, dst, L, flag, update)
            ref = y0.copy()
            op.apply(u, ref, flag, update == hg.ADD)
            assert relmax(dst.download(L), ref) < 1e-12, (flag, update)
    A.inverse_diagonal(L, dst)
    assert relmax(dst.download(L), 1.0 / op.matrix().diagonal()) < 1e-12
    for f in (src, dst):
        f.destroy()
    A.destroy()


@pytest.mark.parametrize("name", ["cube", "octahedron"])
def test_gauss_seidel_fixed_point_and_energy(ctx, meshes, name):
    """the discrete solution is a fixed point of the sweep, and every sweep lowers the
    energy norm of the error (Gauss-Seidel on an SPD system)"""
    st, om = meshes[name]
    L = 2
    A = hg.P2ConstantLaplaceOperator(st, L, L)
    M = po.P2Op(om, L).matrix()
    inner = om.flags(L + 1) == po.INNER
    rng = np.random.default_rng(91)
    b = np.where(inner, rng.uniform(-1, 1, len(inner)), 0.0)
    xs = np.zeros_like(b)
    Mi = M[inner][:, inner].tocsc()
    xs[inner] = spsolve(Mi, b[inner])
    X, B = hg.P2Function("x", st, L, L), hg.P2Function("b", st, L, L)
    B.upload(L, b)
    X.upload(L, xs)
    print(f"P2 Gauss-Seidel fixed point {name}", flush=True)
    A.smooth_gs(X, B, L, hg.INNER)
    assert relmax(X.download(L), xs) < 1e-11
    x = np.where(inner, rng.uniform(-1, 1, len(inner)), 0.0)
    X.upload(L, x)
    energy = []
    for _ in range(4):
        e = (X.download(L) - xs)[inner]
        energy.append(e @ (Mi @ e))
        A.smooth_gs(X, B, L, hg.INNER)
    e = (X.download(L) - xs)[inner]
    energy.append(e @ (Mi @ e))
    assert all(energy[k + 1] < energy[k] for k in range(4)), energy
    for f in (X, B):
        f.destroy()
    A.destroy()


def test_p2_gmg_meets_the_reference_rate(ctx, meshes):
    """P2GMG3DConvergenceTest.cpp:60-140 on the GPU: octahedron, levels 0-3, V(2,2)
    multicoloured Gauss-Seidel, CG on level 0: every cycle reduces the squared Inner
    residual below 3e-2 of its previous value, and from the second cycle on within
    15 % of the rate of the oracle's lexicographic sweep"""
    st, om = meshes["octahedron"]
    minL, maxL = 0, 3
    A = hg.P2ConstantLaplaceOperator(st, minL, maxL)
    P = om.coords(maxL + 1)
    fl = om.flags(maxL + 1)
    u0 = np.where(fl == po.INNER, np.random.default_rng(0).uniform(0, 1, len(P)),
                  np.sin(P[:, 0]) * np.sinh(P[:, 1]) * P[:, 2])
    X, B = hg.P2Function("x", st, minL, maxL), hg.P2Function("b", st, minL, maxL)
    X.upload(maxL, u0)
    B.upload(maxL, np.zeros(len(u0)))
    print("P2 constant GMG", flush=True)
    k, hist = hg.p2_const_gmg_solve(A, X, B, minL, maxL, 2, 2, cycles=4)
    rates = (hist[1:] / hist[:-1]) ** 2
    g = po.P2GSGMG(om, minL, maxL)
    u, inner, last, ro = u0.copy(), fl == po.INNER, None, []
    for _ in range(4):
        r = g.residual(maxL, u, np.zeros_like(u))
        last = r[inner] @ r[inner] if last is None else last
        g.cycle(maxL, u, np.zeros_like(u))
        r = g.residual(maxL, u, np.zeros_like(u))
        ro.append((r[inner] @ r[inner]) / last)
        last = r[inner] @ r[inner]
    print("GPU squared rates", rates, "oracle", ro, flush=True)
    assert k == 4 and np.all(rates < 3e-2), rates
    # after the first cycle the multicoloured sweep smooths like the lexicographic one
    assert np.all(np.abs(rates[1:] / np.array(ro[1:]) - 1) < 0.15), (rates, ro)
    for f in (X, B):
        f.destroy()
    A.destroy()
