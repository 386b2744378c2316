"""Parity of the free-slip Stokes path against the CPU oracle, through the C ABI:
the shell with a FreeSlip inner and a no-slip outer surface, the normal
projection (P2ProjectNormalOperator), the projected operator (free-slip
wrapper: P A P, P B^T, B P, pressure mean out of the full apply), and the
free-slip Uzawa smoother, MINRES and GMG of the FS solver layout."""
import numpy as np
import pytest

import hyteg_amd as hg
from hyteg_amd import problems
from meshes import relmax
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-12
VISC = (np.log(1e4) / 0.45, -np.log(1e4) / 0.45)
FLAGS = dict(inner_flag=po.FREESLIP, outer_flag=po.DIRICHLET)


@pytest.fixture(scope="module")
def fs_pair(ctx):
    return (hg.PrimitiveStorage.shell(ctx, **problems.SHELL, **FLAGS), po.Mesh.shell(**problems.SHELL, **FLAGS))


def test_surface_flags_match(ctx, fs_pair):
    st, om = fs_pair
    for L in (0, 1, 3):
        assert np.array_equal(st.flags(L), om.flags(L)), L


@pytest.mark.parametrize("L", [1, 2])
def test_project_normal_parity(ctx, fs_pair, L):
    st, om = fs_pair
    n2 = om.ndofs(L + 1)
    u0 = np.random.default_rng(40 + L).uniform(-1, 1, 3 * n2)
    f = hg.P2P1TaylorHoodFunction("u", st, L, L)
    P = hg.P2ProjectNormalOperator(st)
    for flag in (hg.FREESLIP, hg.FREESLIP | hg.DIRICHLET, hg.DIRICHLET):
        print(f"project normal L{L} flag {flag}", flush=True)
        f.upload(L, u0, np.zeros(om.ndofs(L)))
        P.project(f, L, flag)
        ref = u0.copy()
        om.p2_project_normal(L, ref, flag)
        assert relmax(f.download(L)[0], ref) < 1e-14
    with pytest.raises(hg.HygError):
        P.project(f, L, hg.INNER | hg.FREESLIP)
    f.destroy()


@pytest.mark.parametrize("L", [1, 2])
def test_freeslip_apply_parity(ctx, fs_pair, L):
    st, om = fs_pair
    A = hg.P2P1StokesOperator(st, L, L, *VISC, freeslip=True)
    op = po.StokesOp(om, L, *VISC, freeslip=True)
    rng = np.random.default_rng(50 + L)
    u, p = rng.uniform(-1, 1, 3 * op.n2), rng.uniform(-1, 1, op.n1)
    yu0, yp0 = rng.uniform(-1, 1, 3 * op.n2), rng.uniform(-1, 1, op.n1)
    src = hg.P2P1TaylorHoodFunction("src", st, L, L)
    dst = hg.P2P1TaylorHoodFunction("dst", st, L, L)
    src.upload(L, u, p)
    for mask in (1, 2, 4, 7):
        for flag in (hg.GMG_FLAG, hg.ALL, hg.INNER):
            for update in (hg.REPLACE, hg.ADD):
                print(f"free-slip apply L{L} mask {mask} flag {flag} update {update}", flush=True)
                dst.upload(L, yu0, yp0)
                A.apply(src, dst, L, flag, update, mask)
                ru, rp = yu0.copy(), yp0.copy()
                op.apply(u, p, ru, rp, flag, mask, update == hg.ADD)
                gu, gp = dst.download(L)
                assert relmax(gu, ru) < TOL, (L, mask, flag, update, relmax(gu, ru))
                assert np.abs(gp - rp).max() < TOL * max(np.abs(rp).max(), np.abs(yp0).max()), (L, mask, flag, update)
    assert relmax(src.download(L)[0], u) == 0.0  # the source is not projected in place
    for f in (src, dst):
        f.destroy()
    A.destroy()


def _rhs(om, L, seed):
    rng = np.random.default_rng(seed)
    fl = om.flags(L + 1)
    live = (fl == po.INNER) | (fl == po.FREESLIP)
    bu = np.concatenate([np.where(live, rng.uniform(-1, 1, len(fl)), 0.0) for _ in range(3)])
    om.p2_project_normal(L, bu)
    return bu, np.zeros(om.ndofs(L))


def test_freeslip_radius_uzawa_minres_parity(ctx, fs_pair):
    st, om = fs_pair
    L = 1
    A = hg.P2P1StokesOperator(st, L, L, *VISC, freeslip=True)
    op = po.StokesOp(om, L, *VISC, freeslip=True)
    rng = np.random.default_rng(60)
    x0 = rng.uniform(-1, 1, 3 * op.n2)
    X = hg.P2P1TaylorHoodFunction("x", st, L, L)
    T = hg.P2P1TaylorHoodFunction("t", st, L, L)
    B = hg.P2P1TaylorHoodFunction("b", st, L, L)
    X.upload(L, x0, np.zeros(op.n1))
    print("free-slip radius parity", flush=True)
    rho = hg.stokes_estimate_radius(A, X, T, L, 10)
    rho_o = po.stokes_power_iteration(op, op.inv_diag(), x0.copy(), 10)
    assert abs(rho - rho_o) < 1e-10 * rho_o
    u, p = rng.uniform(-1, 1, 3 * op.n2), rng.uniform(-1, 1, op.n1)
    bu, bp = rng.uniform(-1, 1, 3 * op.n2), rng.uniform(-1, 1, op.n1)
    B.upload(L, bu, bp)
    coeffs = po.chebyshev_coefficients(3, 0.3 * rho_o, 1.2 * rho_o)
    print("free-slip Uzawa step parity", flush=True)
    X.upload(L, u, p)
    hg.stokes_uzawa_smooth(A, X, B, L, rho_o, omega=0.3, velocity_iterations=2, order=3, project=True)
    uo, pr = u.copy(), p.copy()
    po.uzawa_smooth(op, uo, pr, bu, bp, op.inv_diag(), coeffs, po.p1_pspg_inv_diag(om, L), 0.3, 2, po.GMG_FLAG, True)
    gu, gp = X.download(L)
    assert relmax(gu, uo) < 1e-11 and relmax(gp, pr) < 1e-11, (relmax(gu, uo), relmax(gp, pr))
    print("free-slip MINRES parity", flush=True)
    L0 = 0
    A0 = hg.P2P1StokesOperator(st, L0, L0, *VISC, freeslip=True)
    op0 = po.StokesOp(om, L0, *VISC, freeslip=True)
    bu, bp = _rhs(om, L0, 61)
    X0 = hg.P2P1TaylorHoodFunction("x0", st, L0, L0)
    B0 = hg.P2P1TaylorHoodFunction("b0", st, L0, L0)
    X0.upload(L0, np.zeros(3 * op0.n2), np.zeros(op0.n1))
    B0.upload(L0, bu, bp)
    its = hg.stokes_minres_solve(A0, X0, B0, L0, max_iter=40, rel_tol=1e-30)
    u, p = np.zeros(3 * op0.n2), np.zeros(op0.n1)
    its_o = po.stokes_minres(op0, u, p, bu, bp, po.p1_lumped_inv_mass(om, L0), 40, 1e-30)
    gu, gp = X0.download(L0)
    assert its == its_o
    assert relmax(gu, u) < 1e-9 and relmax(gp, p) < 1e-9, (relmax(gu, u), relmax(gp, p))
    for f in (X, T, B, X0, B0):
        f.destroy()
    A.destroy()
    A0.destroy()


def test_freeslip_gmg_parity(ctx, fs_pair):
    """three V(3,3) cycles of the free-slip Uzawa GMG on the shell at level 2 with
    the 1e4 viscosity contrast: residual history within 1 % of the oracle's per
    cycle, and the velocity stays tangential on the FreeSlip surface"""
    st, om = fs_pair
    minL, maxL = 0, 2
    A = hg.P2P1StokesOperator(st, minL, maxL, *VISC, freeslip=True)
    bu, bp = _rhs(om, maxL, 62)
    X = hg.P2P1TaylorHoodFunction("x", st, minL, maxL)
    B = hg.P2P1TaylorHoodFunction("b", st, minL, maxL)
    X.upload(maxL, np.zeros(len(bu)), np.zeros(len(bp)))
    B.upload(maxL, bu, bp)
    rho = 2.9
    print("free-slip Stokes GMG parity", flush=True)
    k, hist = hg.stokes_gmg_solve(A, X, B, minL, maxL, rho, cycles=3)
    g = po.StokesGMG(om, minL, maxL, *VISC, rho=rho, coarse_iter=10, coarse_tol=1e-6, freeslip=True)
    u, p = np.zeros(len(bu)), np.zeros(len(bp))
    ko, ho = g.solve(u, p, bu, bp, cycles=3, tol=0.0)
    assert k == ko == 3
    assert abs(hist[0] / ho[0] - 1) < 1e-12
    f, fo = hist[1:] / hist[:-1], ho[1:] / ho[:-1]
    assert np.all(np.abs(f / fo - 1) < 0.01), (f, fo)
    gu, gp = X.download(maxL)
    assert relmax(gu, u) < 1e-6, relmax(gu, u)
    n2 = om.ndofs(maxL + 1)
    Xc = om.coords(maxL + 1)
    fs = om.flags(maxL + 1) == po.FREESLIP
    U = gu.reshape(3, n2).T
    assert np.abs((U * Xc).sum(1)[fs]).max() < 1e-13 * np.abs(U).max()
    X.destroy()
    B.destroy()
    A.destroy()
