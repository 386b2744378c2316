"""Parity of the HIP P2-P1 Stokes operator (blended full-viscous velocity block
with eta = exp(c0 + c1 r), divergence and its transpose, velocity inverse
diagonal, PSPG and lumped-mass pressure diagonals) against the CPU oracle,
through the C ABI.  FP64: <= 1e-12 relative max-norm."""
import numpy as np
import pytest

import hyteg_amd as hg
from hyteg_amd import problems
from meshes import relmax
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-12
VISC = (np.log(1e4) / 0.45, -np.log(1e4) / 0.45)


@pytest.fixture(scope="module")
def shell_pair(ctx):
    return hg.PrimitiveStorage.shell(ctx, **problems.SHELL), po.Mesh.shell(**problems.SHELL)


@pytest.fixture(scope="module")
def cube_pair(ctx):
    return hg.PrimitiveStorage.cuboid(ctx), po.Mesh.cuboid()


def _apply_parity(st, om, L, visc):
    A = hg.P2P1StokesOperator(st, L, L, *visc)
    op = po.StokesOp(om, L, *visc)
    rng = np.random.default_rng(20 + L)
    u, p = rng.uniform(-1, 1, 3 * op.n2), rng.uniform(-1, 1, op.n1)
    yu0, yp0 = rng.uniform(-1, 1, 3 * op.n2), rng.uniform(-1, 1, op.n1)
    src = hg.P2P1TaylorHoodFunction("src", st, L, L)
    dst = hg.P2P1TaylorHoodFunction("dst", st, L, L)
    src.upload(L, u, p)
    for mask in (1, 2, 3, 4, 7):
        for flag in (hg.GMG_FLAG, hg.ALL, hg.DIRICHLET):
            for update in (hg.REPLACE, hg.ADD):
                print(f"Stokes apply L{L} mask {mask} flag {flag} update {update}", flush=True)
                dst.upload(L, yu0, yp0)
                A.apply(src, dst, L, flag, update, mask)
                ru, rp = yu0.copy(), yp0.copy()
                op.apply(u, p, ru, rp, flag, mask, update == hg.ADD)
                gu, gp = dst.download(L)
                assert relmax(gu, ru) < TOL, (L, mask, flag, update, relmax(gu, ru))
                assert relmax(gp, rp) < TOL, (L, mask, flag, update, relmax(gp, rp))
    inv = hg.P2P1TaylorHoodFunction("inv", st, L, L)
    A.velocity_inverse_diagonal(L, inv)
    assert relmax(inv.download(L)[0], op.inv_diag()) < TOL
    A.pspg_inverse_diagonal(L, inv.p)
    assert relmax(inv.p.download(L), po.p1_pspg_inv_diag(om, L)) < TOL
    A.lumped_inverse_mass(L, inv.p)
    assert relmax(inv.p.download(L), po.p1_lumped_inv_mass(om, L)) < TOL
    for f in (src, dst, inv):
        f.destroy()
    A.destroy()


@pytest.mark.parametrize("L", [1, 2])
def test_shell_apply_parity(ctx, shell_pair, L):
    _apply_parity(*shell_pair, L, VISC)


@pytest.mark.parametrize("L", [2, 3])
def test_cube_apply_parity(ctx, cube_pair, L):
    _apply_parity(*cube_pair, L, (0.2, 0.7))


def test_pressure_sum_and_constant(ctx, shell_pair):
    st, om = shell_pair
    L = 2
    f = hg.P1Function("p", st, L, L)
    x = np.random.default_rng(5).uniform(-1, 1, st.ndofs(L))
    f.upload(L, x)
    assert abs(f.sum(L) - x.sum()) < 1e-12 * np.abs(x).sum()
    th = hg.P2P1TaylorHoodFunction("th", st, L, L)
    th.p.upload(L, x)
    th.project_mean_pressure(L)
    assert np.abs(th.p.download(L) - (x - x.mean())).max() < 1e-13
    f.destroy()
    th.destroy()


def _stokes_rhs(om, L, seed):
    rng = np.random.default_rng(seed)
    fl = om.flags(L + 1)
    bu = np.concatenate([np.where(fl == po.INNER, rng.uniform(-1, 1, len(fl)), 0.0) for _ in range(3)])
    return bu, np.zeros(om.ndofs(L))


def test_uzawa_step_and_radius_parity(ctx, shell_pair):
    st, om = shell_pair
    L = 1
    A = hg.P2P1StokesOperator(st, L, L, *VISC)
    op = po.StokesOp(om, L, *VISC)
    rng = np.random.default_rng(30)
    x0 = rng.uniform(-1, 1, 3 * op.n2)
    X = hg.P2P1TaylorHoodFunction("x", st, L, L)
    T = hg.P2P1TaylorHoodFunction("t", st, L, L)
    X.upload(L, x0, np.zeros(op.n1))
    print("Stokes radius parity", flush=True)
    rho = hg.stokes_estimate_radius(A, X, T, L, 10)
    xo = x0.copy()
    rho_o = po.stokes_power_iteration(op, op.inv_diag(), xo, 10)
    assert abs(rho - rho_o) < 1e-10 * rho_o
    u, p = rng.uniform(-1, 1, 3 * op.n2), rng.uniform(-1, 1, op.n1)
    bu, bp = rng.uniform(-1, 1, 3 * op.n2), rng.uniform(-1, 1, op.n1)
    B = hg.P2P1TaylorHoodFunction("b", st, L, L)
    B.upload(L, bu, bp)
    coeffs = po.chebyshev_coefficients(3, 0.3 * rho_o, 1.2 * rho_o)
    for project in (False, True):
        print(f"Uzawa step parity project={project}", flush=True)
        X.upload(L, u, p)
        hg.stokes_uzawa_smooth(A, X, B, L, rho_o, omega=0.3, velocity_iterations=2, order=3, project=project)
        uo, po_ = u.copy(), p.copy()
        po.uzawa_smooth(op, uo, po_, bu, bp, op.inv_diag(), coeffs, po.p1_pspg_inv_diag(om, L), 0.3, 2, po.GMG_FLAG,
                        project)
        gu, gp = X.download(L)
        assert relmax(gu, uo) < 1e-11 and relmax(gp, po_) < 1e-11, (relmax(gu, uo), relmax(gp, po_))
    for f in (X, T, B):
        f.destroy()


def test_minres_parity(ctx, shell_pair):
    st, om = shell_pair
    L = 0
    A = hg.P2P1StokesOperator(st, L, L, *VISC)
    op = po.StokesOp(om, L, *VISC)
    bu, bp = _stokes_rhs(om, L, 31)
    X = hg.P2P1TaylorHoodFunction("x", st, L, L)
    B = hg.P2P1TaylorHoodFunction("b", st, L, L)
    X.upload(L, np.zeros(3 * op.n2), np.zeros(op.n1))
    B.upload(L, bu, bp)
    print("MINRES parity", flush=True)
    its = hg.stokes_minres_solve(A, X, B, L, max_iter=40, rel_tol=1e-30)
    u, p = np.zeros(3 * op.n2), np.zeros(op.n1)
    its_o = po.stokes_minres(op, u, p, bu, bp, po.p1_lumped_inv_mass(om, L), 40, 1e-30)
    gu, gp = X.download(L)
    assert its == its_o
    assert relmax(gu, u) < 1e-9 and relmax(gp, p) < 1e-9, (relmax(gu, u), relmax(gp, p))
    X.destroy()
    B.destroy()


def test_gmg_parity_shell(ctx, shell_pair):
    """three V(3,3) cycles of the Uzawa GMG (smoothing increment 2, MINRES on level 0)
    on the shell at level 2 with the 1e4 viscosity contrast: the residual history
    matches the oracle's (each convergence factor within 1 %)"""
    st, om = shell_pair
    minL, maxL = 0, 2
    A = hg.P2P1StokesOperator(st, minL, maxL, *VISC)
    bu, bp = _stokes_rhs(om, maxL, 32)
    X = hg.P2P1TaylorHoodFunction("x", st, minL, maxL)
    B = hg.P2P1TaylorHoodFunction("b", st, minL, maxL)
    X.upload(maxL, np.zeros(len(bu)), np.zeros(len(bp)))
    B.upload(maxL, bu, bp)
    rho = 2.9
    print("Stokes GMG parity", flush=True)
    k, hist = hg.stokes_gmg_solve(A, X, B, minL, maxL, rho, cycles=3)
    g = po.StokesGMG(om, minL, maxL, *VISC, rho=rho, coarse_iter=10, coarse_tol=1e-6)
    u, p = np.zeros(len(bu)), np.zeros(len(bp))
    ko, ho = g.solve(u, p, bu, bp, cycles=3, tol=0.0)
    assert k == ko == 3
    f, fo = hist[1:] / hist[:-1], ho[1:] / ho[:-1]
    assert np.all(np.abs(f / fo - 1) < 0.01), (f, fo)
    gu, gp = X.download(maxL)
    assert relmax(gu, u) < 1e-6, relmax(gu, u)
    X.destroy()
    B.destroy()
