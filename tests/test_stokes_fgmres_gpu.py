"""Parity of the 1/eta pressure mass (mask 8 of the Stokes apply), of
invertElementwise and of FGMRES with the block-factorisation preconditioner
(velocity GMG V-cycle + CG on the pressure mass) against the CPU oracle, through
the C ABI, on the no-slip and the free-slip shell."""
import numpy as np
import pytest

import hyteg_amd as hg
from hyteg_amd import problems
from meshes import relmax
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

VISC = (np.log(1e4) / 0.45, -np.log(1e4) / 0.45)
FS = dict(inner_flag=po.FREESLIP, outer_flag=po.DIRICHLET)


@pytest.fixture(scope="module")
def shells(ctx):
    return {fs: (hg.PrimitiveStorage.shell(ctx, **problems.SHELL, **(FS if fs else {})),
                 po.Mesh.shell(**problems.SHELL, **(FS if fs else {}))) for fs in (False, True)}


@pytest.mark.parametrize("L", [1, 2])
def test_kmass_and_invert_parity(ctx, shells, L):
    st, om = shells[False]
    A = hg.P2P1StokesOperator(st, L, L, *VISC)
    op = po.StokesOp(om, L, *VISC)
    rng = np.random.default_rng(70 + L)
    p, y0 = rng.uniform(-1, 1, op.n1), rng.uniform(-1, 1, op.n1)
    src = hg.P1Function("p", st, L, L)
    dst = hg.P1Function("y", st, L, L)
    src.upload(L, p)
    for flag in (hg.ALL, hg.INNER, hg.DIRICHLET):
        for update in (hg.REPLACE, hg.ADD):
            print(f"K-mass L{L} flag {flag} update {update}", flush=True)
            dst.upload(L, y0)
            A.kmass_apply(src, dst, L, flag, update)
            ref = y0.copy()
            op.apply_raw(None, p, None, ref, flag, op.K_MASS, update == hg.ADD)
            assert relmax(dst.download(L), ref) < 1e-12, (flag, update)
    v = rng.uniform(0.5, 2.0, op.n1) * np.where(rng.uniform(size=op.n1) < 0.1, 0.0, 1.0)
    dst.upload(L, v)
    ctx.check(hg.lib().hg_function_invert(ctx.h, dst.h, L, hg.ALL), "invert")
    assert relmax(dst.download(L), np.where(v != 0, 1.0 / np.where(v != 0, v, 1.0), 0.0)) < 1e-15
    for f in (src, dst):
        f.destroy()
    A.destroy()


def _rhs(om, L, seed, fs):
    rng = np.random.default_rng(seed)
    fl = om.flags(L + 1)
    live = (fl == po.INNER) | (fl == po.FREESLIP)
    bu = np.concatenate([np.where(live, rng.uniform(-1, 1, len(fl)), 0.0) for _ in range(3)])
    if fs:
        om.p2_project_normal(L, bu)
    return bu, np.zeros(om.ndofs(L))


@pytest.mark.parametrize("fs", [False, True])
def test_fgmres_block_factorisation_parity(ctx, shells, fs):
    """8 FGMRES iterations on the shell at level 2 (1e4 viscosity contrast): the
    approximate residual history matches the oracle's and so does the iterate"""
    st, om = shells[fs]
    minL, maxL = 0, 2
    A = hg.P2P1StokesOperator(st, minL, maxL, *VISC, freeslip=fs)
    bu, bp = _rhs(om, maxL, 71, fs)
    X = hg.P2P1TaylorHoodFunction("x", st, minL, maxL)
    B = hg.P2P1TaylorHoodFunction("b", st, minL, maxL)
    X.upload(maxL, np.zeros(len(bu)), np.zeros(len(bp)))
    B.upload(maxL, bu, bp)
    rho = 2.9
    kw = dict(a_pre=3, a_post=3, order=3, a_coarse_iter=5, a_coarse_tol=1e-6, schur_iter=5, schur_tol=1e-6)
    print(f"FGMRES parity freeslip={fs}", flush=True)
    its, hist = hg.stokes_fgmres_solve(A, X, B, minL, maxL, rho, max_iter=8, restart=50, tol=1e-30, **kw)
    s = po.StokesFGMRES(om, minL, maxL, *VISC, rho=rho, order=3, a_pre=3, a_post=3, a_coarse_iter=5,
                        a_coarse_tol=1e-6, schur_iter=5, schur_tol=1e-6, freeslip=fs)
    u, p = np.zeros(len(bu)), np.zeros(len(bp))
    its_o, ho = s.solve(u, p, bu, bp, max_iter=8, restart=50, tol=1e-30)
    assert its == its_o == 8
    assert np.all(np.abs(hist / ho - 1) < 1e-6), (hist, ho)
    gu, gp = X.download(maxL)
    assert relmax(gu, u) < 1e-6 and relmax(gp, p) < 1e-6, (relmax(gu, u), relmax(gp, p))
    X.destroy()
    B.destroy()
    A.destroy()
