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


@pytest.mark.parametrize("dtype,tol", [(hg.F32, 2e-5), (hg.F16, 4e-3)])
def test_low_precision_velocity_block(ctx, shells, dtype, tol):
    """the velocity block on FP32 / FP16 vectors (FP32 arithmetic) against the FP64
    oracle, with and without the free-slip projection (the tolerance covers the
    rounding of the inputs to the storage type and FP32 arithmetic)"""
    for fs in (False, True):
        st, om = shells[fs]
        L = 2
        A = hg.P2P1StokesOperator(st, L, L, *VISC)
        op = po.StokesOp(om, L, *VISC, freeslip=fs)
        u = np.random.default_rng(80).uniform(-1, 1, 3 * op.n2)
        src = [hg.P2Function(f"s{c}", st, L, L, dtype) for c in range(3)]
        dst = [hg.P2Function(f"d{c}", st, L, L, dtype) for c in range(3)]
        for c in range(3):
            src[c].upload(L, u[c * op.n2:(c + 1) * op.n2])
        uq = np.concatenate([f.download(L) for f in src])  # the rounded input
        print(f"low-precision velocity block dtype {dtype} freeslip {fs}", flush=True)
        hs = np.ascontiguousarray([f.h for f in src] + [0], dtype=np.int64)
        hd = np.ascontiguousarray([f.h for f in dst] + [0], dtype=np.int64)
        fn = hg.lib().hg_stokes_freeslip_apply if fs else hg.lib().hg_stokes_apply
        if fs:
            # the wrapper runs FP64 only; the reduced-precision P A P is checked through its pieces
            P = hg.P2ProjectNormalOperator(st)
            P.project(src, L)
            om.p2_project_normal(L, uq)
            fn = hg.lib().hg_stokes_apply
        ctx.check(fn(ctx.h, A.h, hs.ctypes.data, hd.ctypes.data, L, hg.GMG_FLAG, 1, hg.REPLACE), "apply")
        if fs:
            P.project(dst, L)
        got = np.concatenate([f.download(L) for f in dst])
        ref = np.zeros_like(uq)
        op.apply_raw(uq, None, ref, None, hg.GMG_FLAG, 1)
        if fs:
            om.p2_project_normal(L, ref)
        m = op.vmask(hg.GMG_FLAG)
        assert np.abs(got[m] - ref[m]).max() < tol * np.abs(ref[m]).max(), (dtype, fs)
        for f in src + dst:
            f.destroy()
        A.destroy()


@pytest.mark.parametrize("dtype", [hg.F32, hg.F16])
def test_mixed_precision_fgmres_reaches_fp64_accuracy(ctx, shells, dtype):
    """FGMRES (FP64) with the velocity V-cycle in FP32 arithmetic on FP32 / FP16
    vectors reaches a 1e-8 relative residual on the free-slip shell at level 2
    (constant viscosity), and its (projected) solution matches the FP64 path's to 1e-6"""
    st, om = shells[True]
    minL, maxL = 0, 2
    A = hg.P2P1StokesOperator(st, minL, maxL, freeslip=True)
    bu, bp = _rhs(om, maxL, 81, True)
    X = hg.P2P1TaylorHoodFunction("x", st, minL, maxL)
    B = hg.P2P1TaylorHoodFunction("b", st, minL, maxL)
    B.upload(maxL, bu, bp)
    sols = {}
    for vd in (hg.F64, dtype):
        X.upload(maxL, np.zeros(len(bu)), np.zeros(len(bp)))
        print(f"mixed FGMRES velocity dtype {vd}", flush=True)
        its, hist = hg.stokes_fgmres_solve(A, X, B, minL, maxL, 2.9, max_iter=60, restart=60, tol=1e-8,
                                           velocity_dtype=vd)
        assert its < 60 and hist[-1] <= 1e-8 * hist[0], (vd, its, hist[-1] / hist[0])
        sols[vd] = X.download(maxL)
    # the normal velocity on the FreeSlip surface is not seen by the projected
    # operator (the FP64 V-cycle's Chebyshev corrections leave some there): compare P u
    for vd in sols:
        om.p2_project_normal(maxL, sols[vd][0])
    assert relmax(sols[dtype][0], sols[hg.F64][0]) < 1e-6
    # the true residual of the mixed solution, FP64 oracle
    op = po.StokesOp(om, maxL, freeslip=True)
    ru, rp = po.stokes_residual(op, sols[dtype][0], sols[dtype][1], bu, bp, po.GMG_FLAG)
    assert np.sqrt(ru @ ru + rp @ rp) < 2e-8 * np.sqrt(bu @ bu)
    X.destroy()
    B.destroy()
    A.destroy()
