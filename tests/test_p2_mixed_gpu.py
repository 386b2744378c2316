"""Reduced-precision P2 path (FP32 arithmetic, FP32 / FP16 storage, FP32
accumulation) and the FP64 iterative refinement around it:

* on tile levels (>= 2) the FP32 / FP16 applies use the reduced element form
  (affine micro-cells of the blended vertices, k constant per micro-cell, 4-point
  rule), checked against the oracle's restatement of that form; below, and with
  the form switched off, the reference form in FP32 arithmetic, checked against
  the FP64 HIP path (which the other GPU tests pin to the oracle);
* the refinement reaches the oracle's FP64 solution."""
import numpy as np
import pytest

import hyteg_amd as hg
from hyteg_amd import problems
from meshes import relmax
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

# relative max-norm bounds of one reduced-precision apply against the FP64 result of
# the same element form (FP16: input and output rounded to binary16).  The reduced
# form contracts the rounded FP32 metric with barycentric differences whose terms
# cancel to about 1/100 of their size, so its FP32 apply lands near 200 units of
# FP32 rounding (measured on the MI355X: 1.2e-5 on shell levels 2 and 3); the
# reference form in FP32 stays within 2e-6
APPLY_TOL = {hg.F32: 2e-6, hg.F16: 4e-3}
REDUCED_TOL = {hg.F32: 4e-5, hg.F16: 4e-3}


@pytest.fixture(scope="module")
def shell(ctx):
    return hg.PrimitiveStorage.shell(ctx, **problems.SHELL)


@pytest.mark.parametrize("dtype", [hg.F32, hg.F16])
@pytest.mark.parametrize("flag,update", [(hg.ALL, hg.REPLACE), (hg.GMG_FLAG, hg.REPLACE), (hg.ALL, hg.ADD),
                                         (hg.DIRICHLET, hg.REPLACE)])
def test_reduced_form_apply_matches_the_oracle(ctx, shell, dtype, flag, update):
    k = problems.shell_coefficient(0)
    om = po.Mesh.shell(**problems.SHELL)
    for L in (2, 3):
        A = hg.P2DivKGradOperator(shell, L, L, *k)
        n = shell.p2_ndofs(L)
        x = problems.canonical_uniform(n, 21)
        y0 = problems.canonical_uniform(n, 22)
        src = hg.P2Function("src", shell, L, L, dtype=dtype)
        dst = hg.P2Function("dst", shell, L, L, dtype=dtype)
        src.upload(L, x)
        dst.upload(L, y0)
        xs, ys = src.download(L), dst.download(L)  # as stored (binary16 for FP16)
        print(f"reduced apply dtype {dtype} level {L} flag {flag} update {update}", flush=True)
        A.apply(src, dst, L, flag, update)
        ref = ys.copy()
        po.P2Op(om, L, po.KCoeff(*k), reduced=True).apply(xs, ref, flag, update == hg.ADD)
        err = relmax(dst.download(L), ref)
        print(f"   relative difference to the oracle's reduced form {err:.2e}", flush=True)
        assert err < REDUCED_TOL[dtype]
        for f in (src, dst):
            f.destroy()
        A.destroy()


@pytest.mark.parametrize("dtype", [hg.F32, hg.F16])
def test_reference_form_in_low_precision_and_transfer(ctx, shell, dtype):
    """level 1 (below the tile levels) and, with the reduced form switched off,
    level 2: the reference form in FP32 arithmetic agrees with FP64"""
    k = problems.shell_coefficient(0)
    for L, switch_off in ((1, False), (2, True)):
        A = hg.P2DivKGradOperator(shell, 1, L, *k)
        if switch_off:
            A.low_precision_form(False)
        n = shell.p2_ndofs(L)
        x = problems.canonical_uniform(n, 21)
        ref = {}
        for dt in (hg.F64, dtype):
            src = hg.P2Function("src", shell, 1, L, dtype=dt)
            dst = hg.P2Function("dst", shell, 1, L, dtype=dt)
            src.upload(L, x)
            A.apply(src, dst, L, hg.ALL)
            if L > 1:
                hg.restrict(src, L, hg.GMG_FLAG)
            ref[dt] = (dst.download(L), src.download(L - 1) if L > 1 else None)
        print(f"dtype {dtype} level {L}: apply {relmax(ref[dtype][0], ref[hg.F64][0]):.2e}", flush=True)
        assert relmax(ref[dtype][0], ref[hg.F64][0]) < APPLY_TOL[dtype]
        if L > 1:
            assert relmax(ref[dtype][1], ref[hg.F64][1]) < APPLY_TOL[dtype]
        if dtype == hg.F32 and switch_off:
            # stored and in-kernel coefficients agree in FP32 as well
            A.store_coefficients(False)
            src = hg.P2Function("src", shell, L, L, dtype=dtype)
            dst = hg.P2Function("dst", shell, L, L, dtype=dtype)
            src.upload(L, x)
            A.apply(src, dst, L, hg.ALL)
            assert relmax(dst.download(L), ref[dtype][0]) < 1e-5


@pytest.mark.parametrize("dtype", [hg.F32, hg.F16])
def test_iterative_refinement_reaches_fp64_solution(ctx, shell, dtype):
    L = 2
    k = problems.shell_coefficient(0)
    A = hg.P2DivKGradOperator(shell, 0, L, *k)
    fl = shell.flags(L + 1)
    b = np.where(fl == hg.INNER, problems.canonical_uniform(len(fl), 0), 0.0)
    fb = hg.P2Function("b", shell, 0, L)
    fb.upload(L, b)
    t = hg.P2Function("t", shell, L, L)
    t.interpolate_random(1, -1.0, 1.0, L, hg.ALL)
    rho = hg.p2_estimate_radius(A, t, L, 20)
    x = hg.P2Function("x", shell, 0, L)
    print(f"IR dtype {dtype}", flush=True)
    k_ir, hist = hg.p2_ir_solve(A, x, fb, 0, L, rho, low_dtype=dtype, max_outer=40, tol=1e-10)
    assert hist[-1] <= 1e-10 * hist[0], hist
    # the oracle's FP64 V-cycles to the same residual reduction
    xo = np.zeros(len(fl))
    ko, ho = po.Mesh.shell(**problems.SHELL).p2_gmg(0, L, po.KCoeff(*k), xo, b, 2, 2, 3, rho, cycles=40, tol=1e-10)
    err = relmax(x.download(L), xo)
    print(f"IR dtype {dtype}: {k_ir} outer steps (oracle FP64 GMG {ko} cycles), error vs oracle {err:.2e}", flush=True)
    assert err < 1e-7
    assert k_ir <= ko + 3
