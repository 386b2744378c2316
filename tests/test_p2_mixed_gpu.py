"""Reduced-precision P2 path (FP32 arithmetic, FP32 / FP16 storage) and the FP64
iterative refinement around it, checked against the FP64 HIP path (which the
other GPU tests pin to the oracle) and the oracle's FP64 solution."""
import numpy as np
import pytest

import hyteg_amd as hg
from hyteg_amd import problems
from meshes import relmax
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

# relative max-norm bounds of one reduced-precision apply against FP64
APPLY_TOL = {hg.F32: 2e-6, hg.F16: 4e-3}


@pytest.fixture(scope="module")
def shell(ctx):
    return hg.PrimitiveStorage.shell(ctx, **problems.SHELL)


@pytest.mark.parametrize("dtype", [hg.F32, hg.F16])
def test_low_precision_apply_and_transfer(ctx, shell, dtype):
    L = 2
    A = hg.P2DivKGradOperator(shell, 1, L, *problems.shell_coefficient(0))
    n = shell.p2_ndofs(L)
    x = problems.canonical_uniform(n, 21)
    ref = {}
    for dt in (hg.F64, dtype):
        src = hg.P2Function("src", shell, 1, L, dtype=dt)
        dst = hg.P2Function("dst", shell, 1, L, dtype=dt)
        src.upload(L, x)
        A.apply(src, dst, L, hg.ALL)
        hg.restrict(src, L, hg.GMG_FLAG)
        ref[dt] = (dst.download(L), src.download(L - 1))
    print(f"dtype {dtype}: apply {relmax(ref[dtype][0], ref[hg.F64][0]):.2e}", flush=True)
    assert relmax(ref[dtype][0], ref[hg.F64][0]) < APPLY_TOL[dtype]
    assert relmax(ref[dtype][1], ref[hg.F64][1]) < APPLY_TOL[dtype]
    if dtype == hg.F32:
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
