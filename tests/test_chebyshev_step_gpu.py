"""The fused Chebyshev step (hg_function_chebyshev_step, one pass per smoother
step) equals the separate assign / setToZero / multElementwise / add sequence
of ChebyshevSmoother::solve (ChebyshevSmoother.hpp:186-231) bit for bit, for
FP64, FP32 and FP16 storage, first and later steps, and several DoF flags."""
import numpy as np
import pytest

import hyteg_amd as hg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [hg.F64, hg.F32, hg.F16])
def test_fused_step_equals_the_separate_kernels(ctx, dtype):
    st = hg.PrimitiveStorage.shell(ctx, ntan=3, nrad=2, rmin=0.5, rmax=1.0)
    L = 2
    n = st.p2_ndofs(L)
    rng = np.random.default_rng(31)
    names = ("x", "t", "s", "b", "d", "x2", "t2", "s2")
    f = {k: hg.P2Function(k, st, L, L, dtype) for k in names}
    lib = hg.lib()
    for first in (True, False):
        for flag in (hg.GMG_FLAG, hg.ALL, hg.INNER, hg.DIRICHLET):
            print(f"chebyshev step dtype {dtype} first {first} flag {flag}", flush=True)
            vals = {k: rng.uniform(-1, 1, n) for k in ("x", "t", "s", "b", "d")}
            for k in ("x", "t", "s", "b", "d"):
                f[k].upload(L, vals[k])
            for k in ("x", "t", "s"):
                f[k + "2"].upload(L, vals[k])
            c = 0.37
            # separate kernels on (x2, t2, s2)
            if first:
                sc = np.array([1.0, -1.0])
                hs = np.array([f["b"].h, f["s2"].h], dtype=np.int64)
                ctx.check(lib.hg_function_assign(ctx.h, f["s2"].h, 2, sc.ctypes.data, hs.ctypes.data, L, flag), "assign")
            ctx.check(lib.hg_function_interpolate_const(ctx.h, f["t2"].h, 0.0, L, hg.ALL), "zero")
            ctx.check(lib.hg_function_mult(ctx.h, f["t2"].h, f["d"].h, f["s2"].h, L, flag), "mult")
            sc = np.array([c])
            hs = np.array([f["t2"].h], dtype=np.int64)
            ctx.check(lib.hg_function_add(ctx.h, f["x2"].h, 1, sc.ctypes.data, hs.ctypes.data, L, flag), "add")
            # fused
            ctx.check(lib.hg_function_chebyshev_step(ctx.h, f["x"].h, f["t"].h, f["s"].h, f["b"].h if first else 0,
                                                     f["d"].h, c, L, flag), "fused")
            assert np.array_equal(f["x"].download(L), f["x2"].download(L)), (first, flag)
            assert np.array_equal(f["t"].download(L), f["t2"].download(L)), (first, flag)
    for g in f.values():
        g.destroy()
