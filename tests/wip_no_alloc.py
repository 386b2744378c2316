"""The allocation-free dot product: same results as hg_function_dot_batched, bit
for bit, on two calls in a row (the second with a smaller input), through the
C ABI."""
import ctypes as C

import numpy as np
import pytest

import hyteg_amd as hg
from hyteg_amd._lib import lib

pytestmark = pytest.mark.gpu


def _dots(ctx, f, g, level, flag, nv, fn):
    out = np.zeros(nv)
    st = fn(ctx.h, f.h, g.h, level, flag, out.ctypes.data_as(C.c_void_p))
    assert st == 0, lib().hg_last_error(ctx.h)
    return out


def test_noalloc_dot_equals_the_allocating_dot(ctx):
    st = hg.PrimitiveStorage.cuboid(ctx)
    cases = [(5, 8, hg.F64, hg.INNER), (3, 1, hg.F64, hg.ALL), (4, 2, hg.F16, hg.GMG_FLAG)]
    for L, nv, dt, flag in cases:
        print(f"noalloc dot: level {L}, {nv} vectors, dtype {dt}, flag {flag}", flush=True)
        f = hg.P1Function("f", st, L, L, dtype=dt, nvec=nv)
        g = hg.P1Function("g", st, L, L, dtype=dt, nvec=nv)
        f.interpolate_random(5, -1.0, 1.0, L)
        g.interpolate_random(9, -1.0, 1.0, L)
        a = _dots(ctx, f, g, L, flag, nv, lib().hg_function_dot_batched)
        b = _dots(ctx, f, g, L, flag, nv, lib().hg_function_dot_batched_noalloc)
        assert np.array_equal(a, b), (a, b)
        F, G = f.download(L).reshape(-1, nv), g.download(L).reshape(-1, nv)
        m = (st.flags(L) & flag) != 0
        assert np.abs(b - (F[m] * G[m]).sum(axis=0)).max() <= 1e-12 * np.abs(F[m] * G[m]).sum(axis=0).max()
        f.destroy()
        g.destroy()
