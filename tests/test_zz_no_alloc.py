"""The allocation-free dot product through the C ABI: the same results as
hg_function_dot_batched, bit for bit, on calls in a row (each later one with a
smaller input than the one before), and a request beyond the scratch array is
refused."""
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
    # level, vectors, value type, flag: each case smaller than the one before
    cases = [(6, 8, hg.F64, hg.INNER), (5, 4, hg.F32, hg.ALL), (4, 2, hg.F16, hg.GMG_FLAG), (3, 1, hg.F64, hg.ALL)]
    for L, nv, dt, flag in cases:
        print(f"noalloc dot: level {L}, {nv} vectors, dtype {dt}, flag {flag}", flush=True)
        f = hg.P1Function("f", st, L, L, dtype=dt, nvec=nv)
        g = hg.P1Function("g", st, L, L, dtype=dt, nvec=nv)
        f.interpolate_random(5, -1.0, 1.0, L)
        g.interpolate_random(9, -1.0, 1.0, L)
        b = _dots(ctx, f, g, L, flag, nv, lib().hg_function_dot_batched_noalloc)
        a = _dots(ctx, f, g, L, flag, nv, lib().hg_function_dot_batched)
        b2 = _dots(ctx, f, g, L, flag, nv, lib().hg_function_dot_batched_noalloc)
        assert np.array_equal(a, b), (a, b)
        assert np.array_equal(a, b2), (a, b2)
        F = f.download(L).astype(np.float64).reshape(-1, nv)
        G = g.download(L).astype(np.float64).reshape(-1, nv)
        m = (st.flags(L) & flag) != 0
        ref = (F[m] * G[m]).sum(axis=0)
        assert np.abs(b - ref).max() <= 1e-12 * np.abs(F[m] * G[m]).sum(axis=0).max() + 1e-300
        f.destroy()
        g.destroy()
