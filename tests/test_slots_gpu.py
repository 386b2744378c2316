"""Device-resident Krylov scalars through the C ABI: a dot product written into
a slot equals hg_function_dot bit for bit, a linear combination with slot
quotients equals hg_function_add / hg_function_assign with the host-formed
coefficients bit for bit, the slot allocator hands out disjoint ranges and
refuses what it does not own, and CG on the device scalars (CGSolver) takes
the iterates of the host-scalar loop (PCG parity against the oracle)."""
import ctypes as C

import numpy as np
import pytest

import hyteg_amd as hg
from hyteg_amd import problems
from meshes import relmax
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _alloc(ctx, n):
    first = C.c_int(-1)
    ctx.check(hg.lib().hg_slots_alloc(ctx.h, n, C.byref(first)), "alloc")
    return first.value


def _read(ctx, first, n):
    out = np.zeros(n)
    ctx.check(hg.lib().hg_slots_read(ctx.h, first, n, out.ctypes.data), "read")
    return out


@pytest.mark.parametrize("L", [0, 2])
def test_dot_and_lincomb_slots_match_host_scalars(ctx, L):
    st = hg.PrimitiveStorage.shell(ctx, ntan=3, nrad=2, rmin=0.5, rmax=1.0)
    n = st.p2_ndofs(L)
    rng = np.random.default_rng(5 + L)
    f = [hg.P2Function(f"f{i}", st, L, L) for i in range(4)]
    for g in f:
        g.upload(L, rng.uniform(-1, 1, n))
    s0 = _alloc(ctx, 3)
    print(f"slot dots L{L}", flush=True)
    for flag in (hg.ALL, hg.INNER):
        ref = np.zeros(1)
        ctx.check(hg.lib().hg_function_dot(ctx.h, f[0].h, f[1].h, L, flag, ref.ctypes.data), "dot")
        ctx.check(hg.lib().hg_function_dot_slot(ctx.h, f[0].h, f[1].h, L, flag, s0), "dot slot")
        assert _read(ctx, s0, 1)[0] == ref[0]
    ctx.check(hg.lib().hg_function_dot_slot(ctx.h, f[0].h, f[1].h, L, hg.INNER, s0), "dot slot")
    ctx.check(hg.lib().hg_function_dot_slot(ctx.h, f[2].h, f[2].h, L, hg.ALL, s0 + 1), "dot slot")
    a, b = _read(ctx, s0, 2)
    # add: f3 += -(a / b) f1 + 1 * f2, against the host-scalar add on a copy
    base = f[3].download(L)
    ref_f = hg.P2Function("r", st, L, L)
    ref_f.upload(L, base)
    srcs = np.array([f[1].h, f[2].h], dtype=np.int64)
    for flag, add in ((hg.INNER, 1), (hg.ALL, 0)):
        f[3].upload(L, base)
        ref_f.upload(L, base)
        signs = np.array([-1.0, 1.0])
        num = np.array([s0, -1], dtype=np.int32)
        den = np.array([s0 + 1, -1], dtype=np.int32)
        ctx.check(hg.lib().hg_function_lincomb_slots(ctx.h, f[3].h, 2, signs.ctypes.data, num.ctypes.data,
                                                     den.ctypes.data, srcs.ctypes.data, L, flag, add), "lincomb")
        coef = np.array([-(a / b), 1.0])
        fn = hg.lib().hg_function_add if add else hg.lib().hg_function_assign
        ctx.check(fn(ctx.h, ref_f.h, 2, coef.ctypes.data, srcs.ctypes.data, L, flag), "host lincomb")
        assert np.array_equal(f[3].download(L), ref_f.download(L)), (flag, add)
    ctx.check(hg.lib().hg_slots_free(ctx.h, s0, 3), "free")
    for g in f + [ref_f]:
        g.destroy()


def test_slot_allocator(ctx):
    a = _alloc(ctx, 4)
    b = _alloc(ctx, 4)
    assert set(range(a, a + 4)).isdisjoint(range(b, b + 4))
    ctx.check(hg.lib().hg_slots_free(ctx.h, a, 4), "free")
    c = _alloc(ctx, 2)
    assert a <= c <= a + 2  # the freed range is reused
    out = np.zeros(1)
    assert hg.lib().hg_slots_read(ctx.h, a + 3, 1, out.ctypes.data) != 0  # freed, not owned
    assert hg.lib().hg_slots_alloc(ctx.h, 257, C.byref(C.c_int())) != 0
    assert hg.lib().hg_slots_read(ctx.h, 300, 1, out.ctypes.data) != 0
    for first, n in ((b, 4), (c, 2)):
        ctx.check(hg.lib().hg_slots_free(ctx.h, first, n), "free")
    # all 256 slots in one range, then nothing is left
    d = _alloc(ctx, 256)
    assert hg.lib().hg_slots_alloc(ctx.h, 1, C.byref(C.c_int())) != 0
    ctx.check(hg.lib().hg_slots_free(ctx.h, d, 256), "free")


def test_cg_on_device_scalars_matches_the_oracle_loop(ctx):
    """CGSolver (device scalars, Jacobi preconditioner) against the oracle's PCG
    after a fixed number of iterations: P2 div(k grad) on the shell at level 1"""
    st, om = hg.PrimitiveStorage.shell(ctx, **problems.SHELL), po.Mesh.shell(**problems.SHELL)
    L = 1
    k = problems.shell_coefficient(0)
    A = hg.P2DivKGradOperator(st, 0, L, *k)
    op = po.P2Op(om, L, po.KCoeff(*k))
    fl = om.flags(L + 1)
    b = np.where(fl == hg.INNER, problems.canonical_uniform(len(fl), 4), 0.0)
    X, B = hg.P2Function("x", st, L, L), hg.P2Function("b", st, L, L)
    B.upload(L, b)
    for its in (1, 4, 15):
        X.upload(L, np.zeros(len(b)))
        print(f"device-scalar PCG {its} iterations", flush=True)
        it = hg.p2_cg_solve(A, X, B, L, max_iter=its, rel_tol=1e-30, abs_tol=1e-300)
        x = np.zeros(len(b))
        it_o = op.cg(x, b, max_iter=its, rel_tol=1e-30, abs_tol=1e-300)
        assert it == it_o == its - 1
        assert relmax(X.download(L), x) < 1e-11, its
    for g in (X, B):
        g.destroy()
    A.destroy()
