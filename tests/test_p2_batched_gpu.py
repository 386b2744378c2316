"""Batched P2 div(k grad) (interleaved vectors, one coefficient slot per vector)
through the C ABI: every vector of a batched apply equals the oracle's apply
with that slot's k (FP64, <= 1e-12 relative max-norm), including after a slot
is given new coefficients; the batched inverse diagonal, P2 transfers and
per-vector Chebyshev steps equal the single-vector routines vector by vector."""
import ctypes as C

import numpy as np
import pytest

import hyteg_amd as hg
from hyteg_amd import problems
from hyteg_amd._lib import lib
from hyteg_amd.api import P2DivKGradOperatorBatched
from meshes import relmax
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-12
NV = 4


@pytest.fixture(scope="module")
def shell_pair(ctx):
    return hg.PrimitiveStorage.shell(ctx, **problems.SHELL), po.Mesh.shell(**problems.SHELL)


def _k(seed):
    """coefficient modes of a slot; odd seeds give k = 1"""
    return problems.shell_coefficient(seed=seed) if seed % 2 == 0 else ((), (), ())


def _oracle_apply(om, L, k, x):
    ref = np.zeros_like(x)
    po.P2Op(om, L, po.KCoeff(*k) if len(k[0]) else None).apply(x.copy(), ref, hg.ALL, False)
    return ref


@pytest.mark.parametrize("L", [1, 2, 3])
def test_batched_apply_matches_the_oracle_per_slot(ctx, shell_pair, L):
    st, om = shell_pair
    A = P2DivKGradOperatorBatched(st, L, L, NV)
    seeds = [100 + s for s in range(NV)]
    for s in range(NV):
        A.set_slot(s, *_k(seeds[s]))
    n = st.p2_ndofs(L)
    x = np.random.default_rng(L).uniform(-1, 1, (n, NV))
    src = hg.P2Function("s", st, L, L, nvec=NV)
    dst = hg.P2Function("d", st, L, L, nvec=NV)
    src.upload(L, x)
    for rnd in range(2):
        print(f"batched P2 apply L{L} round {rnd}", flush=True)
        A.apply(src, dst, L, hg.ALL)
        y = dst.download(L)
        for s in range(NV):
            err = relmax(y[:, s], _oracle_apply(om, L, _k(seeds[s]), x[:, s]))
            assert err < TOL, (L, rnd, s, err)
        # new coefficients for slot 1 after its tables were built
        seeds[1] = 200 + 2 * L
        A.set_slot(1, *_k(seeds[1]))


def test_batched_inverse_diagonal_transfers_and_chebyshev_step(ctx, shell_pair):
    st, _ = shell_pair
    L = 3
    A = P2DivKGradOperatorBatched(st, L - 1, L, NV)
    singles = []
    for s in range(NV):
        A.set_slot(s, *_k(300 + s))
        singles.append(hg.P2DivKGradOperator(st, L - 1, L, *_k(300 + s)))
    print("batched inverse diagonal", flush=True)
    db = hg.P2Function("db", st, L - 1, L, nvec=NV)
    ds = hg.P2Function("ds", st, L - 1, L)
    for lv in (L - 1, L):
        A.inverse_diagonal(lv, db)
        got = db.download(lv)
        for s in range(NV):
            singles[s].inverse_diagonal(lv, ds)
            assert relmax(got[:, s], ds.download(lv)) < 1e-14, (lv, s)
    print("batched P2 restriction and prolongation", flush=True)
    rng = np.random.default_rng(7)
    fb = hg.P2Function("fb", st, L - 1, L, nvec=NV)
    fs = hg.P2Function("fs", st, L - 1, L)
    xf = rng.uniform(-1, 1, (st.p2_ndofs(L), NV))
    xc = rng.uniform(-1, 1, (st.p2_ndofs(L - 1), NV))
    fb.upload(L, xf)
    fb.upload(L - 1, xc)
    hg.restrict(fb, L, hg.ALL)
    rb = fb.download(L - 1)
    hg.prolongate(fb, L - 1, hg.ALL, hg.ADD)
    pb = fb.download(L)
    for s in range(NV):
        fs.upload(L, xf[:, s])
        fs.upload(L - 1, xc[:, s])
        hg.restrict(fs, L, hg.ALL)
        assert np.array_equal(rb[:, s], fs.download(L - 1)), s
        hg.prolongate(fs, L - 1, hg.ALL, hg.ADD)
        assert np.array_equal(pb[:, s], fs.download(L)), s
    print("per-vector Chebyshev step", flush=True)
    xs = [hg.P2Function(nm, st, L, L, nvec=NV) for nm in ("x", "t", "s", "b")]
    vals = [rng.uniform(-1, 1, (st.p2_ndofs(L), NV)) for _ in xs]
    for f, v in zip(xs, vals):
        f.upload(L, v)
    A.inverse_diagonal(L, db)
    cs = np.array([0.7, -1.3, 2.1, 0.25])
    first = C.c_int()
    ctx.check(lib().hg_slots_alloc(ctx.h, NV, C.byref(first)), "slots")
    ctx.check(lib().hg_slots_write(ctx.h, first.value, NV, cs.ctypes.data_as(C.c_void_p)), "slots")
    x, t, s_, b = xs
    ctx.check(lib().hg_function_chebyshev_step_slots(ctx.h, x.h, t.h, s_.h, b.h, db.h, first.value, L, hg.INNER),
              "chebyshev_step_slots")
    xb, tb = x.download(L), t.download(L)
    ys = [hg.P2Function(nm, st, L, L) for nm in ("x1", "t1", "s1", "b1")]
    dg = db.download(L)
    d1 = hg.P2Function("d1", st, L, L)
    for k in range(NV):
        for f, v in zip(ys, vals):
            f.upload(L, v[:, k])
        d1.upload(L, dg[:, k])
        x1, t1, s1, b1 = ys
        ctx.check(lib().hg_function_chebyshev_step(ctx.h, x1.h, t1.h, s1.h, b1.h, d1.h, cs[k], L, hg.INNER),
                  "chebyshev_step")
        assert np.array_equal(xb[:, k], x1.download(L)), k
        assert np.array_equal(tb[:, k], t1.download(L)), k
    ctx.check(lib().hg_slots_free(ctx.h, first.value, NV), "slots")


def test_inactive_slots_give_zero_and_leave_the_others_unchanged(ctx, shell_pair):
    st, om = shell_pair
    L = 3
    A = P2DivKGradOperatorBatched(st, L, L, NV)
    for s in range(NV):
        A.set_slot(s, *_k(400 + s))
    x = np.random.default_rng(11).uniform(-1, 1, (st.p2_ndofs(L), NV))
    src = hg.P2Function("s", st, L, L, nvec=NV)
    dst = hg.P2Function("d", st, L, L, nvec=NV)
    src.upload(L, x)
    A.apply(src, dst, L, hg.ALL)
    full = dst.download(L)
    ctx.check(lib().hg_p2_divkgrad_slot_active(ctx.h, A.h, 2, 0), "slot_active")
    A.apply(src, dst, L, hg.ALL)
    part = dst.download(L)
    assert not part[:, 2].any()
    for s in (0, 1, 3):
        assert relmax(part[:, s], full[:, s]) < 1e-14, s
    ctx.check(lib().hg_p2_divkgrad_slot_active(ctx.h, A.h, 2, 1), "slot_active")
    A.apply(src, dst, L, hg.ALL)
    assert relmax(dst.download(L)[:, 2], _oracle_apply(om, L, _k(402), x[:, 2])) < TOL
