"""Interleaved multi-vector (batched) P1 path: every kernel applied to a function
with nvec interleaved vectors gives, vector by vector, what it gives for nvec
single-vector functions (the single-vector path is pinned to the oracle by the
other GPU tests); batched GMG and FP16 iterative refinement reach the FP64
single-vector solutions."""
import numpy as np
import pytest

import hyteg_amd as hg
from hyteg_amd import problems
from meshes import octahedron, relmax

pytestmark = pytest.mark.gpu

NV = 8


@pytest.fixture(scope="module", params=["cube", "octa"])
def storage(ctx, request):
    if request.param == "cube":
        return hg.PrimitiveStorage.cuboid(ctx)
    return hg.PrimitiveStorage.from_cells(ctx, *octahedron())


def _singles(st, L0, L, X):
    fs = []
    for k in range(X.shape[1]):
        f = hg.P1Function(f"s{k}", st, L0, L)
        f.upload(L, X[:, k])
        fs.append(f)
    return fs


def test_batched_blas_random_and_dot(ctx, storage):
    L = 3
    n = storage.ndofs(L)
    f = hg.P1Function("f", storage, L, L, nvec=NV)
    g = hg.P1Function("g", storage, L, L, nvec=NV)
    f.interpolate_random(40, -1.0, 1.0, L)
    X = f.download(L)
    for k in range(NV):
        assert np.array_equal(X[:, k], problems.canonical_uniform(n, 40 + k))
    Y = np.random.default_rng(1).standard_normal((n, NV))
    g.upload(L, Y)
    assert np.array_equal(g.download(L), Y)
    fl = storage.flags(L)
    m = fl == hg.INNER
    d = f.dot(g, L, hg.INNER)
    assert np.abs(d - (X[m] * Y[m]).sum(axis=0)).max() < 1e-12 * np.abs(d).max()
    f.assign([2.0, -1.0], [f, g], L, hg.INNER)
    ref = np.where(m[:, None], 2 * X - Y, X)
    assert relmax(f.download(L), ref) < 1e-15
    f.scale(np.arange(1.0, NV + 1.0), L, hg.INNER)
    ref = np.where(m[:, None], ref * np.arange(1.0, NV + 1.0)[None, :], ref)
    assert relmax(f.download(L), ref) < 1e-15


def test_batched_kernels_equal_single_vector_kernels(ctx, storage):
    L = 3
    A = hg.P1ConstantLaplaceOperator(storage, 0, L)
    n, nc = storage.ndofs(L), storage.ndofs(L - 1)
    rng = np.random.default_rng(2)
    X = rng.uniform(-1, 1, (n, NV))
    B = rng.uniform(-1, 1, (n, NV))
    C = rng.uniform(-1, 1, (nc, NV))
    fb = hg.P1Function("x", storage, 0, L, nvec=NV)
    yb = hg.P1Function("y", storage, 0, L, nvec=NV)
    bb = hg.P1Function("b", storage, 0, L, nvec=NV)
    singles = _singles(storage, 0, L, X)
    ys = [hg.P1Function(f"y{k}", storage, 0, L) for k in range(NV)]
    bs = _singles(storage, 0, L, B)
    fb.upload(L, X)
    bb.upload(L, B)
    print("batched apply", flush=True)
    A.apply(fb, yb, L, hg.ALL)
    for k in range(NV):
        A.apply(singles[k], ys[k], L, hg.ALL)
    Yb = yb.download(L)
    for k in range(NV):
        assert np.array_equal(Yb[:, k], ys[k].download(L))
    print("batched sor", flush=True)
    A.smooth_sor(fb, bb, 1.3, L, hg.GMG_FLAG)
    for k in range(NV):
        A.smooth_sor(singles[k], bs[k], 1.3, L, hg.GMG_FLAG)
    Xb = fb.download(L)
    for k in range(NV):
        assert np.array_equal(Xb[:, k], singles[k].download(L))
    print("batched transfer", flush=True)
    fb.upload(L - 1, C)
    for k in range(NV):
        singles[k].upload(L - 1, C[:, k])
    hg.prolongate(fb, L - 1, hg.GMG_FLAG, hg.ADD)
    hg.restrict(fb, L, hg.GMG_FLAG)
    for k in range(NV):
        hg.prolongate(singles[k], L - 1, hg.GMG_FLAG, hg.ADD)
        hg.restrict(singles[k], L, hg.GMG_FLAG)
    for lev in (L - 1, L):
        Zb = fb.download(lev)
        for k in range(NV):
            assert np.array_equal(Zb[:, k], singles[k].download(lev))
    print("fp16 batched apply", flush=True)
    fh = hg.P1Function("xh", storage, L, L, dtype=hg.F16, nvec=NV)
    yh = hg.P1Function("yh", storage, L, L, dtype=hg.F16, nvec=NV)
    fh.upload(L, X)
    A.apply(fh, yh, L, hg.ALL)
    assert relmax(yh.download(L), Yb) < 4e-3


def test_batched_gmg_and_fp16_refinement_reach_single_solutions(ctx):
    st = hg.PrimitiveStorage.cuboid(ctx)
    L = 4
    A = hg.P1ConstantLaplaceOperator(st, 0, L)
    fl = st.flags(L)
    n = len(fl)
    B = np.stack([np.where(fl == hg.INNER, problems.canonical_uniform(n, k), 0.0) for k in range(NV)], axis=1)
    ref = []
    for k in range(NV):
        x = hg.P1Function("x", st, 0, L)
        b = hg.P1Function("b", st, 0, L)
        b.upload(L, B[:, k])
        hg.gmg_solve(A, x, b, 0, L, 3, 3, cycles=40, tol=1e-12)
        ref.append(x.download(L))
    ref = np.stack(ref, axis=1)
    xb = hg.P1Function("xb", st, 0, L, nvec=NV)
    bb = hg.P1Function("bb", st, 0, L, nvec=NV)
    bb.upload(L, B)
    print("batched fp64 gmg", flush=True)
    k64, h64 = hg.batched_gmg_solve(A, xb, bb, 0, L, 3, 3, cycles=40, tol=1e-12)
    assert (h64[-1] <= 1e-12 * h64[0]).all()
    assert relmax(xb.download(L), ref) < 1e-10
    print("batched fp16 refinement", flush=True)
    xb.interpolate(0.0, L)
    kir, hir = hg.batched_ir_solve(A, xb, bb, 0, L, low_dtype=hg.F16, max_outer=40, tol=1e-10)
    print(f"batched: fp64 {k64} cycles, fp16 refinement {kir} steps", flush=True)
    assert (hir[-1] <= 1e-10 * hir[0]).all()
    assert relmax(xb.download(L), ref) < 1e-8


def test_graph_captured_cycles_equal_eager_cycles(ctx):
    """the V-cycle replayed from a captured HIP graph gives the eager results bit for bit"""
    st = hg.PrimitiveStorage.cuboid(ctx)
    L = 4
    A = hg.P1ConstantLaplaceOperator(st, 0, L)
    out = []
    for graphs in (True, False):
        ctx.set_graphs(graphs)
        x = hg.P1Function("x", st, 0, L, nvec=4)
        b = hg.P1Function("b", st, 0, L, nvec=4)
        b.interpolate_random(3, -1.0, 1.0, L, hg.INNER)
        k1, h1 = hg.batched_gmg_solve(A, x, b, 0, L, cycles=6)
        x2 = hg.P1Function("x2", st, 0, L, nvec=4)
        k2, h2 = hg.batched_ir_solve(A, x2, b, 0, L, low_dtype=hg.F16, max_outer=5, tol=0.0)
        out.append((h1, x.download(L), h2, x2.download(L)))
    ctx.set_graphs(True)
    for a, b_ in zip(out[0], out[1]):
        assert np.array_equal(a, b_)
