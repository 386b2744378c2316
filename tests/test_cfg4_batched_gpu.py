"""The batched multi-RHS configuration at full size against the oracle: 64 P1
Laplacian systems on the unit cube (6 macro-tetrahedra) at level 6, right-hand
sides uniform[-1, 1] on the Inner DoFs from seeds 0..63, interleaved per DoF in
one hierarchy, solved by FP16-storage / FP32-arithmetic V(3,3) cycles inside FP64
iterative refinement to 1e-10; every vector is compared with the oracle's FP64
GMG solution of the same system (lexicographic Gauss-Seidel, solved to 1e-12)."""
import numpy as np
import pytest

import hyteg_amd as hg
from hyteg_amd import problems
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

NV, L = 64, 6


def test_64_interleaved_fp16_refinement_solves_match_the_oracle(ctx):
    st = hg.PrimitiveStorage.cuboid(ctx)
    om = po.Mesh.cuboid()
    A = hg.P1ConstantLaplaceOperator(st, 0, L)
    fl = st.flags(L)
    n = len(fl)
    assert n == 274625
    x = hg.P1Function("x", st, 0, L, nvec=NV)
    b = hg.P1Function("b", st, 0, L, nvec=NV)
    b.interpolate_random(0, -1.0, 1.0, L, hg.INNER)
    B = b.download(L)
    inner = fl == hg.INNER
    for k in range(NV):
        assert np.array_equal(B[:, k], np.where(inner, problems.canonical_uniform(n, k), 0.0))
    print("64 x cube level 6: FP16 refinement", flush=True)
    k_ir, hist = hg.batched_ir_solve(A, x, b, 0, L, low_dtype=hg.F16, max_outer=40, tol=1e-10)
    assert (hist[-1] <= 1e-10 * hist[0]).all(), hist[-1] / hist[0]
    X = x.download(L)
    print(f"refinement steps {k_ir}; oracle solves", flush=True)
    Xo, done = om.gmg_many(0, L, np.zeros((NV, n)), np.ascontiguousarray(B.T), cycles=60, tol=1e-12)
    assert (done < 60).all()
    err = np.abs(X - Xo.T).max(axis=0) / np.abs(Xo.T).max(axis=0)
    print(f"max relative difference per vector: worst {err.max():.2e}, median {np.median(err):.2e}", flush=True)
    assert err.max() <= 1e-8
    # Dirichlet DoFs untouched (zero)
    assert np.all(X[~inner & (fl == hg.DIRICHLET)] == 0.0)
