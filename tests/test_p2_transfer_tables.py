"""CPU checks of the product's translation-invariant P2 transfer stencils
(host-built tables, read through hg_p2_transfer_stencil; no GPU needed).

The restriction weight sums are the values P2QuadraticRestriction3DTest.cpp:46-103
expects for the restriction of 1 at cell-interior DoFs."""
import itertools

import numpy as np
import pytest

import hyteg_amd as hg


def _q(X):
    x, y, z = X[..., 0], X[..., 1], X[..., 2]
    return 0.3 + x - 2 * y + 0.5 * z + x * x - 1.5 * x * y + 2 * y * z - z * z + 0.25 * x * z + y * y


def test_pull_stencils_interpolate_quadratics():
    for rx, ry, rz in itertools.product(range(4), repeat=3):
        off, w = hg.p2_transfer_stencil(0, rx + 4 * ry + 16 * rz)
        assert abs(w.sum() - 1.0) < 1e-15
        # coarse nodes at (2*A + off) / 2 and the fine point at (4*A + r) / 4 in lattice-L units, A = 0
        val = (w * _q(off / 2.0)).sum()
        assert abs(val - _q(np.array([rx, ry, rz]) / 4.0)) < 1e-14
        if rx % 2 == 0 and ry % 2 == 0 and rz % 2 == 0:
            assert len(w) == 1 and w[0] == 1.0 and tuple(off[0]) == (rx // 2, ry // 2, rz // 2)


@pytest.mark.parametrize("cls,expected", [(0, 1 + 14 * 3 / 8 - 14 / 8 - 3 * 24 / 8 - 24 / 8),
                                          (1, 2.5 + 1.5 * 6), (2, 2.5 + 1.5 * 4), (3, 2.5 + 1.5 * 6),
                                          (4, 2.5 + 1.5 * 6), (5, 2.5 + 1.5 * 4), (6, 2.5 + 1.5 * 6),
                                          (7, 2.5 + 1.5 * 4)])
def test_restriction_weights_match_reference_test(cls, expected):
    off, w = hg.p2_transfer_stencil(1, cls)
    assert abs(w.sum() - expected) < 1e-13


def test_restriction_is_transpose_of_pull():
    """every (fine, coarse, weight) triple of the pull tables on a patch appears in the
    restriction tables and vice versa"""
    pull = {}
    for cls in range(64):
        r = (cls & 3, (cls >> 2) & 3, cls >> 4)
        off, w = hg.p2_transfer_stencil(0, cls)
        for A in itertools.product(range(2), repeat=3):
            F = tuple(4 * a + ri for a, ri in zip(A, r))
            for o, wt in zip(off, w):
                C = tuple(2 * a + oi for a, oi in zip(A, o))
                pull[(F, C)] = wt
    rest = {}
    for q in range(8):
        off, w = hg.p2_transfer_stencil(1, q)
        base = (q & 1, (q >> 1) & 1, (q >> 2) & 1)
        for S in itertools.product(range(1, 3), repeat=3):
            C = tuple(2 * s + b for s, b in zip(S, base))
            for o, wt in zip(off, w):
                F = tuple(2 * c + oi for c, oi in zip(C, o))
                rest[(F, C)] = wt
    # compare on fine points fully inside the pull patch [0, 8) and coarse nodes covered by both
    common = [k for k in rest if all(0 <= f < 8 for f in k[0])]
    assert len(common) > 100
    for k in common:
        assert k in pull and abs(pull[k] - rest[k]) < 1e-15, k
    for k, wt in pull.items():
        C = k[1]
        if all(2 <= c < 6 for c in C) and all(0 <= f < 8 for f in k[0]):
            assert k in rest and abs(rest[k] - wt) < 1e-15, k


_CONCURRENT_FIRST_USE = r"""
import sys, threading
sys.path.insert(0, sys.argv[1])
import hyteg_amd as hg
barrier = threading.Barrier(16)
out = {}
def work(i):
    barrier.wait()
    out[i] = [len(hg.p2_transfer_stencil(0, c)[1]) for c in range(64)] + \
             [len(hg.p2_transfer_stencil(1, c)[1]) for c in range(8)]
ts = [threading.Thread(target=work, args=(i,)) for i in range(16)]
for t in ts: t.start()
for t in ts: t.join()
print(sorted(set(tuple(v) for v in out.values())))
"""


def test_tables_built_once_under_concurrent_first_use():
    """the tables are built lazily; threads that ask for them at the same time
    (one context per serving worker) must all see the single-threaded tables"""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    ref = [len(hg.p2_transfer_stencil(0, c)[1]) for c in range(64)] + \
          [len(hg.p2_transfer_stencil(1, c)[1]) for c in range(8)]
    for _ in range(3):
        r = subprocess.run([sys.executable, "-c", _CONCURRENT_FIRST_USE, root], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        assert r.stdout.strip() == str([tuple(ref)]), r.stdout
