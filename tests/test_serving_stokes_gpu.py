"""Stokes requests in the serving stream: isoviscous P2-P1 Taylor-Hood solves on
the blended shell (no-slip surfaces, buoyancy of 12 blobs whose centres come from
the request's seed) mixed with P1 Poisson requests.
Every Stokes solution is checked with the oracle's Stokes operator: b - A x of
the downloaded solution, with b rebuilt from the request's blob centres, has the
relative residual the stream reports; one request solved alone gives the same
vector as in the mix (up to the order of floating-point atomic sums)."""
import numpy as np
import pytest

from hyteg_amd import problems, serving
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-6
C0, C1 = 0.0, 0.0  # isoviscous


def _rhs(mesh, level, seed, sigma=0.05):
    """the velocity right-hand side [ux|uy|uz] the stream builds for a Stokes request"""
    xyz = mesh.coords(level + 1)
    fl = mesh.flags(level + 1)
    c = serving.request_blobs(seed)
    T = np.zeros(len(xyz))
    for cb in c:
        d2 = ((xyz - cb) ** 2).sum(axis=1)
        near = d2 < 36.0 * sigma * sigma
        T[near] += np.exp(-d2[near] / (2.0 * sigma * sigma))
    T[fl != po.INNER] = 0.0
    r = np.linalg.norm(xyz, axis=1)
    return np.concatenate([T * xyz[:, k] / r for k in range(3)])


def test_stokes_requests_in_the_stream(ctx):
    S, P1C = serving.STOKES_SHELL, serving.P1_CUBE
    reqs = [(S, 2), (P1C, 4), (S, 3), (P1C, 3), (S, 2), (P1C, 5)]
    kinds = np.array([k for k, _ in reqs], dtype=np.int32)
    levels = np.array([lv for _, lv in reqs], dtype=np.int32)
    seeds = np.arange(900, 900 + len(reqs), dtype=np.uint64)
    arrivals = np.linspace(0.0, 0.2, len(reqs))
    print("mixed stream with Stokes requests", flush=True)
    mix = serving.serve_stream(kinds, levels, seeds, arrivals, tol=TOL, slots=8, solutions=True)
    print("cycles", mix["cycles"], "relres", mix["rel_residual"], flush=True)
    assert (mix["rel_residual"] <= TOL).all(), mix["rel_residual"]
    mesh = po.Mesh.shell(**problems.SHELL, inner_flag=po.DIRICHLET, outer_flag=po.DIRICHLET)
    for i, (k, lv) in enumerate(reqs):
        if k != S:
            continue
        print(f"oracle residual: request {i} level {lv}", flush=True)
        op = po.StokesOp(mesh, lv, C0, C1)
        bu = _rhs(mesh, lv, seeds[i])
        bp = np.zeros(op.n1)
        x = mix["solutions"][i]
        u, p = x[:3 * op.n2], x[3 * op.n2:]
        assert len(p) == op.n1
        ru, rp = po.stokes_residual(op, u, p, bu, bp, po.GMG_FLAG)
        r0u, r0p = po.stokes_residual(op, np.zeros_like(u), np.zeros_like(p), bu, bp, po.GMG_FLAG)
        rel = np.sqrt((ru ** 2).sum() + (rp ** 2).sum()) / np.sqrt((r0u ** 2).sum() + (r0p ** 2).sum())
        print(f"   oracle relative residual {rel:.3e}, stream {mix['rel_residual'][i]:.3e}", flush=True)
        assert rel <= 1.05 * TOL
        assert abs(rel - mix["rel_residual"][i]) <= 0.05 * mix["rel_residual"][i] + 1e-14
        assert np.abs(p.mean()) <= 1e-10 * max(1.0, np.abs(p).max())
    i = 4
    print(f"alone: request {i}", flush=True)
    one = serving.serve_stream(kinds[i:i + 1], levels[i:i + 1], seeds[i:i + 1], np.zeros(1), tol=TOL, slots=8,
                               solutions=True)
    a, b = one["solutions"][0], mix["solutions"][i]
    assert one["cycles"][0] == mix["cycles"][i]
    assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max()
