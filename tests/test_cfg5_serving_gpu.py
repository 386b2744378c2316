"""The serving stream against the oracle: a mixed stream of P1 Poisson (unit cube
levels 3-5, shell macro mesh levels 3-4) and P2 div(k grad) (cube levels 2-3,
blended shell level 2) requests, arriving while others run, is solved to 1e-10;
every request's downloaded solution is compared with the oracle's FP64 GMG
solution of the same system (right-hand side and coefficient generated from the
request's seed; solved to 1e-12), and for one request of every kind the full
solution vector computed alone equals the one computed in the mix (P1 bit for
bit: their vectors never interact; P2 up to the order of floating-point atomic
sums in the apply).  The shell's P1 level 5 (10.5 M DoFs) and P2 level 3 (1.3 M
DoFs) classes of the benchmark mix are left out: the oracle needs minutes there."""
import numpy as np
import pytest

from hyteg_amd import problems, serving
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _oracle_mesh(kind):
    if kind in (serving.P1_CUBE, serving.P2_CUBE):
        return po.Mesh.cuboid()
    if kind == serving.P2_SHELL:
        return po.Mesh.shell(**problems.SHELL)
    sh = po.Mesh.shell(**problems.SHELL)  # P1: the shell's macro mesh without blending
    return po.Mesh.from_arrays(sh.coords(0), sh.cell_vertices())


def _oracle_solution(kind, level, seed):
    m = _oracle_mesh(kind)
    if kind in (serving.P1_CUBE, serving.P1_SHELL):
        fl = m.flags(level)
        b = np.where(fl == po.INNER, problems.canonical_uniform(len(fl), int(seed)), 0.0)
        x = np.zeros_like(b)
        k, hist = m.gmg(0, level, x, b, 3, 3, cycles=80, tol=1e-12)
        assert hist[-1] <= 1e-12 * hist[0]
        return x
    kc = po.KCoeff(*serving.request_coefficient(seed))
    fl = m.flags(level + 1)
    b = np.where(fl == po.INNER, problems.canonical_uniform(len(fl), int(seed)), 0.0)
    op = po.P2Op(m, level, kc)
    rho = op.power_iteration(op.inv_diag(), problems.canonical_uniform(len(fl), 1), 15)
    x = np.zeros_like(b)
    k, hist = m.p2_gmg(0, level, kc, x, b, 2, 2, 3, rho, cycles=80, tol=1e-12)
    assert hist[-1] <= 1e-12 * hist[0]
    return x


def test_served_solutions_equal_the_oracle_and_do_not_depend_on_the_mix(ctx):
    P1C, P1S, P2S, P2C = serving.P1_CUBE, serving.P1_SHELL, serving.P2_SHELL, serving.P2_CUBE
    reqs = [(P1C, 3), (P1C, 4), (P2C, 2), (P1S, 3), (P1C, 5), (P2S, 2), (P1C, 3), (P2C, 3), (P1S, 4), (P1C, 4),
            (P2C, 3), (P2C, 2), (P1C, 5), (P1S, 3)]
    kinds = np.array([k for k, _ in reqs], dtype=np.int32)
    levels = np.array([lv for _, lv in reqs], dtype=np.int32)
    seeds = np.arange(500, 500 + len(reqs), dtype=np.uint64)
    arrivals = np.linspace(0.0, 0.4, len(reqs))
    print("mixed stream with solutions", flush=True)
    mix = serving.serve_stream(kinds, levels, seeds, arrivals, tol=TOL, slots=8, solutions=True)
    assert (mix["rel_residual"] <= TOL).all(), mix["rel_residual"]
    for i, (k, lv) in enumerate(reqs):
        print(f"oracle: request {i} kind {k} level {lv}", flush=True)
        xo = _oracle_solution(k, lv, seeds[i])
        x = mix["solutions"][i]
        err = np.abs(x - xo).max() / np.abs(xo).max()
        print(f"   relative max difference {err:.2e}", flush=True)
        assert err < 1e-8, (i, k, lv, err)
    for i in (0, 3, 5, 7):
        print(f"alone: request {i} kind {kinds[i]} level {levels[i]}", flush=True)
        one = serving.serve_stream(kinds[i:i + 1], levels[i:i + 1], seeds[i:i + 1], np.zeros(1), tol=TOL, slots=8,
                                   solutions=True)
        a, b = one["solutions"][0], mix["solutions"][i]
        assert one["cycles"][0] == mix["cycles"][i]
        if kinds[i] in (P1C, P1S):
            assert np.array_equal(a, b)
        else:
            assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()
