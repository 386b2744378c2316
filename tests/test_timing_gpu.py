"""The timing tree: while enabled, the multigrid solver records the reference's
named regions (GeometricMultigridSolver.hpp:200-306) as HIP-event pairs on the
context stream; the report nests them by path and counts every visit."""
import numpy as np
import pytest

import hyteg_amd as hg

pytestmark = pytest.mark.gpu


def test_gmg_records_the_reference_regions(ctx):
    st = hg.PrimitiveStorage.cuboid(ctx)
    L, cycles = 4, 3
    A = hg.P1ConstantLaplaceOperator(st, 0, L)
    x, b = hg.P1Function("x", st, 0, L), hg.P1Function("b", st, 0, L)
    b.interpolate_random(1, -1.0, 1.0, L, hg.INNER)
    ctx.set_graphs(False)
    ctx.timing(True)
    try:
        hg.gmg_solve(A, x, b, 0, L, 3, 3, cycles=cycles)
        rep = ctx.timing_report(reset=True)
    finally:
        ctx.timing(False)
        ctx.set_graphs(True)
    print(rep, flush=True)
    top = "Geometric Multigrid Solver"
    assert rep[top][0] == cycles
    for lv in range(1, L + 1):
        path = f"{top}/Level {lv}"
        assert rep[path][0] == 2 * cycles  # entered before and after the coarser levels
        for sub in ("Smoother", "Restriction", "Prolongation"):
            assert rep[f"{path}/{sub}"][0] == (2 if sub == "Smoother" else 1) * cycles
    assert rep[f"{top}/Level 0/Coarse Grid Solver"][0] == cycles
    assert all(ms >= 0.0 for _, ms in rep.values())
    inner = sum(ms for p, (_, ms) in rep.items() if p.count("/") == 1)
    assert inner <= rep[top][1] * (1 + 1e-6) + 1e-3
    # nothing is recorded while disabled; reset emptied the report
    hg.gmg_solve(A, x, b, 0, L, 3, 3, cycles=1)
    assert ctx.timing_report() == {}
