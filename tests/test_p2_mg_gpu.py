"""Parity of the HIP P2 multigrid components with the CPU oracle, through the C ABI:
quadratic prolongation / restriction (FP64, <= 1e-12 relative max-norm),
Chebyshev smoothing steps, power iteration, preconditioned CG, and the
residual reduction factors of the V-cycle (within 5 % per cycle)."""
import numpy as np
import pytest

import hyteg_amd as hg
from hyteg_amd import problems
from meshes import relmax
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def shell_pair(ctx):
    return hg.PrimitiveStorage.shell(ctx, **problems.SHELL), po.Mesh.shell(**problems.SHELL)


@pytest.fixture(scope="module")
def cube_pair(ctx):
    return hg.PrimitiveStorage.cuboid(ctx), po.Mesh.cuboid()


@pytest.mark.parametrize("mesh", ["cube", "shell"])
def test_p2_transfer_parity(ctx, cube_pair, shell_pair, mesh):
    st, om = cube_pair if mesh == "cube" else shell_pair
    rng = np.random.default_rng(5)
    top = 3 if mesh == "cube" else 2
    f = hg.P2Function("f", st, 0, top)
    for Lc in range(top):
        print(f"P2 transfer parity {mesh} {Lc}->{Lc + 1}", flush=True)
        nc, nf = om.ndofs(Lc + 1), om.ndofs(Lc + 2)
        c, fine0 = rng.uniform(-1, 1, nc), rng.uniform(-1, 1, nf)
        for flag, update in ((hg.ALL, hg.REPLACE), (hg.GMG_FLAG, hg.ADD), (hg.GMG_FLAG, hg.REPLACE)):
            f.upload(Lc, c)
            f.upload(Lc + 1, fine0)
            hg.prolongate(f, Lc, flag, update)
            ref = fine0.copy()
            om.p2_prolongate(Lc, c, ref, flag, add=update == hg.ADD)
            assert relmax(f.download(Lc + 1), ref) < TOL, (Lc, flag, update)
        c0 = rng.uniform(-1, 1, nc)
        for flag in (hg.ALL, hg.GMG_FLAG):
            f.upload(Lc + 1, fine0)
            f.upload(Lc, c0)
            hg.restrict(f, Lc + 1, flag)
            ref = c0.copy()
            om.p2_restrict(Lc + 1, fine0, ref, flag)
            assert relmax(f.download(Lc), ref) < TOL, (Lc, flag)


@pytest.mark.parametrize("dtype", ["f64", "f16"])
def test_p2_prolongate_parity_large(ctx, cube_pair, dtype):
    """prolongation into cube P2 levels 5 and 6 (fine lattices of 64 and 128
    micro-edges: chunks of several rows per workgroup) against the oracle; FP16
    values against the oracle applied to the same FP16-rounded coarse values"""
    st, om = cube_pair
    rng = np.random.default_rng(9)
    dt = hg.F64 if dtype == "f64" else hg.F16
    f = hg.P2Function("f", st, 4, 6, dtype=dt)
    for Lc in (4, 5):
        print(f"P2 prolongation parity cube {Lc}->{Lc + 1} {dtype}", flush=True)
        nc, nf = om.ndofs(Lc + 1), om.ndofs(Lc + 2)
        c, fine0 = rng.uniform(-1, 1, nc), rng.uniform(-1, 1, nf)
        if dt == hg.F16:
            c, fine0 = c.astype(np.float16).astype(np.float64), fine0.astype(np.float16).astype(np.float64)
        for flag, update in ((hg.ALL, hg.REPLACE), (hg.GMG_FLAG, hg.ADD)):
            f.upload(Lc, c)
            f.upload(Lc + 1, fine0)
            hg.prolongate(f, Lc, flag, update)
            ref = fine0.copy()
            om.p2_prolongate(Lc, c, ref, flag, add=update == hg.ADD)
            tol = TOL if dt == hg.F64 else 2e-3  # one FP16 rounding of values below 4 in magnitude
            assert relmax(f.download(Lc + 1), ref) < tol, (Lc, flag, update, relmax(f.download(Lc + 1), ref))
    f.destroy()


def _problem(st, om, L, seed=0):
    k = problems.shell_coefficient(0)
    A = hg.P2DivKGradOperator(st, 0, L, *k)
    fl = om.flags(L + 1)
    b = np.where(fl == hg.INNER, problems.canonical_uniform(len(fl), seed), 0.0)
    return A, po.KCoeff(*k), fl, b


def test_p2_power_iteration_chebyshev_and_cg_parity(ctx, shell_pair):
    st, om = shell_pair
    L = 2
    A, k, fl, b = _problem(st, om, L)
    op = po.P2Op(om, L, k)
    inv = op.inv_diag()
    print("power iteration", flush=True)
    x0 = problems.canonical_uniform(len(fl), 1)
    fx = hg.P2Function("x", st, L, L)
    fx.upload(L, x0)
    rho = hg.p2_estimate_radius(A, fx, L, 15)
    xo = x0.copy()
    rho_o = op.power_iteration(inv, xo, 15)
    assert abs(rho - rho_o) < 1e-10 * rho_o
    print("chebyshev", flush=True)
    x = np.where(fl == hg.INNER, problems.canonical_uniform(len(fl), 2), 0.0)
    fb = hg.P2Function("b", st, L, L)
    fb.upload(L, b)
    for order in (1, 3):
        fx.upload(L, x)
        hg.p2_chebyshev_smooth(A, fx, fb, L, order, rho_o, steps=2)
        ref = x.copy()
        coeffs = po.chebyshev_coefficients(order, 0.3 * rho_o, 1.2 * rho_o)
        for _ in range(2):
            op.chebyshev_smooth(inv, coeffs, ref, b)
        assert relmax(fx.download(L), ref) < 1e-11, order
    print("pcg", flush=True)
    L = 1
    A1, _, fl1, b1 = _problem(st, om, L, seed=3)
    op1 = po.P2Op(om, L, k)
    fx1, fb1 = hg.P2Function("x", st, L, L), hg.P2Function("b", st, L, L)
    fb1.upload(L, b1)
    it = hg.p2_cg_solve(A1, fx1, fb1, L, rel_tol=1e-12)
    xo = np.zeros(len(fl1))
    it_o = op1.cg(xo, b1, rel_tol=1e-12)
    assert abs(it - it_o) <= 2, (it, it_o)
    assert relmax(fx1.download(L), xo) < 1e-9


def test_p2_gmg_factors_within_5_percent(ctx, shell_pair):
    st, om = shell_pair
    L = 2
    A, k, fl, b = _problem(st, om, L)
    op = po.P2Op(om, L, k)
    rho = op.power_iteration(op.inv_diag(), problems.canonical_uniform(len(fl), 1), 20)
    x = hg.P2Function("x", st, 0, L)
    fb = hg.P2Function("b", st, 0, L)
    fb.upload(L, b)
    print("gpu gmg", flush=True)
    kg, hg_hist = hg.p2_gmg_solve(A, x, fb, 0, L, rho, pre=2, post=2, order=3, cycles=15, tol=1e-8)
    print("oracle gmg", flush=True)
    xo = np.zeros(len(fl))
    ko, ho = om.p2_gmg(0, L, k, xo, b, 2, 2, 3, rho, cycles=15, tol=1e-8)
    fg, fo = hg_hist[1:] / hg_hist[:-1], ho[1:] / ho[:-1]
    n = min(len(fg), len(fo), 8)
    assert hg_hist[-1] <= 1e-8 * hg_hist[0]
    assert abs(kg - ko) <= 1
    assert np.abs(fg[:n] / fo[:n] - 1).max() < 0.05, (fg, fo)
    assert relmax(x.download(L), xo) < 1e-6
