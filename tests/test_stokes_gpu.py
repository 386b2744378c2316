"""Parity of the HIP P2-P1 Stokes operator (blended full-viscous velocity block
with eta = exp(c0 + c1 r), divergence and its transpose, velocity inverse
diagonal, PSPG and lumped-mass pressure diagonals) against the CPU oracle,
through the C ABI.  FP64: <= 1e-12 relative max-norm."""
import numpy as np
import pytest

import hyteg_amd as hg
from hyteg_amd import problems
from meshes import relmax
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-12
VISC = (np.log(1e4) / 0.45, -np.log(1e4) / 0.45)


@pytest.fixture(scope="module")
def shell_pair(ctx):
    return hg.PrimitiveStorage.shell(ctx, **problems.SHELL), po.Mesh.shell(**problems.SHELL)


@pytest.fixture(scope="module")
def cube_pair(ctx):
    return hg.PrimitiveStorage.cuboid(ctx), po.Mesh.cuboid()


def _apply_parity(st, om, L, visc):
    A = hg.P2P1StokesOperator(st, L, L, *visc)
    op = po.StokesOp(om, L, *visc)
    rng = np.random.default_rng(20 + L)
    u, p = rng.uniform(-1, 1, 3 * op.n2), rng.uniform(-1, 1, op.n1)
    yu0, yp0 = rng.uniform(-1, 1, 3 * op.n2), rng.uniform(-1, 1, op.n1)
    src = hg.P2P1TaylorHoodFunction("src", st, L, L)
    dst = hg.P2P1TaylorHoodFunction("dst", st, L, L)
    src.upload(L, u, p)
    for mask in (1, 2, 3, 4, 7):
        for flag in (hg.GMG_FLAG, hg.ALL, hg.DIRICHLET):
            for update in (hg.REPLACE, hg.ADD):
                print(f"Stokes apply L{L} mask {mask} flag {flag} update {update}", flush=True)
                dst.upload(L, yu0, yp0)
                A.apply(src, dst, L, flag, update, mask)
                ru, rp = yu0.copy(), yp0.copy()
                op.apply(u, p, ru, rp, flag, mask, update == hg.ADD)
                gu, gp = dst.download(L)
                assert relmax(gu, ru) < TOL, (L, mask, flag, update, relmax(gu, ru))
                assert relmax(gp, rp) < TOL, (L, mask, flag, update, relmax(gp, rp))
    inv = hg.P2P1TaylorHoodFunction("inv", st, L, L)
    A.velocity_inverse_diagonal(L, inv)
    assert relmax(inv.download(L)[0], op.inv_diag()) < TOL
    A.pspg_inverse_diagonal(L, inv.p)
    assert relmax(inv.p.download(L), po.p1_pspg_inv_diag(om, L)) < TOL
    A.lumped_inverse_mass(L, inv.p)
    assert relmax(inv.p.download(L), po.p1_lumped_inv_mass(om, L)) < TOL
    for f in (src, dst, inv):
        f.destroy()
    A.destroy()


@pytest.mark.parametrize("L", [1, 2])
def test_shell_apply_parity(ctx, shell_pair, L):
    _apply_parity(*shell_pair, L, VISC)


@pytest.mark.parametrize("L", [2, 3])
def test_cube_apply_parity(ctx, cube_pair, L):
    _apply_parity(*cube_pair, L, (0.2, 0.7))


def test_pressure_sum_and_constant(ctx, shell_pair):
    st, om = shell_pair
    L = 2
    f = hg.P1Function("p", st, L, L)
    x = np.random.default_rng(5).uniform(-1, 1, st.ndofs(L))
    f.upload(L, x)
    assert abs(f.sum(L) - x.sum()) < 1e-12 * np.abs(x).sum()
    th = hg.P2P1TaylorHoodFunction("th", st, L, L)
    th.p.upload(L, x)
    th.project_mean_pressure(L)
    assert np.abs(th.p.download(L) - (x - x.mean())).max() < 1e-13
    f.destroy()
    th.destroy()
