"""hyteg_amd -- MI355X-native matrix-free operator application and geometric
multigrid for hierarchical hybrid tetrahedral grids (HyTeG's hot path).

The numerical work runs in hand-written HIP kernels inside
hyteg_amd/libhyteg_amd.so, reached through its C ABI (include/hyteg_amd.h).
This module is a thin ctypes mirror of that ABI with HyTeG's names
(PrimitiveStorage, P1Function, P1ConstantLaplaceOperator, ...).  There is no
CPU fallback: if the library is missing, importing `lib()` raises.
"""
from ._lib import (ALL, DIRICHLET, F16, F32, F64, FREESLIP, GMG_FLAG, INNER, NEUMANN, ADD, REPLACE,
                   HygError, lib, library_path)
from .api import (Context, P1ConstantLaplaceOperator, P1Function, P2DivKGradOperator, P2Function,
                  P2P1StokesOperator, P2P1TaylorHoodFunction, P2ProjectNormalOperator, PrimitiveStorage, batched_gmg_solve, batched_ir_solve, cg_solve, gmg_solve, p2_cg_solve, p2_chebyshev_smooth, p2_estimate_radius,
                  p2_gmg_solve, p2_ir_solve, p2_transfer_stencil, prolongate, restrict, stokes_estimate_radius,
                  stokes_fgmres_solve, stokes_gmg_solve, stokes_minres_solve, stokes_uzawa_smooth, vtk_write,
                  P2ConstantLaplaceOperator, p2_const_gmg_solve)

__all__ = [
    "ALL", "DIRICHLET", "F16", "F32", "F64", "FREESLIP", "GMG_FLAG", "INNER", "NEUMANN", "ADD", "REPLACE",
    "HygError", "lib", "library_path", "Context", "P1ConstantLaplaceOperator", "P1Function",
    "P2DivKGradOperator", "P2Function", "P2P1StokesOperator", "P2P1TaylorHoodFunction",
    "P2ProjectNormalOperator", "PrimitiveStorage", "batched_gmg_solve", "batched_ir_solve", "cg_solve", "gmg_solve", "p2_cg_solve", "p2_chebyshev_smooth", "p2_estimate_radius",
    "p2_gmg_solve", "p2_ir_solve", "p2_transfer_stencil", "prolongate", "restrict", "stokes_estimate_radius",
    "stokes_fgmres_solve", "stokes_gmg_solve", "stokes_minres_solve", "stokes_uzawa_smooth", "vtk_write",
    "P2ConstantLaplaceOperator", "p2_const_gmg_solve",
]
