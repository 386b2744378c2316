"""ctypes loader and signatures of the C ABI declared in include/hyteg_amd.h."""
import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HEADER = os.path.join(ROOT, "include", "hyteg_amd.h")

INNER, DIRICHLET, NEUMANN, FREESLIP, ALL = 1, 2, 4, 8, 15
GMG_FLAG = INNER | NEUMANN | FREESLIP
REPLACE, ADD = 0, 1
F64, F32, F16 = 0, 1, 2

_p = C.c_void_p
_i = C.c_int
_i64 = C.c_int64
_d = C.c_double

SIGNATURES = {
    "hg_ctx_create": [_i, _p],
    "hg_ctx_destroy": [_p],
    "hg_ctx_synchronize": [_p],
    "hg_graph_begin_capture": [_p],
    "hg_graph_end_capture": [_p, _p],
    "hg_graph_launch": [_p, _i64],
    "hg_graph_destroy": [_p, _i64],
    "hg_ctx_set_graphs": [_p, _i],
    "hg_ctx_graphs_enabled": [_p, _p],
    "hg_ctx_memory": [_p, _p],
    "hg_ctx_event_timing": [_p, _i],
    "hg_ctx_last_kernel_ms": [_p, _p],
    "hg_storage_cuboid": [_p, _p, _p, _i, _i, _i, _p],
    "hg_storage_create": [_p, _i, _p, _i, _p, _p],
    "hg_storage_counts": [_p, _i64, _p],
    "hg_p1_ndofs": [_p, _i64, _i, _p],
    "hg_p1_dof_coords": [_p, _i64, _i, _p],
    "hg_p1_dof_flags": [_p, _i64, _i, _p],
    "hg_p1_function_create": [_p, _i64, _i, _i, _i, _p],
    "hg_function_destroy": [_p, _i64],
    "hg_function_storage": [_p, _i64, _p],
    "hg_function_upload": [_p, _i64, _i, _p],
    "hg_function_download": [_p, _i64, _i, _p],
    "hg_function_assign": [_p, _i64, _i, _p, _p, _i, _i],
    "hg_function_add": [_p, _i64, _i, _p, _p, _i, _i],
    "hg_function_interpolate_const": [_p, _i64, _d, _i, _i],
    "hg_function_dot": [_p, _i64, _i64, _i, _i, _p],
    "hg_function_interpolate_random": [_p, _i64, C.c_uint64, _d, _d, _i, _i],
    "hg_function_copy_from": [_p, _i64, _i64, _i],
    "hg_function_communicate": [_p, _i64, _i],
    "hg_p1_laplace_create": [_p, _i64, _i, _i, _p],
    "hg_storage_shell": [_p, _i, _i, _d, _d, _p],
    "hg_storage_shell_bc": [_p, _i, _i, _d, _d, _i, _i, _p],
    "hg_storage_blend": [_p, _i64, _p],
    "hg_storage_cell_vertices": [_p, _i64, _p],
    "hg_storage_vertex_coords": [_p, _i64, _p],
    "hg_p2_ndofs": [_p, _i64, _i, _p],
    "hg_p2_dof_coords": [_p, _i64, _i, _p],
    "hg_p2_dof_flags": [_p, _i64, _i, _p],
    "hg_p2_function_create": [_p, _i64, _i, _i, _i, _p],
    "hg_function_mult": [_p, _i64, _i64, _i64, _i, _i],
    "hg_p2_divkgrad_create": [_p, _i64, _i, _i, _i, _p, _p, _p, _p],
    "hg_p2_divkgrad_create_batched": [_p, _i64, _i, _i, _i, _p],
    "hg_p2_divkgrad_set_slot": [_p, _i64, _i, _i, _p, _p, _p],
    "hg_p2_divkgrad_slot_active": [_p, _i64, _i, _i],
    "hg_p2_function_create_batched": [_p, _i64, _i, _i, _i, _i, _p],
    "hg_p2_apply": [_p, _i64, _i64, _i64, _i, _i, _i],
    "hg_p2_inverse_diagonal": [_p, _i64, _i, _i64],
    "hg_p2_store_coefficients": [_p, _i64, _i],
    "hg_operator_destroy": [_p, _i64],
    "hg_p1_apply": [_p, _i64, _i64, _i64, _i, _i, _i],
    "hg_p1_smooth_sor": [_p, _i64, _i64, _i64, _d, _i, _i],
    "hg_p1_cell_stencil": [_p, _i64, _i, _i, _p],
    "hg_p1_restrict": [_p, _i64, _i, _i],
    "hg_p1_function_create_batched": [_p, _i64, _i, _i, _i, _i, _p],
    "hg_function_nvec": [_p, _i64, _p],
    "hg_serve_stream": [_i, _i64, _p, _p, _p, _p, _d, _i, _d, _p, _p, _p, _p, _p],
    "hg_function_sum": [_p, _i64, _i, _i, _p],
    "hg_function_add_const": [_p, _i64, _d, _i, _i],
    "hg_stokes_create": [_p, _i64, _i, _i, _d, _d, _p],
    "hg_stokes_apply": [_p, _i64, _p, _p, _i, _i, _i, _i],
    "hg_stokes_velocity_inverse_diagonal": [_p, _i64, _i, _p],
    "hg_stokes_pressure_diagonal": [_p, _i64, _i, _i, _i64],
    "hg_p2_project_normal": [_p, _p, _i, _i],
    "hg_function_invert": [_p, _i64, _i, _i],
    "hg_vtk_write": [_p, C.c_char_p, _i64, _i, _i, _p, _p, _i],
    "hg_p2_const_create": [_p, _i64, _i, _i, _p],
    "hg_p2_const_apply": [_p, _i64, _i64, _i64, _i, _i, _i],
    "hg_p2_const_smooth_sor": [_p, _i64, _i64, _i64, _i, _i, _d],
    "hg_p2_const_inverse_diagonal": [_p, _i64, _i, _i64],
    "hg_p2_const_gmg_solve": [_p, _i64, _i64, _i64, _i, _i, _i, _i, _i, _d, _i, _d, _p, _p],
    "hg_stokes_fgmres_solve": [_p, _i64, _p, _p, _i, _i, _i, _i, _i, _d, _i, _d, _i, _d, _i, _i, _i, _i, _d, _p, _p],
    "hg_stokes_freeslip_apply": [_p, _i64, _p, _p, _i, _i, _i, _i],
    "hg_stokes_estimate_radius": [_p, _i64, _p, _p, _i, _i, _i, _p],
    "hg_stokes_uzawa_smooth": [_p, _i64, _p, _p, _i, _d, _i, _i, _d, _i, _i],
    "hg_stokes_minres_solve": [_p, _i64, _p, _p, _i, _i, _d, _d, _i, _p],
    "hg_stokes_gmg_solve": [_p, _i64, _p, _p, _i, _i, _i, _i, _i, _d, _i, _i, _d, _i, _d, _i, _i, _i, _d, _p, _p],
    "hg_function_copy_vector": [_p, _i64, _i, _i64, _i, _i],
    "hg_function_dot_batched": [_p, _i64, _i64, _i, _i, _p],
    "hg_function_dot_batched_noalloc": [_p, _i64, _i64, _i, _i, _p],
    "hg_function_scale_batched": [_p, _i64, _p, _i, _i],
    "hg_function_chebyshev_step": [_p, _i64, _i64, _i64, _i64, _i64, _d, _i, _i],
    "hg_function_chebyshev_step_slots": [_p, _i64, _i64, _i64, _i64, _i64, _i, _i, _i],
    "hg_slots_alloc": [_p, _i, _p],
    "hg_slots_free": [_p, _i, _i],
    "hg_slots_read": [_p, _i, _i, _p],
    "hg_slots_write": [_p, _i, _i, _p],
    "hg_function_dot_slot": [_p, _i64, _i64, _i, _i, _i],
    "hg_function_lincomb_slots": [_p, _i64, _i, _p, _p, _p, _p, _i, _i, _i],
    "hg_p1_batched_gmg_solve": [_p, _i64, _i64, _i64, _i, _i, _i, _i, _i, _i, _d, _p, _p],
    "hg_p1_batched_ir_solve": [_p, _i64, _i64, _i64, _i, _i, _i, _i, _i, _i, _i, _i, _d, _p, _p],
    "hg_p2_restrict": [_p, _i64, _i, _i],
    "hg_p2_prolongate": [_p, _i64, _i, _i, _i],
    "hg_p2_transfer_stencil": [_i, _i, _i, _p, _p, _p],
    "hg_p2_cg_solve": [_p, _i64, _i64, _i64, _i, _i, _d, _d, _i, _p],
    "hg_p2_estimate_radius": [_p, _i64, _i64, _i, _i, _p],
    "hg_p2_chebyshev_smooth": [_p, _i64, _i64, _i64, _i, _i, _d, _i],
    "hg_p2_ir_solve": [_p, _i64, _i64, _i64, _i, _i, _i, _i, _i, _i, _d, _i, _i, _d, _i, _d, _p, _p],
    "hg_p2_gmg_solve": [_p, _i64, _i64, _i64, _i, _i, _i, _i, _i, _d, _i, _d, _i, _d, _i, _p, _p],
    "hg_p1_prolongate": [_p, _i64, _i, _i, _i],
    "hg_p1_cg_solve": [_p, _i64, _i64, _i64, _i, _i, _d, _d, _p],
    "hg_p1_gmg_solve": [_p, _i64, _i64, _i64, _i, _i, _i, _i, _i, _d, _i, _d, _p, _p],
    "hg_p1_smooth_sor_backwards": [_p, _i64, _i64, _i64, _d, _i, _i],
    "hg_p1_sor_colours": [_p, _i64, _i, _p, _p],
    "hg_p1_inverse_diagonal": [_p, _i64, _i, _i64],
    "hg_p1_smooth_jac": [_p, _i64, _i64, _i64, _i64, _d, _i, _i],
    "hg_p1_gemv": [_p, _i64, _d, _i64, _d, _i64, _i, _i],
    "hg_p2_gemv": [_p, _i64, _d, _i64, _d, _i64, _i, _i],
    "hg_p2_smooth_jac": [_p, _i64, _i64, _i64, _i64, _d, _i, _i],
    "hg_p2_const_smooth_sor_backwards": [_p, _i64, _i64, _i64, _i, _i, _d],
    "hg_p2_const_sor_colours": [_p, _i64, _i, _p, _p],
    "hg_function_lincomb_slots_gated": [_p, _i64, _i, _p, _p, _p, _p, _i, _i, _i, _i],
    "hg_slots_cg_stop": [_p, _i, _i, _d, _d, _i, _i],
    "hg_serve_stream_solutions": [_i, _i64, _p, _p, _p, _p, _d, _i, _d, _p, _p, _p, _p, _p, _p],
    "hg_serve_request_coefficient": [C.c_uint64, _p, _p, _p],
    "hg_serve_request_blobs": [C.c_uint64, _p],
    "hg_p2_low_precision_form": [_p, _i64, _i, _i],
    "hg_timing_enable": [_p, _i],
    "hg_timing_start": [_p, C.c_char_p],
    "hg_timing_stop": [_p, C.c_char_p],
    "hg_timing_report": [_p, C.c_char_p, _i64, _p],
    "hg_timing_reset": [_p],
}

_LIB = None


class HygError(RuntimeError):
    pass


def library_path():
    return os.path.join(HERE, "libhyteg_amd.so")


def declared_symbols():
    """names of all functions declared in include/hyteg_amd.h"""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(hg_[a-z0-9_]+)\s*\(", text)))


def lib():
    """Load libhyteg_amd.so; raises if it was not built (no CPU fallback)."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise HygError(f"{path} is missing: build it with `python hyteg_amd/build.py`")
        L = C.CDLL(path)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _i
        L.hg_last_error.argtypes = [_p]
        L.hg_last_error.restype = C.c_char_p
        L.hg_serve_last_error.argtypes = []
        L.hg_serve_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB
