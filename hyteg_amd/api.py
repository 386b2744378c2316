"""Python mirror of the HyTeG-shaped API over the C ABI (ctypes)."""
import ctypes as C

import numpy as np

from ._lib import ALL, DIRICHLET, F16, F64, FREESLIP, GMG_FLAG, INNER, REPLACE, HygError, lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """One HIP device, one stream (hg_ctx)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        st = lib().hg_ctx_create(device, C.byref(h))
        if st != 0:
            raise HygError(f"hg_ctx_create failed with status {st}")
        self.h = h

    def check(self, status, what):
        if status != 0:
            msg = lib().hg_last_error(self.h).decode()
            raise HygError(f"{what}: status {status}: {msg}")

    def synchronize(self):
        self.check(lib().hg_ctx_synchronize(self.h), "synchronize")

    def memory(self):
        out = C.c_int64()
        self.check(lib().hg_ctx_memory(self.h, C.byref(out)), "memory")
        return out.value

    def event_timing(self, enable=True):
        self.check(lib().hg_ctx_event_timing(self.h, int(enable)), "event_timing")

    def last_ms(self):
        out = C.c_double()
        self.check(lib().hg_ctx_last_kernel_ms(self.h, C.byref(out)), "last_ms")
        return out.value

    def timing(self, enable=True):
        """record the solvers' named timing regions (HIP events on the context stream)"""
        self.check(lib().hg_timing_enable(self.h, int(enable)), "timing")

    def timing_report(self, reset=False):
        """{region path: (count, total device ms)} of the recorded regions"""
        need = C.c_int64()
        self.check(lib().hg_timing_report(self.h, None, 0, C.byref(need)), "timing report")
        buf = C.create_string_buffer(need.value)
        self.check(lib().hg_timing_report(self.h, buf, need.value, None), "timing report")
        if reset:
            self.check(lib().hg_timing_reset(self.h), "timing reset")
        out = {}
        for line in buf.value.decode().splitlines():
            path, count, ms = line.split("\t")
            out[path] = (int(count), float(ms))
        return out

    def set_graphs(self, enable=True):
        """let the batched multigrid solvers replay their V-cycle from a captured HIP graph"""
        self.check(lib().hg_ctx_set_graphs(self.h, int(enable)), "set_graphs")

    def close(self):
        if getattr(self, "h", None):
            lib().hg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PrimitiveStorage:
    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle

    @staticmethod
    def cuboid(ctx, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(1, 1, 1)):
        lo = np.asarray(lo, dtype=np.float64)
        hi = np.asarray(hi, dtype=np.float64)
        h = C.c_int64()
        ctx.check(lib().hg_storage_cuboid(ctx.h, _ptr(lo), _ptr(hi), n[0], n[1], n[2], C.byref(h)), "meshCuboid")
        return PrimitiveStorage(ctx, h.value)

    @staticmethod
    def shell(ctx, ntan=5, nrad=3, rmin=0.55, rmax=1.0, inner_flag=None, outer_flag=None):
        """meshSphericalShell with the icosahedral shell map; inner_flag / outer_flag
        (DIRICHLET, NEUMANN or FREESLIP) set the boundary type of each surface"""
        h = C.c_int64()
        if inner_flag is None and outer_flag is None:
            ctx.check(lib().hg_storage_shell(ctx.h, ntan, nrad, rmin, rmax, C.byref(h)), "meshSphericalShell")
        else:
            ctx.check(lib().hg_storage_shell_bc(ctx.h, ntan, nrad, rmin, rmax, inner_flag or DIRICHLET,
                                                outer_flag or DIRICHLET, C.byref(h)), "meshSphericalShell")
        return PrimitiveStorage(ctx, h.value)

    def blend(self):
        out = np.zeros((self.counts()[3], 14))
        self.ctx.check(lib().hg_storage_blend(self.ctx.h, self.h, _ptr(out)), "blend")
        return out

    def cell_vertices(self):
        out = np.zeros((self.counts()[3], 4), dtype=np.int64)
        self.ctx.check(lib().hg_storage_cell_vertices(self.ctx.h, self.h, _ptr(out)), "cell_vertices")
        return out

    def vertex_coords(self):
        out = np.zeros((self.counts()[0], 3))
        self.ctx.check(lib().hg_storage_vertex_coords(self.ctx.h, self.h, _ptr(out)), "vertex_coords")
        return out

    def p2_ndofs(self, level):
        return self.ndofs(level + 1)

    @staticmethod
    def from_cells(ctx, coords, cells):
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        cells = np.ascontiguousarray(cells, dtype=np.int64)
        h = C.c_int64()
        ctx.check(lib().hg_storage_create(ctx.h, len(coords), _ptr(coords), len(cells), _ptr(cells), C.byref(h)),
                  "storage")
        return PrimitiveStorage(ctx, h.value)

    def counts(self):
        out = np.zeros(4, dtype=np.int64)
        self.ctx.check(lib().hg_storage_counts(self.ctx.h, self.h, _ptr(out)), "counts")
        return tuple(int(v) for v in out)

    def ndofs(self, level):
        out = C.c_int64()
        self.ctx.check(lib().hg_p1_ndofs(self.ctx.h, self.h, level, C.byref(out)), "ndofs")
        return out.value

    def coords(self, level):
        out = np.zeros((self.ndofs(level), 3))
        self.ctx.check(lib().hg_p1_dof_coords(self.ctx.h, self.h, level, _ptr(out)), "coords")
        return out

    def flags(self, level):
        out = np.zeros(self.ndofs(level), dtype=np.int32)
        self.ctx.check(lib().hg_p1_dof_flags(self.ctx.h, self.h, level, _ptr(out)), "flags")
        return out


class P1Function:
    OFFSET = 0

    def __init__(self, name, storage, min_level, max_level, dtype=F64, nvec=1):
        """nvec > 1: that many independent vectors interleaved per DoF"""
        self.name = name
        self.storage = storage
        self.ctx = storage.ctx
        self.min_level, self.max_level = min_level, max_level
        self.nvec = nvec
        h = C.c_int64()
        if nvec != 1:
            create = lib().hg_p2_function_create_batched if self.OFFSET else lib().hg_p1_function_create_batched
            self.ctx.check(create(self.ctx.h, storage.h, min_level, max_level, dtype, nvec, C.byref(h)),
                           f"Function {name}")
        else:
            create = lib().hg_p2_function_create if self.OFFSET else lib().hg_p1_function_create
            self.ctx.check(create(self.ctx.h, storage.h, min_level, max_level, dtype, C.byref(h)), f"Function {name}")
        self.h = h.value

    def ndofs(self, level):
        return self.storage.ndofs(level + self.OFFSET)

    def destroy(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().hg_function_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def _shape(self, level):
        return (self.ndofs(level),) if self.nvec == 1 else (self.ndofs(level), self.nvec)

    def upload(self, level, values):
        """values: (ndofs,) or, for interleaved functions, (ndofs, nvec)"""
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.shape != self._shape(level):
            raise ValueError("wrong number of values")
        self.ctx.check(lib().hg_function_upload(self.ctx.h, self.h, level, _ptr(v)), "upload")

    def download(self, level):
        out = np.zeros(self._shape(level))
        self.ctx.check(lib().hg_function_download(self.ctx.h, self.h, level, _ptr(out)), "download")
        return out

    def _lincomb(self, fn, scalars, funcs, level, flag):
        s = np.ascontiguousarray(scalars, dtype=np.float64)
        f = np.ascontiguousarray([g.h for g in funcs], dtype=np.int64)
        self.ctx.check(fn(self.ctx.h, self.h, len(funcs), _ptr(s), _ptr(f), level, flag), "assign/add")

    def assign(self, scalars, funcs, level, flag=ALL):
        self._lincomb(lib().hg_function_assign, scalars, funcs, level, flag)

    def add(self, scalars, funcs, level, flag=ALL):
        self._lincomb(lib().hg_function_add, scalars, funcs, level, flag)

    def interpolate(self, value, level, flag=ALL):
        self.ctx.check(lib().hg_function_interpolate_const(self.ctx.h, self.h, float(value), level, flag),
                       "interpolate")

    def interpolate_random(self, seed, lo, hi, level, flag=ALL):
        self.ctx.check(lib().hg_function_interpolate_random(self.ctx.h, self.h, seed, lo, hi, level, flag),
                       "interpolate_random")

    def dot(self, other, level, flag=ALL):
        if self.nvec != 1:
            out = np.zeros(self.nvec)
            self.ctx.check(lib().hg_function_dot_batched(self.ctx.h, self.h, other.h, level, flag, _ptr(out)), "dot")
            return out
        out = C.c_double()
        self.ctx.check(lib().hg_function_dot(self.ctx.h, self.h, other.h, level, flag, C.byref(out)), "dot")
        return out.value

    def scale(self, scales, level, flag=ALL):
        """vector k *= scales[k] (interleaved functions)"""
        sc = np.ascontiguousarray(scales, dtype=np.float64)
        if sc.shape != (self.nvec,):
            raise ValueError("one scale per vector")
        self.ctx.check(lib().hg_function_scale_batched(self.ctx.h, self.h, _ptr(sc), level, flag), "scale")

    def copy_from(self, other, level):
        self.ctx.check(lib().hg_function_copy_from(self.ctx.h, self.h, other.h, level), "copyFrom")

    def communicate(self, level):
        self.ctx.check(lib().hg_function_communicate(self.ctx.h, self.h, level), "communicate")

    def mult(self, a, b, level, flag=ALL):
        self.ctx.check(lib().hg_function_mult(self.ctx.h, self.h, a.h, b.h, level, flag), "multElementwise")

    def sum(self, level, flag=ALL):
        """sumGlobal (FP64 single-vector functions)"""
        out = C.c_double()
        self.ctx.check(lib().hg_function_sum(self.ctx.h, self.h, level, flag, C.byref(out)), "sum")
        return out.value

    def add_const(self, value, level, flag=ALL):
        self.ctx.check(lib().hg_function_add_const(self.ctx.h, self.h, float(value), level, flag), "add")


class P2Function(P1Function):
    """P2 DoFs of level L, numbered like the P1 DoFs of level L+1"""
    OFFSET = 1


class P2DivKGradOperator:
    """P2ElementwiseOperator with the blended div(k grad) form,
    k(x) = exp(sum_m a_m sin(omega_m . x + phi_m))"""

    def __init__(self, storage, min_level, max_level, a=(), omega=(), phi=()):
        self.storage = storage
        self.ctx = storage.ctx
        self.a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        self.omega = np.ascontiguousarray(omega, dtype=np.float64).reshape(-1, 3)
        self.phi = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1)
        h = C.c_int64()
        e = np.zeros(1)
        nm = len(self.a)
        pa, po, pp = (_ptr(self.a), _ptr(self.omega), _ptr(self.phi)) if nm else (_ptr(e), _ptr(e), _ptr(e))
        self.ctx.check(lib().hg_p2_divkgrad_create(self.ctx.h, storage.h, min_level, max_level, nm, pa, po, pp,
                                                   C.byref(h)), "P2DivKGradOperator")
        self.h = h.value

    def destroy(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().hg_operator_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def apply(self, src, dst, level, flag, update=REPLACE):
        self.ctx.check(lib().hg_p2_apply(self.ctx.h, self.h, src.h, dst.h, level, flag, update), "apply")

    def inverse_diagonal(self, level, dst):
        self.ctx.check(lib().hg_p2_inverse_diagonal(self.ctx.h, self.h, level, dst.h), "inverse diagonal")

    def gemv(self, alpha, src, beta, dst, level, flag):
        """dst = alpha A src + beta dst on DoFs of type flag"""
        self.ctx.check(lib().hg_p2_gemv(self.ctx.h, self.h, alpha, src.h, beta, dst.h, level, flag), "gemv")

    def apply_scaled(self, alpha, src, dst, level, flag, update=REPLACE):
        self.gemv(alpha, src, 0.0 if update == REPLACE else 1.0, dst, level, flag)

    def smooth_jac(self, dst, rhs, src, relax, level, flag):
        """dst = src + relax D^-1 (rhs - A src)"""
        self.ctx.check(lib().hg_p2_smooth_jac(self.ctx.h, self.h, dst.h, rhs.h, src.h, relax, level, flag),
                       "smooth_jac")

    def low_precision_form(self, reduced=True, from_level=2):
        """element form of FP32 / FP16 applies: reduced (affine micro-cells, k constant,
        4-point rule) on levels >= from_level, or the reference form"""
        self.ctx.check(lib().hg_p2_low_precision_form(self.ctx.h, self.h, int(reduced), from_level),
                       "low_precision_form")

    def store_coefficients(self, enable=True):
        """keep k at the quadrature points (default) or evaluate it in every apply"""
        self.ctx.check(lib().hg_p2_store_coefficients(self.ctx.h, self.h, int(enable)), "store_coefficients")


class P1ConstantLaplaceOperator:
    def __init__(self, storage, min_level, max_level):
        self.storage = storage
        self.ctx = storage.ctx
        h = C.c_int64()
        self.ctx.check(lib().hg_p1_laplace_create(self.ctx.h, storage.h, min_level, max_level, C.byref(h)),
                       "P1ConstantLaplaceOperator")
        self.h = h.value

    def apply(self, src, dst, level, flag, update=REPLACE):
        self.ctx.check(lib().hg_p1_apply(self.ctx.h, self.h, src.h, dst.h, level, flag, update), "apply")

    def smooth_sor(self, x, b, relax, level, flag):
        self.ctx.check(lib().hg_p1_smooth_sor(self.ctx.h, self.h, x.h, b.h, relax, level, flag), "smooth_sor")

    def smooth_gs(self, x, b, level, flag):
        self.smooth_sor(x, b, 1.0, level, flag)

    def smooth_sor_backwards(self, x, b, relax, level, flag):
        self.ctx.check(lib().hg_p1_smooth_sor_backwards(self.ctx.h, self.h, x.h, b.h, relax, level, flag),
                       "smooth_sor_backwards")

    def smooth_gs_backwards(self, x, b, level, flag):
        self.smooth_sor_backwards(x, b, 1.0, level, flag)

    def sor_colours(self, level):
        """(colour of every canonical DoF in sweep order, number of boundary colours)"""
        n = self.storage.ndofs(level)
        col = np.zeros(n, dtype=np.int32)
        nb = C.c_int()
        self.ctx.check(lib().hg_p1_sor_colours(self.ctx.h, self.h, level, _ptr(col), C.byref(nb)), "sor_colours")
        return col, nb.value

    def inverse_diagonal(self, level, dst):
        self.ctx.check(lib().hg_p1_inverse_diagonal(self.ctx.h, self.h, level, dst.h), "inverse diagonal")

    def smooth_jac(self, dst, rhs, src, relax, level, flag):
        """dst = src + relax D^-1 (rhs - A src)"""
        self.ctx.check(lib().hg_p1_smooth_jac(self.ctx.h, self.h, dst.h, rhs.h, src.h, relax, level, flag),
                       "smooth_jac")

    def gemv(self, alpha, src, beta, dst, level, flag):
        """dst = alpha A src + beta dst on DoFs of type flag"""
        self.ctx.check(lib().hg_p1_gemv(self.ctx.h, self.h, alpha, src.h, beta, dst.h, level, flag), "gemv")

    def apply_scaled(self, alpha, src, dst, level, flag, update=REPLACE):
        self.gemv(alpha, src, 0.0 if update == REPLACE else 1.0, dst, level, flag)

    def cell_stencil(self, level, cell):
        out = np.zeros(15)
        self.ctx.check(lib().hg_p1_cell_stencil(self.ctx.h, self.h, level, cell, _ptr(out)), "cell_stencil")
        return out


class P2DivKGradOperatorBatched(P2DivKGradOperator):
    """nslot coefficient slots for P2 functions of nslot interleaved vectors:
    vector s is multiplied by the div(k_s grad) operator of slot s (k_s = 1 until
    set_slot gives it its own modes)"""

    def __init__(self, storage, min_level, max_level, nslot):
        self.storage = storage
        self.ctx = storage.ctx
        self.nslot = nslot
        h = C.c_int64()
        self.ctx.check(lib().hg_p2_divkgrad_create_batched(self.ctx.h, storage.h, min_level, max_level, nslot,
                                                           C.byref(h)), "P2DivKGradOperatorBatched")
        self.h = h.value

    def set_slot(self, slot, a=(), omega=(), phi=()):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        omega = np.ascontiguousarray(omega, dtype=np.float64).reshape(-1, 3)
        phi = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1)
        e = np.zeros(1)
        nm = len(a)
        pa, po, pp = (_ptr(a), _ptr(omega), _ptr(phi)) if nm else (_ptr(e), _ptr(e), _ptr(e))
        self.ctx.check(lib().hg_p2_divkgrad_set_slot(self.ctx.h, self.h, slot, nm, pa, po, pp), "set_slot")


def restrict(f, source_level, flag):
    """P1toP1LinearRestriction / P2toP2QuadraticRestriction by the function's space"""
    fn = lib().hg_p2_restrict if isinstance(f, P2Function) else lib().hg_p1_restrict
    f.ctx.check(fn(f.ctx.h, f.h, source_level, flag), "restrict")


def prolongate(f, source_level, flag, update):
    """P1toP1LinearProlongation / P2toP2QuadraticProlongation by the function's space"""
    fn = lib().hg_p2_prolongate if isinstance(f, P2Function) else lib().hg_p1_prolongate
    f.ctx.check(fn(f.ctx.h, f.h, source_level, flag, update), "prolongate")


def p2_transfer_stencil(kind, cls):
    """(offsets (n,3), weights (n,)) of a P2 transfer stencil (kind 0 pull, 1 restriction)"""
    n = C.c_int()
    off = np.zeros((128, 3), dtype=np.int32)
    w = np.zeros(128)
    if lib().hg_p2_transfer_stencil(kind, cls, 128, _ptr(off), _ptr(w), C.byref(n)) != 0:
        raise ValueError("invalid stencil class")
    return off[: n.value].copy(), w[: n.value].copy()


def p2_cg_solve(A, x, b, level, max_iter=10000, rel_tol=1e-14, abs_tol=1e-300, jacobi=True):
    it = C.c_int()
    A.ctx.check(lib().hg_p2_cg_solve(A.ctx.h, A.h, x.h, b.h, level, max_iter, rel_tol, abs_tol, int(jacobi),
                                     C.byref(it)), "p2 cg")
    return it.value


def p2_chebyshev_smooth(A, x, b, level, order, rho, steps=1):
    A.ctx.check(lib().hg_p2_chebyshev_smooth(A.ctx.h, A.h, x.h, b.h, level, order, rho, steps), "chebyshev")


def p2_estimate_radius(A, x, level, iterations):
    r = C.c_double()
    A.ctx.check(lib().hg_p2_estimate_radius(A.ctx.h, A.h, x.h, level, iterations, C.byref(r)), "estimate radius")
    return r.value


def batched_gmg_solve(A, x, b, min_level, max_level, pre=3, post=3, coarse_sweeps=10, cycles=30, tol=0.0):
    """P1 GMG V(pre,post) with Gauss-Seidel for all interleaved vectors of x, b; returns
    (cycles, residual norms of shape (cycles+1, nvec))"""
    hist = np.zeros((cycles + 1, x.nvec))
    done = C.c_int()
    A.ctx.check(lib().hg_p1_batched_gmg_solve(A.ctx.h, A.h, x.h, b.h, min_level, max_level, pre, post, coarse_sweeps,
                                              cycles, tol, _ptr(hist), C.byref(done)), "batched gmg")
    return done.value, hist[: done.value + 1]


def batched_ir_solve(A, x, b, min_level, max_level, low_dtype=F16, pre=3, post=3, coarse_sweeps=10, inner_cycles=1,
                     max_outer=40, tol=1e-10):
    """FP64 iterative refinement around the low-precision batched V-cycle"""
    hist = np.zeros((max_outer + 1, x.nvec))
    done = C.c_int()
    A.ctx.check(lib().hg_p1_batched_ir_solve(A.ctx.h, A.h, x.h, b.h, min_level, max_level, low_dtype, pre, post,
                                             coarse_sweeps, inner_cycles, max_outer, tol, _ptr(hist),
                                             C.byref(done)), "batched ir")
    return done.value, hist[: done.value + 1]


def p2_ir_solve(A, x, b, min_level, max_level, rho, low_dtype=F16, pre=2, post=2, order=3, inner_cycles=1,
                max_outer=40, tol=1e-8, coarse_max_iter=10000, coarse_tol=1e-14):
    """FP64 iterative refinement around a reduced-precision (FP32 arithmetic, FP32/FP16
    storage) V-cycle; returns (outer steps, FP64 Inner residual norms)"""
    hist = np.zeros(max_outer + 1)
    done = C.c_int()
    A.ctx.check(lib().hg_p2_ir_solve(A.ctx.h, A.h, x.h, b.h, min_level, max_level, low_dtype, pre, post, order, rho,
                                     inner_cycles, max_outer, tol, coarse_max_iter, coarse_tol, _ptr(hist),
                                     C.byref(done)), "p2 iterative refinement")
    return done.value, hist[: done.value + 1]


def p2_gmg_solve(A, x, b, min_level, max_level, rho, pre=2, post=2, order=3, cycles=20, tol=0.0,
                 coarse_max_iter=10000, coarse_tol=1e-14, coarse_jacobi=True):
    hist = np.zeros(cycles + 1)
    done = C.c_int()
    A.ctx.check(lib().hg_p2_gmg_solve(A.ctx.h, A.h, x.h, b.h, min_level, max_level, pre, post, order, rho, cycles,
                                      tol, coarse_max_iter, coarse_tol, int(coarse_jacobi), _ptr(hist),
                                      C.byref(done)), "p2 gmg")
    return done.value, hist[: done.value + 1]


def cg_solve(A, x, b, level, max_iter, rel_tol, abs_tol):
    it = C.c_int()
    A.ctx.check(lib().hg_p1_cg_solve(A.ctx.h, A.h, x.h, b.h, level, max_iter, rel_tol, abs_tol, C.byref(it)), "cg")
    return it.value


def gmg_solve(A, x, b, min_level, max_level, pre=3, post=3, cycles=10, tol=0.0, coarse_max_iter=10000,
              coarse_tol=1e-16):
    hist = np.zeros(cycles + 1)
    done = C.c_int()
    A.ctx.check(lib().hg_p1_gmg_solve(A.ctx.h, A.h, x.h, b.h, min_level, max_level, pre, post, cycles, tol,
                                      coarse_max_iter, coarse_tol, _ptr(hist), C.byref(done)), "gmg")
    return done.value, hist[: done.value + 1]


def _pressure_flag(flag):
    """the pressure of a Taylor-Hood function has no boundary DoFs (createAllInnerBC)"""
    return ALL if flag & INNER else 0


class P2P1TaylorHoodFunction:
    """velocity uvw (three P2 functions) and pressure p (P1, every DoF Inner)"""

    def __init__(self, name, storage, min_level, max_level):
        self.storage = storage
        self.ctx = storage.ctx
        self.uvw = [P2Function(f"{name}_{c}", storage, min_level, max_level) for c in "uvw"]
        self.p = P1Function(f"{name}_p", storage, min_level, max_level)

    def handles(self):
        return np.ascontiguousarray([f.h for f in self.uvw] + [self.p.h], dtype=np.int64)

    def destroy(self):
        for f in self.uvw + [self.p]:
            f.destroy()

    def upload(self, level, u, p):
        """u: [u_x | u_y | u_z] (3 * P2 DoFs), p: P1 DoFs"""
        n2 = self.uvw[0].ndofs(level)
        u = np.asarray(u, dtype=np.float64).reshape(3, n2)
        for c in range(3):
            self.uvw[c].upload(level, u[c])
        self.p.upload(level, p)

    def download(self, level):
        return np.concatenate([f.download(level) for f in self.uvw]), self.p.download(level)

    def assign(self, scalars, funcs, level, flag=ALL):
        for c in range(3):
            self.uvw[c].assign(scalars, [f.uvw[c] for f in funcs], level, flag)
        if _pressure_flag(flag):
            self.p.assign(scalars, [f.p for f in funcs], level, ALL)

    def add(self, scalars, funcs, level, flag=ALL):
        for c in range(3):
            self.uvw[c].add(scalars, [f.uvw[c] for f in funcs], level, flag)
        if _pressure_flag(flag):
            self.p.add(scalars, [f.p for f in funcs], level, ALL)

    def interpolate(self, value, level, flag=ALL):
        for c in range(3):
            self.uvw[c].interpolate(value, level, flag)
        if _pressure_flag(flag):
            self.p.interpolate(value, level, ALL)

    def dot(self, other, level, flag=ALL):
        d = sum(self.uvw[c].dot(other.uvw[c], level, flag) for c in range(3))
        return d + (self.p.dot(other.p, level, ALL) if _pressure_flag(flag) else 0.0)

    def project_mean_pressure(self, level):
        """vertexdof::projectMean of the pressure"""
        n = self.storage.ndofs(level)
        self.p.add_const(-self.p.sum(level, ALL) / n, level, ALL)


class P2ProjectNormalOperator:
    """u -= (u.n) n with the radial normal n = x/|x| on the P2 DoFs whose boundary type is in flag"""

    def __init__(self, storage):
        self.ctx = storage.ctx

    def project(self, f, level, flag=FREESLIP):
        """f: a P2P1TaylorHoodFunction (its velocity is projected) or a list of three P2Functions"""
        comps = f.uvw if hasattr(f, "uvw") else f
        h = np.ascontiguousarray([c.h for c in comps], dtype=np.int64)
        self.ctx.check(lib().hg_p2_project_normal(self.ctx.h, _ptr(h), level, flag), "P2ProjectNormalOperator")


class P2P1StokesOperator:
    """P2P1ElementwiseBlendingFullViscousStokesOperator with viscosity
    eta(x) = exp(c0 + c1 |x|) at the physical (blended) point.  freeslip=True
    makes it the free-slip wrapper (P A P, P B^T, B P, pressure mean projected out
    of apply) and selects the projections of the free-slip solver layout."""

    A_U, BT_P, B_U = 1, 2, 4

    def __init__(self, storage, min_level, max_level, visc_c0=0.0, visc_c1=0.0, freeslip=False):
        self.storage = storage
        self.ctx = storage.ctx
        self.freeslip = bool(freeslip)
        h = C.c_int64()
        self.ctx.check(lib().hg_stokes_create(self.ctx.h, storage.h, min_level, max_level, float(visc_c0),
                                              float(visc_c1), C.byref(h)), "P2P1StokesOperator")
        self.h = h.value

    def destroy(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().hg_operator_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def apply(self, src, dst, level, flag, update=REPLACE, mask=7):
        """mask bit 1: dst.uvw = A src.uvw, 2: dst.uvw += B^T src.p, 4: dst.p = B src.uvw
        (free-slip wrapper: masks 1, 2, 4 or 7)"""
        fn = lib().hg_stokes_freeslip_apply if self.freeslip else lib().hg_stokes_apply
        self.ctx.check(fn(self.ctx.h, self.h, _ptr(src.handles()), _ptr(dst.handles()), level, flag, mask, update),
                       "Stokes apply")

    def velocity_inverse_diagonal(self, level, dst):
        self.ctx.check(lib().hg_stokes_velocity_inverse_diagonal(self.ctx.h, self.h, level, _ptr(dst.handles())),
                       "velocity inverse diagonal")

    def pspg_inverse_diagonal(self, level, dst):
        self.ctx.check(lib().hg_stokes_pressure_diagonal(self.ctx.h, self.h, level, 0, dst.h), "PSPG diagonal")

    def kmass_apply(self, src, dst, level, flag=ALL, update=REPLACE):
        """dst = M_k src on pressure functions, the P1 mass with k = 1/eta"""
        s = np.ascontiguousarray([0, 0, 0, src.h], dtype=np.int64)
        d = np.ascontiguousarray([0, 0, 0, dst.h], dtype=np.int64)
        self.ctx.check(lib().hg_stokes_apply(self.ctx.h, self.h, _ptr(s), _ptr(d), level, flag, 8, update),
                       "K-mass apply")

    def lumped_inverse_mass(self, level, dst):
        self.ctx.check(lib().hg_stokes_pressure_diagonal(self.ctx.h, self.h, level, 1, dst.h), "lumped mass")


def stokes_estimate_radius(A, x, tmp, level, iterations=25):
    """power iteration on D^-1 A of the velocity block; x.uvw is the start vector"""
    r = C.c_double()
    A.ctx.check(lib().hg_stokes_estimate_radius(A.ctx.h, A.h, _ptr(x.handles()), _ptr(tmp.handles()), level,
                                                iterations, int(A.freeslip), C.byref(r)), "Stokes estimateRadius")
    return r.value


def stokes_uzawa_smooth(A, x, b, level, rho, omega=0.3, velocity_iterations=1, order=3, project=True):
    A.ctx.check(lib().hg_stokes_uzawa_smooth(A.ctx.h, A.h, _ptr(x.handles()), _ptr(b.handles()), level, omega,
                                             velocity_iterations, order, rho, int(project), int(A.freeslip)),
                "Uzawa smoother")


def stokes_minres_solve(A, x, b, level, max_iter=10, rel_tol=1e-6, abs_tol=0.0):
    its = C.c_int()
    A.ctx.check(lib().hg_stokes_minres_solve(A.ctx.h, A.h, _ptr(x.handles()), _ptr(b.handles()), level, max_iter,
                                             rel_tol, abs_tol, int(A.freeslip), C.byref(its)), "MINRES")
    return its.value


def stokes_fgmres_solve(A, x, b, min_level, max_level, rho, a_pre=3, a_post=3, order=3, a_coarse_iter=10,
                        a_coarse_tol=1e-6, schur_iter=50, schur_tol=1e-6, max_iter=100, restart=50, tol=1e-7,
                        velocity_dtype=F64):
    """FGMRES with the block-factorisation preconditioner (velocity block: one GMG V-cycle with
    Chebyshev smoothing, in FP64 or, velocity_dtype F32 / F16, in FP32 arithmetic on FP32 / FP16
    vectors; Schur complement: CG on the 1/eta pressure mass); returns (iterations, approximate
    residual history)"""
    hist = np.zeros(max_iter + 1)
    its = C.c_int()
    A.ctx.check(lib().hg_stokes_fgmres_solve(A.ctx.h, A.h, _ptr(x.handles()), _ptr(b.handles()), min_level, max_level,
                                             a_pre, a_post, order, rho, a_coarse_iter, a_coarse_tol, schur_iter,
                                             schur_tol, int(A.freeslip), int(velocity_dtype), max_iter, restart, tol,
                                             _ptr(hist), C.byref(its)), "Stokes FGMRES")
    return its.value, hist[:its.value + 1]


def stokes_gmg_solve(A, x, b, min_level, max_level, rho, pre=3, post=3, smooth_increment=2, omega=0.3,
                     velocity_iterations=1, order=3, coarse_max_iter=10, coarse_tol=1e-6, project=True, cycles=10,
                     tol=0.0):
    """stokesGMGUzawaFSSolver layout (with its free-slip projections when A.freeslip);
    returns (cycles done, residual history)"""
    hist = np.zeros(cycles + 1)
    done = C.c_int()
    A.ctx.check(lib().hg_stokes_gmg_solve(A.ctx.h, A.h, _ptr(x.handles()), _ptr(b.handles()), min_level, max_level,
                                          pre, post, smooth_increment, omega, velocity_iterations, order, rho,
                                          coarse_max_iter, coarse_tol, int(project), int(A.freeslip), cycles, tol,
                                          _ptr(hist), C.byref(done)), "Stokes GMG")
    return done.value, hist[:done.value + 1]


def vtk_write(path, storage, level, functions, names=None, binary=False):
    """VTKOutput::write: one .vtu file with the P1 or P2 functions on the level's
    micro-tetrahedra (per-macro-cell points, blended); names default to f0, f1, ..."""
    names = names or [f"f{i}" for i in range(len(functions))]
    h = np.ascontiguousarray([f.h for f in functions], dtype=np.int64)
    enc = [n.encode() for n in names]
    arr = (C.c_char_p * len(enc))(*enc)
    storage.ctx.check(lib().hg_vtk_write(storage.ctx.h, str(path).encode(), storage.h, level, len(functions),
                                         _ptr(h), C.cast(arr, C.c_void_p), int(binary)), "vtk_write")


class P2ConstantLaplaceOperator:
    """P2ConstantLaplaceOperator: the P2 Laplacian of the affine macro-cells with constant
    stencils, Gauss-Seidel / SOR smoothable"""

    def __init__(self, storage, min_level, max_level):
        self.storage = storage
        self.ctx = storage.ctx
        h = C.c_int64()
        self.ctx.check(lib().hg_p2_const_create(self.ctx.h, storage.h, min_level, max_level, C.byref(h)),
                       "P2ConstantLaplaceOperator")
        self.h = h.value

    def destroy(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().hg_operator_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def apply(self, src, dst, level, flag=ALL, update=REPLACE):
        self.ctx.check(lib().hg_p2_const_apply(self.ctx.h, self.h, src.h, dst.h, level, flag, update), "apply")

    def smooth_sor(self, x, b, level, flag=GMG_FLAG, relax=1.0):
        self.ctx.check(lib().hg_p2_const_smooth_sor(self.ctx.h, self.h, x.h, b.h, level, flag, relax), "smooth_sor")

    def smooth_gs(self, x, b, level, flag=GMG_FLAG):
        self.smooth_sor(x, b, level, flag, 1.0)

    def smooth_sor_backwards(self, x, b, level, flag=GMG_FLAG, relax=1.0):
        self.ctx.check(lib().hg_p2_const_smooth_sor_backwards(self.ctx.h, self.h, x.h, b.h, level, flag, relax),
                       "smooth_sor_backwards")

    def sor_colours(self, level):
        """(colour of every canonical P2 DoF in sweep order, number of boundary colours)"""
        n = self.storage.p2_ndofs(level)
        col = np.zeros(n, dtype=np.int32)
        nb = C.c_int()
        self.ctx.check(lib().hg_p2_const_sor_colours(self.ctx.h, self.h, level, _ptr(col), C.byref(nb)),
                       "sor_colours")
        return col, nb.value

    def inverse_diagonal(self, level, dst):
        self.ctx.check(lib().hg_p2_const_inverse_diagonal(self.ctx.h, self.h, level, dst.h), "inverse diagonal")


def p2_const_gmg_solve(A, x, b, min_level, max_level, pre=2, post=2, coarse_iter=1000, coarse_tol=1e-14,
                       cycles=10, tol=0.0):
    """GeometricMultigridSolver<P2ConstantLaplaceOperator> with Gauss-Seidel smoothing and CG on
    min_level; returns (cycles done, Inner residual history)"""
    hist = np.zeros(cycles + 1)
    done = C.c_int()
    A.ctx.check(lib().hg_p2_const_gmg_solve(A.ctx.h, A.h, x.h, b.h, min_level, max_level, pre, post, coarse_iter,
                                            coarse_tol, cycles, tol, _ptr(hist), C.byref(done)), "P2 constant GMG")
    return done.value, hist[:done.value + 1]
