"""ctypes binding of the CPU oracle (oracle/libhg_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker / CPU baseline, never as
the product path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

INNER, DIRICHLET, NEUMANN, FREESLIP, ALL = 1, 2, 4, 8, 15
GMG_FLAG = INNER | NEUMANN | FREESLIP

_d = C.c_double
_i = C.c_int
_i64 = C.c_int64
_p = C.c_void_p


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libhg_oracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        sig = {
            "or_mesh_create": (_p, [_i, _p, _i, _p]),
            "or_mesh_cuboid": (_p, [_d, _d, _d, _d, _d, _d, _i, _i, _i]),
            "or_mesh_free": (None, [_p]),
            "or_mesh_counts": (None, [_p, _p]),
            "or_mesh_cell_vertices": (None, [_p, _p]),
            "or_p1_ndofs": (_i64, [_p, _i]),
            "or_p1_nbnd": (_i64, [_p, _i]),
            "or_p1_coords": (None, [_p, _i, _p]),
            "or_p1_flags": (None, [_p, _i, _p]),
            "or_p1_locate": (_i64, [_p, _i, _i, _i, _i, _i]),
            "or_p1op_create": (_p, [_p, _i]),
            "or_p1op_free": (None, [_p]),
            "or_p1op_nnz": (_i64, [_p]),
            "or_p1op_csr": (None, [_p, _p, _p, _p]),
            "or_p1op_cell_stencil": (None, [_p, _i, _p]),
            "or_p1_apply": (None, [_p, _p, _p, _i, _i]),
            "or_p1_sor": (None, [_p, _p, _p, _i, _d]),
            "or_p1_restrict": (None, [_p, _i, _p, _p, _i]),
            "or_p1_prolongate": (None, [_p, _i, _p, _p, _i, _i]),
            "or_dot": (_d, [_i64, _p, _p, _p, _i]),
            "or_p1_cg": (_i, [_p, _p, _p, _i, _i, _d, _d]),
            "or_p1_gmg": (_i, [_p, _i, _i, _p, _p, _i, _i, _i, _d, _p, _i, _d]),
            "or_p1_owner_vertices": (_i, [_p, _i, _i64, _p]),
            "or_sor_mc_csr": (None, [_i64, _p, _p, _p, _p, _i, _p, _p, _p, _d, _i]),
            "or_p1_gmg_many": (None, [_p, _i, _i, _i, _p, _p, _i, _i, _i, _d, _i, _d, _p]),
            "or_mesh_shell": (_p, [_i, _i, _d, _d]),
            "or_mesh_blend": (None, [_p, _p]),
            "or_shell_map": (None, [_p, _p, _p, _p]),
            "or_quadrature": (None, [_p, _p]),
            "or_p2_element_matrix": (None, [_p, _p, _i, _p, _p, _p, _p]),
            "or_p2_reduced_element_matrix": (None, [_p, _p, _i, _p, _p, _p, _p]),
            "or_p2op_set_reduced": (None, [_p, _i]),
            "or_p2op_create": (_p, [_p, _i, _i, _p, _p, _p]),
            "or_p2op_free": (None, [_p]),
            "or_p2_apply": (None, [_p, _p, _p, _i, _i]),
            "or_p2_inv_diag": (None, [_p, _p]),
            "or_p2op_nelem": (_i64, [_p]),
            "or_p2op_coo": (None, [_p, _p, _p, _p]),
            "or_p2_prolongate": (None, [_p, _i, _p, _p, _i, _i]),
            "or_p2_restrict": (None, [_p, _i, _p, _p, _i]),
            "or_chebyshev_coefficients": (_i, [_i, _d, _d, _p]),
            "or_p2_chebyshev_smooth": (None, [_p, _p, _i, _p, _p, _p, _i]),
            "or_p2_power_iteration": (_d, [_p, _p, _p, _i]),
            "or_p2_cg": (_i, [_p, _p, _p, _i, _i, _d, _d, _i]),
            "or_p2_gmg": (_i, [_p, _i, _i, _i, _p, _p, _p, _p, _p, _i, _i, _i, _d, _i, _d, _p, _i, _d, _i]),
            "or_stokes_element": (None, [_p, _p, _d, _d, _p, _p]),
            "or_stokesop_create": (_p, [_p, _i, _d, _d]),
            "or_stokesop_free": (None, [_p]),
            "or_stokes_apply": (None, [_p, _p, _p, _p, _p, _i, _i, _i]),
            "or_stokes_inv_diag": (None, [_p, _p]),
            "or_p1_pspg_inv_diag": (None, [_p, _i, _p]),
            "or_p1_lumped_inv_mass": (None, [_p, _i, _p]),
            "or_mesh_set_shell_flags": (None, [_p, _d, _i, _i]),
            "or_p2_project_normal": (None, [_p, _i, _p, _i]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Mesh:
    def __init__(self, handle):
        self.h = handle

    @staticmethod
    def cuboid(lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(1, 1, 1)):
        return Mesh(lib().or_mesh_cuboid(*lo, *hi, *n))

    @staticmethod
    def shell(ntan=5, nrad=3, rmin=0.55, rmax=1.0, inner_flag=None, outer_flag=None):
        m = Mesh(lib().or_mesh_shell(ntan, nrad, rmin, rmax))
        if inner_flag is not None or outer_flag is not None:
            lib().or_mesh_set_shell_flags(m.h, 0.5 * (rmin + rmax), inner_flag or DIRICHLET, outer_flag or DIRICHLET)
        return m

    def p2_project_normal(self, level, u, flag=FREESLIP):
        """u -= (u.n) n with n = x/|x| on P2 DoFs of type flag; u = [ux|uy|uz] (in place)"""
        assert u.flags["C_CONTIGUOUS"] and u.dtype == np.float64
        lib().or_p2_project_normal(self.h, level, _ptr(u), flag)

    def blend(self):
        out = np.zeros((self.counts()[3], 14))
        lib().or_mesh_blend(self.h, _ptr(out))
        return out

    @staticmethod
    def from_arrays(coords, cells):
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        cells = np.ascontiguousarray(cells, dtype=np.int64)
        return Mesh(lib().or_mesh_create(len(coords), _ptr(coords), len(cells), _ptr(cells)))

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.or_mesh_free(self.h)
            self.h = None

    def counts(self):
        out = np.zeros(4, dtype=np.int64)
        lib().or_mesh_counts(self.h, _ptr(out))
        return tuple(int(v) for v in out)

    def cell_vertices(self):
        nc = self.counts()[3]
        out = np.zeros((nc, 4), dtype=np.int64)
        lib().or_mesh_cell_vertices(self.h, _ptr(out))
        return out

    def ndofs(self, level):
        return int(lib().or_p1_ndofs(self.h, level))

    def nbnd(self, level):
        return int(lib().or_p1_nbnd(self.h, level))

    def coords(self, level):
        out = np.zeros((self.ndofs(level), 3))
        lib().or_p1_coords(self.h, level, _ptr(out))
        return out

    def flags(self, level):
        out = np.zeros(self.ndofs(level), dtype=np.int32)
        lib().or_p1_flags(self.h, level, _ptr(out))
        return out

    def locate(self, level, c, x, y, z):
        return int(lib().or_p1_locate(self.h, level, c, x, y, z))

    # P2 functions of P2 level L live on the P1 DoFs of level L+1
    def p2_prolongate(self, coarse_level, coarse, fine, flag, add=False):
        lib().or_p2_prolongate(self.h, coarse_level, _ptr(coarse), _ptr(fine), flag, int(add))

    def p2_restrict(self, fine_level, fine, coarse, flag):
        lib().or_p2_restrict(self.h, fine_level, _ptr(fine), _ptr(coarse), flag)

    def p2_gmg(self, min_level, max_level, k, x, b, pre, post, order, rho, cycles=50, tol=1e-8,
               coarse_max_iter=10000, coarse_tol=1e-14, coarse_jacobi=True):
        hist = np.zeros(cycles + 1)
        kk = lib().or_p2_gmg(self.h, min_level, max_level, *(k or KCoeff()).args(), _ptr(x), _ptr(b), pre, post,
                             order, rho, cycles, tol, _ptr(hist), coarse_max_iter, coarse_tol, int(coarse_jacobi))
        if kk < 0:
            raise ValueError("unsupported Chebyshev order")
        return kk, hist[:kk + 1]

    def restrict(self, fine_level, fine, coarse, flag):
        lib().or_p1_restrict(self.h, fine_level, _ptr(fine), _ptr(coarse), flag)

    def prolongate(self, coarse_level, coarse, fine, flag, add):
        lib().or_p1_prolongate(self.h, coarse_level, _ptr(coarse), _ptr(fine), flag, int(add))

    def gmg(self, min_level, max_level, x, b, pre=3, post=3, cycles=10, tol=0.0, coarse_maxit=10000,
            coarse_tol=1e-16):
        hist = np.zeros(cycles + 1)
        k = lib().or_p1_gmg(self.h, min_level, max_level, _ptr(x), _ptr(b), pre, post, cycles, tol,
                            _ptr(hist), coarse_maxit, coarse_tol)
        return k, hist[: k + 1]

    def gmg_many(self, min_level, max_level, X, B, pre=3, post=3, cycles=50, tol=0.0, coarse_maxit=10000,
                 coarse_tol=1e-16):
        """or_p1_gmg for every row of B (shape [nvec, ndofs]) into X, OpenMP over the rows"""
        X = np.ascontiguousarray(X)
        B = np.ascontiguousarray(B)
        done = np.zeros(B.shape[0], dtype=np.int32)
        lib().or_p1_gmg_many(self.h, min_level, max_level, B.shape[0], _ptr(X), _ptr(B), pre, post, cycles, tol,
                             coarse_maxit, coarse_tol, _ptr(done))
        return X, done

    def owner_vertices(self, level, dof):
        vs = np.zeros(4, dtype=np.int32)
        k = lib().or_p1_owner_vertices(self.h, level, int(dof), _ptr(vs))
        return tuple(int(v) for v in vs[:k])


def sor_multicolour(A, colour, mask, x, b, relax=1.0, backward=False):
    """one SOR sweep over the CSR matrix A, rows visited colour by colour
    (colour 0 first, or the last colour first when backward), rows with
    mask == 0 untouched; x is updated in place"""
    A = A.tocsr()
    colour = np.ascontiguousarray(colour, dtype=np.int32)
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    rp = np.ascontiguousarray(A.indptr, dtype=np.int64)
    ci = np.ascontiguousarray(A.indices, dtype=np.int64)
    va = np.ascontiguousarray(A.data, dtype=np.float64)
    bb = np.ascontiguousarray(b, dtype=np.float64)
    assert x.dtype == np.float64 and x.flags.c_contiguous
    ncol = int(colour.max()) + 1 if len(colour) else 0
    lib().or_sor_mc_csr(A.shape[0], _ptr(rp), _ptr(ci), _ptr(va), _ptr(colour), ncol, _ptr(m), _ptr(x), _ptr(bb),
                        relax, int(backward))
    return x


def colour_conflicts(A, colour, mask=None):
    """pairs (p, q), p != q, coupled by a nonzero of A (in either direction)
    with colour[p] == colour[q] among the rows with mask != 0: a multicolour
    sweep that updates one colour's rows simultaneously equals the sequential
    sweep only when this is empty"""
    A = A.tocoo()
    colour = np.asarray(colour)
    keep = (A.row != A.col) & (A.data != 0) & (colour[A.row] == colour[A.col])
    if mask is not None:
        mask = np.asarray(mask).astype(bool)
        keep &= mask[A.row] & mask[A.col]
    pairs = {(int(min(p, q)), int(max(p, q))) for p, q in zip(A.row[keep], A.col[keep])}
    return sorted(pairs)


class P1Op:
    def __init__(self, mesh, level):
        self.mesh = mesh
        self.level = level
        self.h = lib().or_p1op_create(mesh.h, level)
        self.n = mesh.ndofs(level)

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.or_p1op_free(self.h)
            self.h = None

    def apply(self, src, dst, flag, add=False):
        lib().or_p1_apply(self.h, _ptr(src), _ptr(dst), flag, int(add))

    def sor(self, x, b, flag, relax=1.0):
        lib().or_p1_sor(self.h, _ptr(x), _ptr(b), flag, relax)

    def cg(self, x, b, flag, maxit, rel, abs_):
        return lib().or_p1_cg(self.h, _ptr(x), _ptr(b), flag, maxit, rel, abs_)

    def cell_stencil(self, c):
        out = np.zeros(15)
        lib().or_p1op_cell_stencil(self.h, c, _ptr(out))
        return out

    def csr(self):
        import scipy.sparse as sp
        nnz = int(lib().or_p1op_nnz(self.h))
        rp = np.zeros(self.n + 1, dtype=np.int64)
        cols = np.zeros(nnz, dtype=np.int64)
        vals = np.zeros(nnz)
        lib().or_p1op_csr(self.h, _ptr(rp), _ptr(cols), _ptr(vals))
        return sp.csr_matrix((vals, cols, rp), shape=(self.n, self.n))


class KCoeff:
    """k(x) = exp(sum_m a_m sin(omega_m . x + phi_m)); empty means k = 1"""

    def __init__(self, a=(), omega=(), phi=()):
        self.a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        self.omega = np.ascontiguousarray(omega, dtype=np.float64).reshape(-1, 3)
        self.phi = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1)
        self.nm = len(self.a)

    def __call__(self, x):
        x = np.atleast_2d(x)
        if self.nm == 0:
            return np.ones(len(x))
        return np.exp((self.a[None, :] * np.sin(x @ self.omega.T + self.phi[None, :])).sum(axis=1))

    def args(self):
        e = np.zeros(1)
        if self.nm == 0:
            return 0, _ptr(e), _ptr(e), _ptr(e)
        return self.nm, _ptr(self.a), _ptr(self.omega), _ptr(self.phi)


def shell_map(blend14, x):
    blend14 = np.ascontiguousarray(blend14, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    F = np.zeros(3)
    DF = np.zeros(9)
    lib().or_shell_map(_ptr(blend14), _ptr(x), _ptr(F), _ptr(DF))
    return F, DF.reshape(3, 3)


def quadrature():
    pts = np.zeros((11, 4))
    w = np.zeros(11)
    lib().or_quadrature(_ptr(pts), _ptr(w))
    return pts, w


def p2_element_matrix(p, blend=None, k=None, reduced=False):
    """10x10 element matrix of the reference form, or (reduced) of the reduced form
    of the low-precision applies"""
    p = np.ascontiguousarray(p, dtype=np.float64)
    k = k or KCoeff()
    A = np.zeros((10, 10))
    b = None if blend is None else np.ascontiguousarray(blend, dtype=np.float64)
    f = lib().or_p2_reduced_element_matrix if reduced else lib().or_p2_element_matrix
    f(_ptr(p), None if b is None else _ptr(b), *k.args(), _ptr(A))
    return A


class P2Op:
    def __init__(self, mesh, level, k=None, reduced=False):
        """reduced: element matrices of the reduced form of the low-precision applies"""
        self.mesh = mesh
        self.level = level
        self.k = k or KCoeff()
        self.h = lib().or_p2op_create(mesh.h, level, *self.k.args())
        if reduced:
            lib().or_p2op_set_reduced(self.h, 1)
        self.n = mesh.ndofs(level + 1)

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.or_p2op_free(self.h)
            self.h = None

    def apply(self, src, dst, flag, add=False):
        lib().or_p2_apply(self.h, _ptr(src), _ptr(dst), flag, int(add))

    def inv_diag(self):
        out = np.zeros(self.n)
        lib().or_p2_inv_diag(self.h, _ptr(out))
        return out

    def chebyshev_smooth(self, inv_diag, coeffs, x, b, flag=GMG_FLAG):
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
        lib().or_p2_chebyshev_smooth(self.h, _ptr(inv_diag), len(coeffs), _ptr(coeffs), _ptr(x), _ptr(b), flag)

    def power_iteration(self, inv_diag, x, its):
        return float(lib().or_p2_power_iteration(self.h, _ptr(inv_diag), _ptr(x), its))

    def cg(self, x, b, flag=GMG_FLAG, max_iter=10000, rel_tol=1e-14, abs_tol=1e-300, jacobi=True):
        return int(lib().or_p2_cg(self.h, _ptr(x), _ptr(b), flag, max_iter, rel_tol, abs_tol, int(jacobi)))

    def matrix(self):
        import scipy.sparse as sp
        ne = int(lib().or_p2op_nelem(self.h))
        r = np.zeros(100 * ne, dtype=np.int64)
        c = np.zeros(100 * ne, dtype=np.int64)
        v = np.zeros(100 * ne)
        lib().or_p2op_coo(self.h, _ptr(r), _ptr(c), _ptr(v))
        A = sp.csr_matrix((v, (r, c)), shape=(self.n, self.n))
        A.sum_duplicates()
        return A


def chebyshev_coefficients(order, lower, upper):
    out = np.zeros(order)
    if lib().or_chebyshev_coefficients(order, lower, upper, _ptr(out)):
        raise ValueError("unsupported Chebyshev order")
    return out


def dot(a, b, flags, flag):
    return float(lib().or_dot(len(a), _ptr(a), _ptr(b), _ptr(flags), flag))


# ---- P2-P1 Stokes ------------------------------------------------------------------
def stokes_element(p, blend=None, c0=0.0, c1=0.0):
    """(A 30x30, B 4x30) of the blended full-viscous Stokes element, eta = exp(c0 + c1 |F(x)|)"""
    p = np.ascontiguousarray(p, dtype=np.float64)
    A = np.zeros((30, 30))
    B = np.zeros((4, 30))
    b = None if blend is None else np.ascontiguousarray(blend, dtype=np.float64)
    lib().or_stokes_element(_ptr(p), None if b is None else _ptr(b), c0, c1, _ptr(A), _ptr(B))
    return A, B


def p1_pspg_inv_diag(mesh, level):
    out = np.zeros(mesh.ndofs(level))
    lib().or_p1_pspg_inv_diag(mesh.h, level, _ptr(out))
    return out


def p1_lumped_inv_mass(mesh, level):
    out = np.zeros(mesh.ndofs(level))
    lib().or_p1_lumped_inv_mass(mesh.h, level, _ptr(out))
    return out


class StokesOp:
    """P2P1ElementwiseBlendingFullViscousStokesOperator on P2 level `level`: velocity
    [ux|uy|uz] in the P1 numbering of level+1, pressure in that of level (all Inner)"""

    A_U, BT_P, B_U, K_MASS = 1, 2, 4, 8

    def __init__(self, mesh, level, c0=0.0, c1=0.0, freeslip=False):
        self.mesh = mesh
        self.level = level
        self.freeslip = freeslip
        self.h = lib().or_stokesop_create(mesh.h, level, c0, c1)
        self.n2 = mesh.ndofs(level + 1)
        self.n1 = mesh.ndofs(level)
        self.fl2 = np.tile(mesh.flags(level + 1), 3)

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.or_stokesop_free(self.h)
            self.h = None

    def apply(self, u, p, yu, yp, flag, mask=7, add=False):
        if self.freeslip:
            self._apply_freeslip(u, p, yu, yp, flag, mask, add)
            return
        self.apply_raw(u, p, yu, yp, flag, mask, add)

    def apply_raw(self, u, p, yu, yp, flag, mask=7, add=False):
        lib().or_stokes_apply(self.h, None if u is None else _ptr(u), None if p is None else _ptr(p),
                              None if yu is None else _ptr(yu), None if yp is None else _ptr(yp), flag, mask, int(add))

    def project(self, u, flag=FREESLIP):
        """P2ProjectNormalOperator::project on [ux|uy|uz] (in place)"""
        self.mesh.p2_project_normal(self.level, u, flag)

    def _apply_freeslip(self, u, p, yu, yp, flag, mask, add):
        """P2P1FullStokesFreeslipWrapperTemplate (terraneo/operators/P2P1StokesOperatorWithWrapper.hpp:41-386):
        mask 1: P A P u, 2: P B^T p, 4: B P u, 7: the blocks on (P u, p), P on the velocity
        result, pressure mean of the result projected out; results written on flag"""
        assert mask in (1, 2, 4, 7)
        tu = None
        if mask & 5:
            tu = u.copy()
            self.project(tu)
        if mask == 4:
            self.apply_raw(tu, None, None, yp, flag, 4, add)
            return
        du = np.zeros(3 * self.n2)
        dp = np.zeros(self.n1) if mask == 7 else None
        self.apply_raw(tu, p if mask & 2 else None, du, dp, flag, mask, False)
        self.project(du)
        m = self.vmask(flag)
        yu[m] = yu[m] + du[m] if add else du[m]
        if mask == 7:
            if flag & INNER:
                yp[:] = yp + dp if add else dp
            project_mean(yp)

    def inv_diag(self):
        out = np.zeros(3 * self.n2)
        lib().or_stokes_inv_diag(self.h, _ptr(out))
        return out

    def vmask(self, flag):
        return (self.fl2 & flag) != 0

    @staticmethod
    def pmask(p, flag):
        return np.full(len(p), bool(flag & INNER))


def project_mean(p):
    """vertexdof::projectMean: subtract the arithmetic mean over all DoFs"""
    p -= p.sum() / len(p)


def stokes_chebyshev(op, inv_diag, coeffs, u, bu, flag=GMG_FLAG, project=None):
    """ChebyshevSmoother::solve on the velocity operator (order = len(coeffs)); the
    corrections are projected when `project` (default: the operator is free-slip)"""
    if project is None:
        project = op.freeslip
    m = op.vmask(flag)
    t2 = np.zeros_like(u)
    op.apply(u, None, t2, None, flag, StokesOp.A_U)
    t2[m] = bu[m] - t2[m]
    t1 = np.zeros_like(u)
    t1[m] = inv_diag[m] * t2[m]
    if project:  # ChebyshevSmoother<., P2ProjectNormalOperator> (ChebyshevSmoother.hpp:186-231)
        op.project(t1)
    u[m] += coeffs[0] * t1[m]
    for k in range(1, len(coeffs)):
        op.apply(t1, None, t2, None, flag, StokesOp.A_U)
        t1[:] = 0.0
        t1[m] = inv_diag[m] * t2[m]
        if project:
            op.project(t1)
        u[m] += coeffs[k] * t1[m]


def stokes_power_iteration(op, inv_diag, x, its):
    """chebyshev::estimateRadius on the velocity operator: power iteration on D^-1 A (flag All)"""
    x /= np.sqrt(x @ x)
    aux = np.zeros_like(x)
    op.apply(x, None, aux, None, ALL, StokesOp.A_U)
    aux *= inv_diag
    radius = 0.0
    for _ in range(its):
        x[:] = aux / np.sqrt(aux @ aux)
        op.apply(x, None, aux, None, ALL, StokesOp.A_U)
        aux *= inv_diag
        radius = float(x @ aux)
    return radius


def uzawa_smooth(op, u, p, bu, bp, inv_diag, coeffs, pspg_inv, omega, vel_iter=1, flag=GMG_FLAG, project=True):
    """UzawaSmoother (block variant without stabilisation, UzawaSmoother.hpp:330-470) with a
    Chebyshev velocity smoother"""
    m = op.vmask(flag)
    ru = np.zeros_like(u)
    op.apply(None, p, ru, None, flag, StokesOp.BT_P)
    ru[m] = bu[m] - ru[m]
    if op.freeslip:  # UzawaSmootherWithFreeSlipProjection (UzawaSmoother.hpp:549-611)
        op.project(ru)
    for _ in range(vel_iter):
        stokes_chebyshev(op, inv_diag, coeffs, u, ru, flag)
    if not flag & INNER:
        return
    rp = np.zeros_like(p)
    op.apply(u, None, None, rp, flag, StokesOp.B_U)
    rp = omega * (bp - rp)
    if project:
        project_mean(rp)
    p += pspg_inv * rp
    if project:
        project_mean(p)


def stokes_residual(op, u, p, bu, bp, flag):
    ru = np.zeros_like(u)
    rp = np.zeros_like(p)
    op.apply(u, p, ru, rp, flag, 7)
    m = op.vmask(flag)
    ru[~m] = 0.0
    ru[m] = bu[m] - ru[m]
    rp = bp - rp if flag & INNER else np.zeros_like(p)
    return ru, rp



def stokes_minres(op, u, p, bu, bp, minv, max_iter, rel_tol, abs_tol=0.0, flag=GMG_FLAG):
    """MinResSolver (MinresSolver.hpp:117-245) with StokesPressureBlockPreconditioner<P1LumpedInvMass>;
    returns the iteration index at exit"""
    n2 = len(u)
    m = np.concatenate([op.vmask(flag), np.full(len(p), bool(flag & INNER))])
    A = lambda z: np.concatenate(stokes_apply_all(op, z[:n2], z[n2:], flag))  # noqa: E731

    def prec(v):
        z = np.zeros_like(v)
        z[m] = v[m]
        if flag & INNER:
            z[n2:] = minv * v[n2:]
        return z

    x = np.concatenate([u, p])
    b = np.concatenate([bu, bp])
    v = np.zeros_like(x)
    v[m] = b[m] - A(x)[m]
    vm = np.zeros_like(x)
    wm = np.zeros_like(x)
    w = np.zeros_like(x)
    z = prec(v)
    gamma_old, gamma_new = 1.0, np.sqrt(z[m] @ v[m])
    res_start = eta = gamma_new
    s_old = s_new = 0.0
    c_old = c_new = 1.0
    it = max_iter
    if gamma_new < abs_tol:
        return 0
    for i in range(max_iter):
        z[m] /= gamma_new
        vp = np.zeros_like(x)
        vp[m] = A(z)[m]
        delta = vp[m] @ z[m]
        vp[m] = vp[m] - (delta / gamma_new) * v[m] - (gamma_new / gamma_old) * vm[m]
        zp = prec(vp)
        gamma_old, gamma_new = gamma_new, np.sqrt(zp[m] @ vp[m])
        a0 = c_new * delta - c_old * s_new * gamma_old
        a1 = np.sqrt(a0 * a0 + gamma_new * gamma_new)
        a2 = s_new * delta + c_old * c_new * gamma_old
        a3 = s_old * gamma_old
        c_old, c_new = c_new, a0 / a1
        s_old, s_new = s_new, gamma_new / a1
        wp = np.zeros_like(x)
        wp[m] = z[m] / a1 - (a3 / a1) * wm[m] - (a2 / a1) * w[m]
        x[m] += c_new * eta * wp[m]
        eta = -s_new * eta
        vm, v = v, vp
        wm, w = w, wp
        z = zp
        if abs(eta) / res_start < rel_tol or abs(eta) < abs_tol:
            it = i
            break
    u[:] = x[:n2]
    p[:] = x[n2:]
    return it


def stokes_apply_all(op, u, p, flag):
    yu = np.zeros_like(u)
    yp = np.zeros_like(p)
    op.apply(u, p, yu, yp, flag, 7)
    return yu, yp


class StokesGMG:
    """GeometricMultigridSolver<P2P1 Stokes> with the Uzawa smoother (Chebyshev velocity
    smoother, PSPG inverse diagonal), P2P1StokesToP2P1Stokes restriction / prolongation and
    pressure-preconditioned MINRES on the coarsest level (stokesGMGUzawaFSSolver's layout,
    StokesFSGMGUzawaSolverTemplate.hpp:63-190, without free-slip projections)"""

    def __init__(self, mesh, min_level, max_level, c0=0.0, c1=0.0, order=3, rho=None, omega=0.3, pre=3, post=3,
                 inc=2, vel_iter=1, coarse_iter=10, coarse_tol=1e-6, power_its=25, seed=42, project=True,
                 freeslip=False):
        self.m, self.minL, self.maxL = mesh, min_level, max_level
        self.pre, self.post, self.inc, self.omega, self.vel_iter = pre, post, inc, omega, vel_iter
        self.coarse_iter, self.coarse_tol, self.project = coarse_iter, coarse_tol, project
        self.freeslip = freeslip
        self.ops = {L: StokesOp(mesh, L, c0, c1, freeslip) for L in range(min_level, max_level + 1)}
        self.invd = {L: self.ops[L].inv_diag() for L in self.ops}
        self.pspg = {L: p1_pspg_inv_diag(mesh, L) for L in self.ops}
        self.minv = p1_lumped_inv_mass(mesh, min_level)
        if rho is None:
            f = self.ops[max_level]
            x = np.random.default_rng(seed).uniform(-1, 1, 3 * f.n2)
            rho = stokes_power_iteration(f, self.invd[max_level], x, power_its)
        self.rho = rho
        self.coeffs = chebyshev_coefficients(order, 0.3 * rho, 1.2 * rho)

    def smooth(self, L, u, p, bu, bp):
        uzawa_smooth(self.ops[L], u, p, bu, bp, self.invd[L], self.coeffs, self.pspg[L], self.omega, self.vel_iter,
                     GMG_FLAG, self.project)

    def cycle(self, L, u, p, bu, bp):
        op = self.ops[L]
        if L == self.minL:
            stokes_minres(op, u, p, bu, bp, self.minv, self.coarse_iter, self.coarse_tol, 0.0, GMG_FLAG)
            return
        for _ in range(self.pre + self.inc * (self.maxL - L)):
            self.smooth(L, u, p, bu, bp)
        ru, rp = stokes_residual(op, u, p, bu, bp, GMG_FLAG)
        c = self.ops[L - 1]
        cu = np.zeros(3 * c.n2)
        for k in range(3):
            coarse = np.zeros(c.n2)
            self.m.p2_restrict(L, np.ascontiguousarray(ru[k * op.n2:(k + 1) * op.n2]), coarse, GMG_FLAG)
            cu[k * c.n2:(k + 1) * c.n2] = coarse
        cp = np.zeros(c.n1)
        self.m.restrict(L, rp, cp, ALL)
        if self.freeslip:  # P2P1StokesToP2P1StokesRestrictionWithFreeSlipProjection
            c.project(cu)
            project_mean(cp)
        xu, xp = np.zeros_like(cu), np.zeros_like(cp)
        self.cycle(L - 1, xu, xp, cu, cp)
        if self.freeslip:
            # P2P1StokesToP2P1StokesProlongationWithFreeSlipProjection::prolongateAndAdd:
            # prolongate a copy, project it, add it on the flag
            tu, tp = np.zeros_like(u), np.zeros_like(p)
            for k in range(3):
                fine = np.zeros(op.n2)
                self.m.p2_prolongate(L - 1, np.ascontiguousarray(xu[k * c.n2:(k + 1) * c.n2]), fine, GMG_FLAG)
                tu[k * op.n2:(k + 1) * op.n2] = fine
            self.m.prolongate(L - 1, xp, tp, ALL, False)
            op.project(tu)
            project_mean(tp)
            m = op.vmask(GMG_FLAG)
            u[m] += tu[m]
            p += tp
        else:
            for k in range(3):
                fine = np.ascontiguousarray(u[k * op.n2:(k + 1) * op.n2])
                self.m.p2_prolongate(L - 1, np.ascontiguousarray(xu[k * c.n2:(k + 1) * c.n2]), fine, GMG_FLAG, True)
                u[k * op.n2:(k + 1) * op.n2] = fine
            self.m.prolongate(L - 1, xp, p, ALL, True)
        for _ in range(self.post + self.inc * (self.maxL - L)):
            self.smooth(L, u, p, bu, bp)

    def residual_norm(self, u, p, bu, bp):
        ru, rp = stokes_residual(self.ops[self.maxL], u, p, bu, bp, INNER | NEUMANN | FREESLIP)
        return float(np.sqrt(ru @ ru + rp @ rp))

    def solve(self, u, p, bu, bp, cycles=20, tol=1e-7):
        hist = [self.residual_norm(u, p, bu, bp)]
        k = 0
        while k < cycles and hist[-1] > tol * hist[0]:
            self.cycle(self.maxL, u, p, bu, bp)
            if self.project:
                project_mean(p)
            hist.append(self.residual_norm(u, p, bu, bp))
            k += 1
        return k, np.array(hist)


# ---- FGMRES with the block-factorisation preconditioner ---------------------------
def cg_solve(apply, prec, x, b, m, max_iter, rel_tol, abs_tol=0.0):
    """CGSolver (CGSolver.hpp:94-203) on the DoFs of mask m; prec(r) -> z; returns the
    iteration index at exit"""
    if max_iter == 0:
        return 0
    r = np.zeros_like(x)
    r[m] = b[m] - apply(x)[m]
    z = np.zeros_like(x)
    z[m] = prec(r)[m]
    p = z.copy()
    prsold = r[m] @ z[m]
    res_start = np.sqrt(r[m] @ r[m])
    it = 0
    if res_start < abs_tol:
        return 0
    for i in range(max_iter):
        ap = apply(p)
        alpha = prsold / (p[m] @ ap[m])
        x[m] += alpha * p[m]
        r[m] -= alpha * ap[m]
        rsnew = r[m] @ r[m]
        it = i
        if np.sqrt(rsnew) / res_start < rel_tol or np.sqrt(rsnew) < abs_tol:
            break
        z[:] = 0.0
        z[m] = prec(r)[m]
        prsnew = r[m] @ z[m]
        p[m] = z[m] + (prsnew / prsold) * p[m]
        prsold = prsnew
    return it


def minres_solve(apply, prec, x, b, m, max_iter, rel_tol, abs_tol=0.0):
    """MinResSolver (MinresSolver.hpp:117-245) on the DoFs of mask m; prec(v) -> z"""
    def pre(v):
        z = np.zeros_like(v)
        z[m] = prec(v)[m]
        return z

    v = np.zeros_like(x)
    v[m] = b[m] - apply(x)[m]
    vm, wm, w = np.zeros_like(x), np.zeros_like(x), np.zeros_like(x)
    z = pre(v)
    gamma_old, gamma_new = 1.0, np.sqrt(z[m] @ v[m])
    res_start = eta = gamma_new
    s_old = s_new = 0.0
    c_old = c_new = 1.0
    it = max_iter
    if gamma_new < abs_tol:
        return 0
    for i in range(max_iter):
        z[m] /= gamma_new
        vp = np.zeros_like(x)
        vp[m] = apply(z)[m]
        delta = vp[m] @ z[m]
        vp[m] = vp[m] - (delta / gamma_new) * v[m] - (gamma_new / gamma_old) * vm[m]
        zp = pre(vp)
        gamma_old, gamma_new = gamma_new, np.sqrt(zp[m] @ vp[m])
        a0 = c_new * delta - c_old * s_new * gamma_old
        a1 = np.sqrt(a0 * a0 + gamma_new * gamma_new)
        a2 = s_new * delta + c_old * c_new * gamma_old
        a3 = s_old * gamma_old
        c_old, c_new = c_new, a0 / a1
        s_old, s_new = s_new, gamma_new / a1
        wp = np.zeros_like(x)
        wp[m] = z[m] / a1 - (a3 / a1) * wm[m] - (a2 / a1) * w[m]
        x[m] += c_new * eta * wp[m]
        eta = -s_new * eta
        vm, v = v, vp
        wm, w = w, wp
        z = zp
        if abs(eta) / res_start < rel_tol or abs(eta) < abs_tol:
            it = i
            break
    return it


def fgmres(apply, prec, x, b, m, max_krylov, rel_tol, abs_tol=0.0, restart=1000, arnoldi_tol=1e-16, hist=None):
    """FGMRESSolver (FGMRESSolver.hpp:55-500): right-preconditioned flexible GMRES with
    modified Gram-Schmidt and Givens rotations; hist collects the approximate
    residuals; returns the iteration count"""
    state = {}

    def init(reset):
        r0 = np.zeros_like(x)
        r0[m] = b[m] - apply(x)[m]
        beta = np.sqrt(r0[m] @ r0[m])
        if not reset:
            state["init"] = beta
            if hist is not None:
                hist.append(beta)
        v0 = np.zeros_like(x)
        v0[m] = r0[m] / beta
        state.update(V=[v0], Z=[], H=[], rot=[], g=[beta], y=[])

    def approximation():
        for yi, zi in zip(state["y"], state["Z"]):
            x[m] += yi * zi[m]
        state["y"] = []

    init(False)
    its = max_krylov - 1
    for j in range(1, max_krylov):
        if j % restart == 0:
            approximation()
            init(True)
            continue
        k = j % restart
        V, Z, H, rot, g = state["V"], state["Z"], state["H"], state["rot"], state["g"]
        z = np.zeros_like(x)
        z[m] = prec(V[k - 1])[m]
        Z.append(z)
        w = np.zeros_like(x)
        w[m] = apply(z)[m]
        col = np.zeros(k + 1)
        for i in range(k):
            col[i] = w[m] @ V[i][m]
            w[m] -= col[i] * V[i][m]
        wn = np.sqrt(w[m] @ w[m])
        col[k] = wn
        for i, (c, s_) in enumerate(rot):
            a, bb = col[i], col[i + 1]
            col[i], col[i + 1] = c * a + s_ * bb, -s_ * a + c * bb
        c, s_ = 1.0, 0.0
        if col[k] != 0.0:
            r = np.sqrt(col[k - 1] ** 2 + col[k] ** 2)
            s_, c = col[k] / r, col[k - 1] / r
        rot.append((c, s_))
        a, bb = col[k - 1], col[k]
        col[k - 1], col[k] = c * a + s_ * bb, -s_ * a + c * bb
        g.append(-s_ * g[k - 1])
        g[k - 1] = c * g[k - 1]
        H.append(col)
        y = np.zeros(k)
        for i in range(k - 1, -1, -1):
            t = g[i] - sum(H[mm][i] * y[mm] for mm in range(i + 1, k))
            y[i] = t / H[i][i]
        state["y"] = list(y)
        residual = abs(g[k])
        if hist is not None:
            hist.append(residual)
        v = np.zeros_like(x)
        v[m] = w[m] / wn
        V.append(v)
        if wn <= arnoldi_tol or residual <= abs_tol or residual / state["init"] <= rel_tol:
            its = j
            break
    approximation()
    return its


class VelocityGMG:
    """GeometricMultigridSolver on the velocity block (StokesMCFGMRESSolver's A-block
    solver): Chebyshev smoothing, P2toP2QuadraticVector transfers, MINRES with
    identity preconditioner on the coarsest level"""

    def __init__(self, mesh, ops, invd, coeffs, min_level, max_level, pre=3, post=3, coarse_iter=10, coarse_tol=1e-6,
                 project=False):
        self.m, self.ops, self.invd, self.coeffs = mesh, ops, invd, coeffs
        self.minL, self.maxL, self.pre, self.post = min_level, max_level, pre, post
        self.coarse_iter, self.coarse_tol = coarse_iter, coarse_tol
        # P2toP2QuadraticVector{Restriction,Prolongation}WithProjection (post-projection of
        # the transferred function on the FreeSlip DoFs)
        self.project = project

    def cycle(self, L, u, b):
        op = self.ops[L]
        m = op.vmask(GMG_FLAG)

        def A(z):
            y = np.zeros_like(z)
            op.apply(z, None, y, None, GMG_FLAG, StokesOp.A_U)
            return y
        if L == self.minL:
            minres_solve(A, lambda v: v.copy(), u, b, m, self.coarse_iter, self.coarse_tol)
            return
        for _ in range(self.pre):
            stokes_chebyshev(op, self.invd[L], self.coeffs, u, b, GMG_FLAG, project=False)
        r = np.zeros_like(u)
        r[m] = b[m] - A(u)[m]
        c = self.ops[L - 1]
        cb = np.zeros(3 * c.n2)
        for k in range(3):
            coarse = np.zeros(c.n2)
            self.m.p2_restrict(L, np.ascontiguousarray(r[k * op.n2:(k + 1) * op.n2]), coarse, GMG_FLAG)
            cb[k * c.n2:(k + 1) * c.n2] = coarse
        if self.project:
            c.project(cb)
        cx = np.zeros_like(cb)
        self.cycle(L - 1, cx, cb)
        for k in range(3):
            fine = np.ascontiguousarray(u[k * op.n2:(k + 1) * op.n2])
            self.m.p2_prolongate(L - 1, np.ascontiguousarray(cx[k * c.n2:(k + 1) * c.n2]), fine, GMG_FLAG, True)
            u[k * op.n2:(k + 1) * op.n2] = fine
        if self.project:
            op.project(u)
        for _ in range(self.post):
            stokes_chebyshev(op, self.invd[L], self.coeffs, u, b, GMG_FLAG, project=False)


class StokesFGMRES:
    """FGMRESSolver<Stokes> with BlockFactorisationPreconditioner (StokesMCFGMRESSolver,
    terraneo/solvers/StokesMCFGMRESSolver.hpp:50-160; StokesBlockPreconditioners.hpp:366-448):
    x.p = -S^-1 b.p (CG on the 1/eta pressure mass, inverse lumped diagonal as
    preconditioner), x.u = one velocity GMG V-cycle on b.u - B^T x.p"""

    def __init__(self, mesh, min_level, max_level, c0=0.0, c1=0.0, rho=None, order=3, a_pre=3, a_post=3,
                 a_coarse_iter=10, a_coarse_tol=1e-6, schur_iter=50, schur_tol=1e-6, freeslip=False, power_its=25,
                 seed=42):
        self.m, self.minL, self.maxL, self.freeslip = mesh, min_level, max_level, freeslip
        self.ops = {L: StokesOp(mesh, L, c0, c1, freeslip) for L in range(min_level, max_level + 1)}
        self.invd = {L: self.ops[L].inv_diag() for L in self.ops}
        if rho is None:
            x = np.random.default_rng(seed).uniform(-1, 1, 3 * self.ops[max_level].n2)
            rho = stokes_power_iteration(self.ops[max_level], self.invd[max_level], x, power_its)
        self.rho = rho
        # the TerraNeo interval factors (TerraNeoParameters.hpp:82-83): [0.1 rho, 1.2 rho]
        coeffs = chebyshev_coefficients(order, 0.1 * rho, 1.2 * rho)
        self.vgmg = VelocityGMG(mesh, self.ops, self.invd, coeffs, min_level, max_level, a_pre, a_post, a_coarse_iter,
                                a_coarse_tol, project=freeslip)
        self.schur_iter, self.schur_tol = schur_iter, schur_tol
        op = self.ops[max_level]
        one, d = np.ones(op.n1), np.zeros(op.n1)
        op.apply_raw(None, one, None, d, ALL, StokesOp.K_MASS)
        self.lumped_inv = np.where(d != 0.0, 1.0 / np.where(d != 0.0, d, 1.0), 0.0)

    def kmass(self, p):
        y = np.zeros_like(p)
        self.ops[self.maxL].apply_raw(None, p, None, y, ALL, StokesOp.K_MASS)
        return y

    def precondition(self, vu, vp):
        op = self.ops[self.maxL]
        m = op.vmask(GMG_FLAG)
        zp = np.zeros_like(vp)
        cg_solve(self.kmass, lambda r: self.lumped_inv * r, zp, vp, np.ones(len(vp), bool), self.schur_iter,
                 self.schur_tol, 1e-16)
        tu = np.zeros_like(vu)
        op.apply(None, zp, tu, None, GMG_FLAG, StokesOp.BT_P)
        tu[m] = vu[m] + tu[m]
        zp = -zp
        if self.freeslip:
            op.project(tu)
        zu = np.zeros_like(vu)
        self.vgmg.cycle(self.maxL, zu, tu)
        return zu, zp

    def solve(self, u, p, bu, bp, max_iter=100, restart=50, tol=1e-7):
        op = self.ops[self.maxL]
        n2 = len(u)
        m = np.concatenate([op.vmask(GMG_FLAG), np.ones(len(p), bool)])
        A = lambda z: np.concatenate(stokes_apply_all(op, z[:n2], z[n2:], GMG_FLAG))  # noqa: E731

        def prec(v):
            zu, zp = self.precondition(v[:n2].copy(), v[n2:].copy())
            return np.concatenate([zu, zp])
        x = np.concatenate([u, p])
        hist = []
        its = fgmres(A, prec, x, np.concatenate([bu, bp]), m, max_iter + 1, tol, 0.0, restart, hist=hist)
        u[:] = x[:n2]
        p[:] = x[n2:]
        return its, np.array(hist)


# ---- P2 Gauss-Seidel and the P2 GMG of P2GMG3DConvergenceTest ---------------------
def gs_sweep(A, x, b, m):
    """one forward Gauss-Seidel sweep in canonical order over the rows of mask m
    (SORSmoothable::smooth_gs restated on the assembled matrix A):
    (D + L)_mm x_m = b_m - U_mm x_m - A_m,~m x_~m, the other entries unchanged"""
    import scipy.sparse as sp
    from scipy.sparse.linalg import spsolve_triangular
    Am = A[m]
    M = Am[:, m]
    rhs = b[m] - sp.triu(M, 1) @ x[m] - Am[:, ~m] @ x[~m]
    x[m] = spsolve_triangular(sp.tril(M).tocsr(), rhs, lower=True)


class P2GSGMG:
    """GeometricMultigridSolver<P2ConstantLaplaceOperator> with GaussSeidelSmoother, CG
    (no preconditioner) on the coarsest level and quadratic transfers
    (P2GMG3DConvergenceTest.cpp:60-100, GeometricMultigridSolver.hpp:207-300); the
    sweeps run in canonical order, or colour by colour when `colours` maps each
    level to a colour per DoF"""

    def __init__(self, mesh, min_level, max_level, pre=2, post=2, coarse_iter=1000, coarse_tol=1e-14, colours=None):
        self.m, self.minL, self.maxL, self.pre, self.post = mesh, min_level, max_level, pre, post
        self.colours = colours
        self.coarse_iter, self.coarse_tol = coarse_iter, coarse_tol
        self.ops = {L: P2Op(mesh, L) for L in range(min_level, max_level + 1)}
        self.A = {L: self.ops[L].matrix() for L in self.ops}
        self.mask = {L: (mesh.flags(L + 1) & GMG_FLAG) != 0 for L in self.ops}

    def residual(self, L, x, b):
        r = np.zeros_like(x)
        m = self.mask[L]
        r[m] = b[m] - (self.A[L] @ x)[m]
        return r

    def cycle(self, L, x, b):
        if L == self.minL:
            self.ops[L].cg(x, b, GMG_FLAG, self.coarse_iter, self.coarse_tol, 1e-16, jacobi=False)
            return
        for _ in range(self.pre):
            self.sweep(L, x, b)
        r = self.residual(L, x, b)
        cb = np.zeros(self.ops[L - 1].n)
        self.m.p2_restrict(L, r, cb, GMG_FLAG)
        cx = np.zeros_like(cb)
        self.cycle(L - 1, cx, cb)
        self.m.p2_prolongate(L - 1, cx, x, GMG_FLAG, True)
        for _ in range(self.post):
            self.sweep(L, x, b)

    def sweep(self, L, x, b):
        if self.colours is None:
            gs_sweep(self.A[L], x, b, self.mask[L])
        else:
            sor_multicolour(self.A[L], self.colours[L], self.mask[L], x, b)
