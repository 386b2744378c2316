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
            "or_mesh_shell": (_p, [_i, _i, _d, _d]),
            "or_mesh_blend": (None, [_p, _p]),
            "or_shell_map": (None, [_p, _p, _p, _p]),
            "or_quadrature": (None, [_p, _p]),
            "or_p2_element_matrix": (None, [_p, _p, _i, _p, _p, _p, _p]),
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
    def shell(ntan=5, nrad=3, rmin=0.55, rmax=1.0):
        return Mesh(lib().or_mesh_shell(ntan, nrad, rmin, rmax))

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


def p2_element_matrix(p, blend=None, k=None):
    p = np.ascontiguousarray(p, dtype=np.float64)
    k = k or KCoeff()
    A = np.zeros((10, 10))
    b = None if blend is None else np.ascontiguousarray(blend, dtype=np.float64)
    lib().or_p2_element_matrix(_ptr(p), None if b is None else _ptr(b), *k.args(), _ptr(A))
    return A


class P2Op:
    def __init__(self, mesh, level, k=None):
        self.mesh = mesh
        self.level = level
        self.k = k or KCoeff()
        self.h = lib().or_p2op_create(mesh.h, level, *self.k.args())
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
