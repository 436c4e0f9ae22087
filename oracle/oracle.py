"""ctypes binding of the oracle (oracle/pepsmg_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, as the checker.  The product library
(peps_mg_amd, libpepsmg.so) never imports or calls it.

Array conventions follow the reference's Fortran storage, seen from numpy
in C order: a padded grid array (0:nx+1, 0:ny+1, 0:nz+1) is a float64
array of shape (nz+2, ny+2, nx+2); the AoS matrix coeff(0:6,nx,ny,nz) is
shape (nz, ny, nx, 7).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.ora_axis_create.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp]
        L.ora_matrix_create.argtypes = [C.c_int] * 3 + [_dp] * 7
        L.ora_vv_dot.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int]
        L.ora_vv_dot.restype = C.c_double
        L.ora_mv_mul.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, C.c_int]
        L.ora_residual.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int]
        L.ora_rbgs_iterator.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
        L.ora_rbgs_solver.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_double, C.c_double, _dp]
        L.ora_rbgs_solver.restype = C.c_int
        L.ora_cg_iterator.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.ora_cg_solver.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _dp]
        L.ora_cg_solver.restype = C.c_int
        L.ora_mg_create.argtypes = ([C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int] + [_dp] * 9 +
                                    [C.c_double] * 3 + [C.c_int] * 5)
        L.ora_mg_create.restype = C.c_int
        L.ora_mg_destroy.argtypes = [C.c_void_p]
        L.ora_mg_level_dims.argtypes = [C.c_void_p, C.c_int, _ip]
        L.ora_mg_level_metrics.argtypes = [C.c_void_p, C.c_int, _dp]
        L.ora_mg_lvc.argtypes = [C.c_void_p, _ip]
        L.ora_mg_restriction.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
        L.ora_mg_prolongation.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
        L.ora_mg_solve_vcycle.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.c_int, C.c_double,
                                          C.c_double, _dp, _ip, _dp]
        L.ora_mg_solve_vcycle.restype = C.c_int
        L.ora_uniform_field.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_uint64]
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def padded(nx, ny, nz):
    return np.zeros((nz + 2, ny + 2, nx + 2), dtype=np.float64)


def _dims(a):
    nz, ny, nx = a.shape
    return nx - 2, ny - 2, nz - 2


# --- geometry.f90:139-281 -------------------------------------------------
def axis_create(n, o, l, a):
    dm = np.zeros(n + 2); dg = np.zeros(n + 2); g = np.zeros(n + 2)
    lib().ora_axis_create(n, o, l, a, _p(dm), _p(dg), _p(g))
    return dm, dg, g


class Grid:
    """Fine-level metrics of one subdomain (1 rank = the global domain)."""

    def __init__(self, nx, ny, nz, o=(0.0, 0.0, 0.0), l=(1.0, 1.0, 1.0), a=(1.0, 1.0, 1.0)):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.o = tuple(float(v) for v in o)
        self.dxm, self.dxg, self.xg = axis_create(nx, o[0], l[0], a[0])
        self.dym, self.dyg, self.yg = axis_create(ny, o[1], l[1], a[1])
        self.dzm, self.dzg, self.zg = axis_create(nz, o[2], l[2], a[2])
        # geometry.f90:373-375: subdomain origin
        self.sox = self.xg[1] - 0.5 * self.dxm[1]
        self.soy = self.yg[1] - 0.5 * self.dym[1]
        self.soz = self.zg[1] - 0.5 * self.dzm[1]

    def matrix(self):
        return matrix_create(self.nx, self.ny, self.nz, self.dxm, self.dxg, self.dym, self.dyg,
                             self.dzm, self.dzg)


# --- matrix.f90:39-94 -----------------------------------------------------
def matrix_create(nx, ny, nz, dxm, dxg, dym, dyg, dzm, dzg):
    coeff = np.zeros((nz, ny, nx, 7))
    lib().ora_matrix_create(nx, ny, nz, _p(np.ascontiguousarray(dxm)), _p(np.ascontiguousarray(dxg)),
                            _p(np.ascontiguousarray(dym)), _p(np.ascontiguousarray(dyg)),
                            _p(np.ascontiguousarray(dzm)), _p(np.ascontiguousarray(dzg)), _p(coeff))
    return coeff


# --- poisson_matrix_operator.f90 -----------------------------------------
def vv_dot(x, y):
    nx, ny, nz = _dims(x)
    return lib().ora_vv_dot(_p(x), _p(y), nx, ny, nz)


def mv_mul(coeff, x):
    nx, ny, nz = _dims(x)
    y = np.empty_like(x)
    lib().ora_mv_mul(_p(y), _p(coeff), _p(x), nx, ny, nz)
    return y


def residual(coeff, x, b):
    nx, ny, nz = _dims(x)
    r = np.empty_like(x)
    lib().ora_residual(_p(r), _p(coeff), _p(x), _p(b), nx, ny, nz)
    return r


# --- rbgs_poisson_matrix.f90 ----------------------------------------------
def rbgs_iterator(sol, coeff, rhs, maxiteration, omega):
    """In place on sol, like the reference (intent(inout))."""
    nx, ny, nz = _dims(sol)
    lib().ora_rbgs_iterator(_p(sol), _p(coeff), _p(rhs), nx, ny, nz, maxiteration, omega)
    return sol


def rbgs_solver(sol, coeff, rhs, maxiteration, tolerance, omega):
    nx, ny, nz = _dims(sol)
    rel = C.c_double(0.0)
    it = lib().ora_rbgs_solver(_p(sol), _p(coeff), _p(rhs), nx, ny, nz, maxiteration, tolerance,
                               omega, C.byref(rel))
    return it, rel.value


# --- cg_poisson_matrix.f90 ------------------------------------------------
def cg_iterator(sol, coeff, rhs, maxiteration):
    nx, ny, nz = _dims(sol)
    lib().ora_cg_iterator(_p(sol), _p(coeff), _p(rhs), nx, ny, nz, maxiteration)
    return sol


def cg_solver(sol, coeff, rhs, maxiteration, tolerance):
    nx, ny, nz = _dims(sol)
    rel = C.c_double(0.0)
    it = lib().ora_cg_solver(_p(sol), _p(coeff), _p(rhs), nx, ny, nz, maxiteration, tolerance,
                             C.byref(rel))
    return it, rel.value


# --- multigrid.f90 --------------------------------------------------------
class Multigrid:
    """multigrid_create / multigrid_solve_vcycle / multigrid_destroy."""

    def __init__(self, grid, nlevel, ncycle, aggr=0, cga_level=0, smoother=0):
        self.grid = grid
        h = C.c_void_p()
        rc = lib().ora_mg_create(C.byref(h), grid.nx, grid.ny, grid.nz,
                                 _p(grid.dxm), _p(grid.dxg), _p(grid.xg),
                                 _p(grid.dym), _p(grid.dyg), _p(grid.yg),
                                 _p(grid.dzm), _p(grid.dzg), _p(grid.zg),
                                 grid.sox, grid.soy, grid.soz, nlevel, ncycle, aggr, cga_level,
                                 smoother)
        if rc != 0:
            raise ValueError(f"oracle multigrid_create failed with code {rc}")
        self.h = h
        self.nlevel = nlevel
        self.ncycle = ncycle

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_mg_destroy(self.h)
            self.h = None

    def level_dims(self, l):
        d = (C.c_int * 3)()
        lib().ora_mg_level_dims(self.h, l, d)
        return tuple(d)

    def level_metrics(self, l):
        nx, ny, nz = self.level_dims(l)
        out = np.zeros(2 * (nx + ny + nz + 6))
        lib().ora_mg_level_metrics(self.h, l, _p(out))
        s = np.cumsum([0, nx + 2, nx + 2, ny + 2, ny + 2, nz + 2, nz + 2])
        keys = ["dxm", "dxg", "dym", "dyg", "dzm", "dzg"]
        return {k: out[s[i]:s[i + 1]].copy() for i, k in enumerate(keys)}

    def lvc(self):
        d = (C.c_int * 3)()
        lib().ora_mg_lvc(self.h, d)
        return tuple(d)

    def restriction(self, val_f, level):
        nx, ny, nz = self.level_dims(level + 1)
        out = padded(nx, ny, nz)
        lib().ora_mg_restriction(self.h, _p(out), _p(val_f), level)
        return out

    def prolongation(self, val_c, level):
        nx, ny, nz = self.level_dims(level)
        out = padded(nx, ny, nz)
        lib().ora_mg_prolongation(self.h, _p(out), _p(val_c), level)
        return out

    def solve_vcycle(self, sol, rsd, coeff0, rhs, maxiteration, tolerance, omega):
        """Returns (cycles, rho history, coarse iteration counts, res0)."""
        hist = np.zeros(self.ncycle)
        cit = np.zeros(self.ncycle, dtype=np.int32)
        res0 = C.c_double(0.0)
        n = lib().ora_mg_solve_vcycle(self.h, _p(sol), _p(rsd), _p(coeff0), _p(rhs), maxiteration,
                                      tolerance, omega, _p(hist), cit.ctypes.data_as(_ip),
                                      C.byref(res0))
        return n, hist[:n].copy(), cit[:n].copy(), res0.value


def uniform_field(nx, ny, nz, seed):
    out = padded(nx, ny, nz)
    lib().ora_uniform_field(_p(out), nx, ny, nz, seed)
    return out
