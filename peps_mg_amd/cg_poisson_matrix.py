"""Conjugate gradient (module cg_poisson_matrix, src/cg_poisson_matrix.f90) on the GPU."""
import ctypes as C

from . import _lib
from .geometry import grid_descriptor


def cg_iterator_poisson_matrix(sol, a_poisson, rhs, dm, maxiteration, is_aggregated=(True,) * 3):
    """src/cg_poisson_matrix.f90:157-241; sol updated in place."""
    g = grid_descriptor(dm)
    n = (dm.nx, dm.ny, dm.nz)
    _lib.check(_lib.lib().pepsmg_cg_iterator(C.byref(g), _lib.padded_ptr(sol, *n, "sol"),
                                             _lib.coeff_ptr(a_poisson, *n), _lib.padded_ptr(rhs, *n, "rhs"),
                                             maxiteration))


def cg_solver_poisson_matrix(sol, a_poisson, rhs, dm, maxiteration, tolerance, is_aggregated=(True,) * 3):
    """src/cg_poisson_matrix.f90:47-146; returns (iter, relative residual)."""
    g = grid_descriptor(dm)
    n = (dm.nx, dm.ny, dm.nz)
    it = C.c_int32(0)
    rel = C.c_double(0)
    _lib.check(_lib.lib().pepsmg_cg_solver(C.byref(g), _lib.padded_ptr(sol, *n, "sol"),
                                           _lib.coeff_ptr(a_poisson, *n), _lib.padded_ptr(rhs, *n, "rhs"),
                                           maxiteration, tolerance, C.byref(it), C.byref(rel)))
    return it.value, rel.value
