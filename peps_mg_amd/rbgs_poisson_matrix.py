"""Red/black Gauss-Seidel/SOR (module rbgs_poisson_matrix,
src/rbgs_poisson_matrix.f90) on the GPU.  is_aggregated is accepted for
interface compatibility; one process is aggregated in every direction."""
import ctypes as C

from . import _lib
from .geometry import grid_descriptor


def rbgs_iterator_poisson_matrix(sol, a_poisson, rhs, dm, maxiteration, omega, is_aggregated=(True,) * 3):
    """src/rbgs_poisson_matrix.f90:191-279; sol updated in place."""
    g = grid_descriptor(dm)
    n = (dm.nx, dm.ny, dm.nz)
    _lib.check(_lib.lib().pepsmg_rbgs_iterator(C.byref(g), _lib.padded_ptr(sol, *n, "sol"),
                                               _lib.coeff_ptr(a_poisson, *n), _lib.padded_ptr(rhs, *n, "rhs"),
                                               maxiteration, omega))


def rbgs_iterator_fp32(sol, rhs, dm, maxiteration, omega):
    """The red/black iterator in FP32 storage and arithmetic (the inner
    smoother of the mixed-precision path; no counterpart in the reference):
    sol and rhs are rounded to FP32, the FP32 result is written back into
    sol (FP64 padded arrays, updated in place)."""
    g = grid_descriptor(dm)
    n = (dm.nx, dm.ny, dm.nz)
    _lib.check(_lib.lib().pepsmg_rbgs_iterator_fp32(C.byref(g), _lib.padded_ptr(sol, *n, "sol"),
                                                    _lib.padded_ptr(rhs, *n, "rhs"), maxiteration, omega))


def rbgs_solver_poisson_matrix(sol, a_poisson, rhs, dm, maxiteration, tolerance, omega,
                               is_aggregated=(True,) * 3):
    """src/rbgs_poisson_matrix.f90:47-179; returns (iter, relative residual)."""
    g = grid_descriptor(dm)
    n = (dm.nx, dm.ny, dm.nz)
    it = C.c_int32(0)
    rel = C.c_double(0)
    _lib.check(_lib.lib().pepsmg_rbgs_solver(C.byref(g), _lib.padded_ptr(sol, *n, "sol"),
                                             _lib.coeff_ptr(a_poisson, *n), _lib.padded_ptr(rhs, *n, "rhs"),
                                             maxiteration, tolerance, omega, C.byref(it), C.byref(rel)))
    return it.value, rel.value
