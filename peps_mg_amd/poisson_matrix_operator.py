"""Matrix-vector product and dot product (module poisson_matrix_operator,
src/poisson_matrix_operator.f90) on the GPU."""
import ctypes as C

from . import _lib
from .geometry import grid_descriptor


def vv_dot_3d_matrix(x, y, nx, ny, nz, is_serial=(True,) * 3):
    """src/poisson_matrix_operator.f90:40-99; returns the interior dot product."""
    out = C.c_double(0)
    _lib.check(_lib.lib().pepsmg_vv_dot(nx, ny, nz, _lib.padded_ptr(x, nx, ny, nz, "x"),
                                        _lib.padded_ptr(y, nx, ny, nz, "y"), C.byref(out)))
    return out.value


def mv_mul_poisson_matrix(y, a_poisson, x, dm, is_serial=(True,) * 3):
    """src/poisson_matrix_operator.f90:110-161; y written in place (ghosts 0)."""
    g = grid_descriptor(dm)
    n = (dm.nx, dm.ny, dm.nz)
    _lib.check(_lib.lib().pepsmg_mv_mul(C.byref(g), _lib.padded_ptr(y, *n, "y"), _lib.coeff_ptr(a_poisson, *n),
                                        _lib.padded_ptr(x, *n, "x")))
