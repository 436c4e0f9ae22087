"""MI355X implementation of PEPS_MG's geometric-multigrid V-cycle.

Modules mirror the reference's Fortran modules: geometry, matrix,
multigrid, rbgs_poisson_matrix, cg_poisson_matrix, poisson_matrix_operator.
Compute runs in libpepsmg.so (HIP, gfx950) through its C-ABI
(include/pepsmg.h).
"""
from . import _lib  # noqa: F401
