"""Host-side argument checks of the Python mirror (no GPU call is made):
each padded array must have the grid's shape (nz+2, ny+2, nx+2) in C order
and float64 C-contiguous storage, the AoS matrix the shape (nz, ny, nx, 7);
anything else raises PepsmgError(ERR_ARG) before the C-ABI copies
(nx+2)(ny+2)(nz+2) values through the pointer."""
import numpy as np
import pytest

from peps_mg_amd import _lib, geometry, matrix
from peps_mg_amd.cg_poisson_matrix import cg_iterator_poisson_matrix, cg_solver_poisson_matrix
from peps_mg_amd.poisson_matrix_operator import mv_mul_poisson_matrix, vv_dot_3d_matrix
from peps_mg_amd.rbgs_poisson_matrix import (rbgs_iterator_fp32, rbgs_iterator_poisson_matrix,
                                             rbgs_solver_poisson_matrix)

NX, NY, NZ = 12, 10, 6


def sdm():
    return geometry.geometry_subdomain_create(geometry.geometry_domain_create(NX, NY, NZ, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0))


GOOD = (NZ + 2, NY + 2, NX + 2)
BAD = [np.zeros((NZ, NY, NX)),                          # interior only
       np.zeros((NX + 2, NY + 2, NZ + 2)),              # Fortran extents read as C order
       np.zeros(GOOD, dtype=np.float32),
       np.zeros((NZ + 2, NY + 2, 2 * (NX + 2)))[:, :, ::2],   # not contiguous
       [[0.0]]]


CALLS = {
    "rbgs_iterator": lambda s, a, b: rbgs_iterator_poisson_matrix(a, None, b, s, 1, 1.0),
    "rbgs_iterator_fp32": lambda s, a, b: rbgs_iterator_fp32(a, b, s, 1, 1.0),
    "rbgs_solver": lambda s, a, b: rbgs_solver_poisson_matrix(a, None, b, s, 10, 1e-10, 1.0),
    "cg_iterator": lambda s, a, b: cg_iterator_poisson_matrix(a, None, b, s, 1),
    "cg_solver": lambda s, a, b: cg_solver_poisson_matrix(a, None, b, s, 10, 1e-10),
    "mv_mul": lambda s, a, b: mv_mul_poisson_matrix(a, None, b, s),
    "vv_dot": lambda s, a, b: vv_dot_3d_matrix(a, b, NX, NY, NZ),
}


@pytest.mark.parametrize("name", sorted(CALLS))
@pytest.mark.parametrize("k", range(len(BAD)))
@pytest.mark.parametrize("which", [0, 1])
def test_wrong_array_rejected(name, k, which):
    s = sdm()
    args = [np.zeros(GOOD), np.zeros(GOOD)]
    args[which] = BAD[k]
    with pytest.raises(_lib.PepsmgError) as e:
        CALLS[name](s, *args)
    assert e.value.code == _lib.ERR_ARG


def test_wrong_matrix_shape_rejected():
    s = sdm()
    a = matrix.matrix_heptadiagonal_create(s)
    a.coeff = a.coeff.reshape(NX, NY, NZ, 7)
    with pytest.raises(_lib.PepsmgError) as e:
        mv_mul_poisson_matrix(np.zeros(GOOD), a, np.zeros(GOOD), s)
    assert e.value.code == _lib.ERR_ARG
