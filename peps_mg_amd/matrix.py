"""Heptadiagonal 7-point matrix (module matrix, src/matrix.f90).

The GPU path is matrix-free; this AoS array is what the reference's
callers build and pass in, and the library checks it against the grid
metrics.  coeff has numpy shape (nz, ny, nx, 7) = Fortran coeff(0:6,nx,ny,nz).
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class MatrixHeptadiagonal:
    """type matrix_heptadiagonal (src/matrix.f90:27-30)."""
    dof: int
    coeff: np.ndarray


def matrix_heptadiagonal_create(sdm):
    """matrix_heptadiagonal_create (src/matrix.f90:39-94), vectorised with
    the same per-entry arithmetic."""
    nx, ny, nz = sdm.nx, sdm.ny, sdm.nz
    ix = np.arange(1, nx + 1)
    jy = np.arange(1, ny + 1)
    kz = np.arange(1, nz + 1)
    dxmp = 1.0 / (sdm.dxm[ix] * sdm.dxg[ix])
    dxmn = 1.0 / (sdm.dxm[ix] * sdm.dxg[ix + 1])
    dymp = 1.0 / (sdm.dym[jy] * sdm.dyg[jy])
    dymn = 1.0 / (sdm.dym[jy] * sdm.dyg[jy + 1])
    dzmp = 1.0 / (sdm.dzm[kz] * sdm.dzg[kz])
    dzmn = 1.0 / (sdm.dzm[kz] * sdm.dzg[kz + 1])
    coeff = np.empty((nz, ny, nx, 7))
    X = lambda v: v[None, None, :]
    Y = lambda v: v[None, :, None]
    Z = lambda v: v[:, None, None]
    c1 = np.broadcast_to(X(dxmp), (nz, ny, nx))
    c2 = np.broadcast_to(X(dxmn), (nz, ny, nx))
    c3 = np.broadcast_to(Y(dymp), (nz, ny, nx))
    c4 = np.broadcast_to(Y(dymn), (nz, ny, nx))
    c5 = np.broadcast_to(Z(dzmp), (nz, ny, nx))
    c6 = np.broadcast_to(Z(dzmn), (nz, ny, nx))
    coeff[..., 0] = -(((((c1 + c2) + c3) + c4) + c5) + c6)
    coeff[..., 1] = c1
    coeff[..., 2] = c2
    coeff[..., 3] = c3
    coeff[..., 4] = c4
    coeff[..., 5] = c5
    coeff[..., 6] = c6
    return MatrixHeptadiagonal(nx * ny * nz, coeff)


def matrix_heptadiagonal_destroy(a_poisson):
    """matrix_heptadiagonal_destroy (src/matrix.f90:100-109)."""
    a_poisson.dof = 0
    a_poisson.coeff = None
