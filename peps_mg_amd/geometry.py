"""Domain and subdomain types (module geometry, src/geometry.f90).

One process owns the whole grid, so the subdomain is the global domain and
is aggregated in every direction.  Arrays use the reference's padded
Fortran layout a(0:nx+1, 0:ny+1, 0:nz+1), seen from numpy in C order as
shape (nz+2, ny+2, nx+2).
"""
from dataclasses import dataclass, field

import numpy as np


def _axis(n, o, l, a):
    """One axis of geometry_domain_create (src/geometry.f90:176-209)."""
    dm = np.zeros(n + 2)
    dg = np.zeros(n + 2)
    g = np.zeros(n + 2)
    g[0] = o
    if a == 1.0:
        for i in range(1, n + 1):
            dm[i] = l / float(n)
    else:
        dm[1] = 0.5 * l * (a - 1.0) / (a ** (n // 2) - 1.0)
        for i in range(2, n // 2 + 1):
            dm[i] = dm[i - 1] * a
            dm[n - i + 1] = dm[i]
        dm[n] = dm[1]
    for i in range(1, n + 1):
        dg[i] = 0.5 * (dm[i] + dm[i - 1])
        g[i] = g[i - 1] + dg[i]
    dg[n + 1] = 0.5 * (dm[n + 1] + dm[n])
    g[n + 1] = g[n] + dg[n + 1]
    return dm, dg, g


@dataclass
class Domain:
    """type domain (src/geometry.f90:30-53)."""
    nx: int
    ny: int
    nz: int
    lx: float
    ly: float
    lz: float
    ox: float
    oy: float
    oz: float
    dxm: np.ndarray
    dym: np.ndarray
    dzm: np.ndarray
    dxg: np.ndarray
    dyg: np.ndarray
    dzg: np.ndarray
    xg: np.ndarray
    yg: np.ndarray
    zg: np.ndarray
    is_periodic: tuple = (False, False, False)


@dataclass
class Subdomain:
    """type subdomain (src/geometry.f90:59-118) on one process."""
    nx: int
    ny: int
    nz: int
    lx: float
    ly: float
    lz: float
    ox: float
    oy: float
    oz: float
    dxm: np.ndarray
    dym: np.ndarray
    dzm: np.ndarray
    dxg: np.ndarray
    dyg: np.ndarray
    dzg: np.ndarray
    xg: np.ndarray
    yg: np.ndarray
    zg: np.ndarray
    x: np.ndarray = field(repr=False, default=None)
    b: np.ndarray = field(repr=False, default=None)
    r: np.ndarray = field(repr=False, default=None)
    is_aggregated: tuple = (True, True, True)
    is_periodic: tuple = (False, False, False)

    def padded(self):
        return np.zeros((self.nz + 2, self.ny + 2, self.nx + 2))


def geometry_domain_create(nx, ny, nz, ox, oy, oz, lx, ly, lz, ax=1.0, ay=1.0, az=1.0,
                           period=(False, False, False)):
    """geometry_domain_create (src/geometry.f90:139-281)."""
    if any(period):
        raise ValueError("periodic boundaries are not supported (homogeneous Dirichlet only)")
    dxm, dxg, xg = _axis(nx, ox, lx, ax)
    dym, dyg, yg = _axis(ny, oy, ly, ay)
    dzm, dzg, zg = _axis(nz, oz, lz, az)
    return Domain(nx, ny, nz, lx, ly, lz, ox, oy, oz, dxm, dym, dzm, dxg, dyg, dzg, xg, yg, zg,
                  tuple(period))


def geometry_subdomain_create(gdm):
    """geometry_subdomain_create (src/geometry.f90:304-423) for one process:
    the subdomain is the whole domain; x, b, r are zeroed padded arrays."""
    s = Subdomain(gdm.nx, gdm.ny, gdm.nz, gdm.lx, gdm.ly, gdm.lz,
                  gdm.xg[1] - 0.5 * gdm.dxm[1], gdm.yg[1] - 0.5 * gdm.dym[1], gdm.zg[1] - 0.5 * gdm.dzm[1],
                  gdm.dxm.copy(), gdm.dym.copy(), gdm.dzm.copy(), gdm.dxg.copy(), gdm.dyg.copy(),
                  gdm.dzg.copy(), gdm.xg.copy(), gdm.yg.copy(), gdm.zg.copy())
    s.x = s.padded()
    s.b = s.padded()
    s.r = s.padded()
    s.is_periodic = gdm.is_periodic
    return s


def grid_descriptor(sdm):
    from ._lib import make_grid
    return make_grid(sdm.nx, sdm.ny, sdm.nz, sdm.dxm, sdm.dxg, sdm.dym, sdm.dyg, sdm.dzm, sdm.dzg)
