"""Geometric multigrid V-cycle (module multigrid, src/multigrid.f90) on the GPU.

Module-level functions keep the reference's single-instance interface
(multigrid_create / multigrid_solve_vcycle / multigrid_destroy, the
hierarchy held as module state, src/multigrid.f90:40-59); the Multigrid
class is the same solver as an object, for callers that need several.
All work runs in libpepsmg.so through its C-ABI.
"""
import ctypes as C

import numpy as np

from . import _lib
from .geometry import grid_descriptor


class SolveResult:
    """What one solve reports (the reference only prints it,
    src/multigrid.f90:1972)."""

    def __init__(self, stats, hist, citers):
        self.cycles = stats.cycles
        self.refine_steps = stats.refine_steps
        self.converged = bool(stats.converged)
        self.res0 = stats.res0
        self.rho = stats.rho
        n = stats.refine_steps if stats.refine_steps > 0 else stats.cycles
        self.history = hist[:n].copy()
        self.coarse_iters = citers[:stats.cycles].copy()
        self.device_ms = stats.device_ms

    def __repr__(self):
        return (f"SolveResult(cycles={self.cycles}, refine_steps={self.refine_steps}, "
                f"converged={self.converged}, rho={self.rho:.3e}, device_ms={self.device_ms:.3f})")


class Multigrid:
    """A multigrid hierarchy on the GPU for one subdomain."""

    def __init__(self, sdm, nlevel, ncycle, aggr_method=0, single_aggr_level=0,
                 adaptive_aggr_level=(0, 0, 0), **options):
        L = _lib.lib()
        self.sdm = sdm
        self.grid = grid_descriptor(sdm)
        self.ncycle = int(ncycle)
        h = C.c_void_p()
        agg = (C.c_int32 * 3)(*adaptive_aggr_level)
        _lib.check(L.pepsmg_multigrid_create(C.byref(self.grid), nlevel, ncycle, aggr_method,
                                             single_aggr_level, agg, C.byref(h)))
        self.h = h
        self.options = _lib.default_options()
        if options:
            self.set_options(**options)

    def set_options(self, **kw):
        for k, v in kw.items():
            if not hasattr(self.options, k):
                raise KeyError(f"unknown option {k}")
            setattr(self.options, k, _lib.option_value(k, v))
        _lib.check(_lib.lib().pepsmg_set_options(self.h, C.byref(self.options)))

    def _stats(self):
        cap = max(self.ncycle, 1) + 64
        hist = np.zeros(cap)
        cit = np.zeros(cap, dtype=np.int32)
        st = _lib.Stats()
        st.history_cap = cap
        st.history = hist.ctypes.data_as(C.POINTER(C.c_double))
        st.coarse_iters = cit.ctypes.data_as(C.POINTER(C.c_int32))
        return st, hist, cit

    def solve_vcycle(self, sol, rsd, a_poisson, rhs, maxiteration, tolerance, omega_sor):
        """multigrid_solve_vcycle on host arrays; sol and rsd updated in place."""
        st, hist, cit = self._stats()
        n = (self.sdm.nx, self.sdm.ny, self.sdm.nz)
        _lib.check(_lib.lib().pepsmg_multigrid_solve_vcycle(
            self.h, _lib.padded_ptr(sol, *n, "sol"), None if rsd is None else _lib.padded_ptr(rsd, *n, "rsd"),
            _lib.coeff_ptr(a_poisson, *n), _lib.padded_ptr(rhs, *n, "rhs"), maxiteration, tolerance, omega_sor,
            C.byref(st)))
        return SolveResult(st, hist, cit)

    def solve_vcycle_device(self, d_sol, d_rsd, d_rhs, maxiteration, tolerance, omega_sor, stream=None):
        """Device arrays given as integer addresses (e.g. torch data_ptr())."""
        st, hist, cit = self._stats()
        _lib.check(_lib.lib().pepsmg_multigrid_solve_vcycle_device(
            self.h, C.c_void_p(d_sol), C.c_void_p(d_rsd) if d_rsd else None, C.c_void_p(d_rhs), maxiteration,
            tolerance, omega_sor, C.c_void_p(stream) if stream else None, C.byref(st)))
        return SolveResult(st, hist, cit)

    def destroy(self):
        if getattr(self, "h", None):
            _lib.lib().pepsmg_multigrid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


_current = None


def multigrid_create(sdm, nlevel, ncycle, aggr_method, single_aggr_level, adaptive_aggr_level):
    """multigrid_create (src/multigrid.f90:75-159)."""
    global _current
    if _current is not None:
        _current.destroy()
    _current = Multigrid(sdm, nlevel, ncycle, aggr_method, single_aggr_level, adaptive_aggr_level)
    return _current


def multigrid_set_options(**kw):
    _current.set_options(**kw)


def multigrid_solve_vcycle(sol, rsd, a_poisson, rhs, sdm, maxiteration, tolerance, omega_sor):
    """multigrid_solve_vcycle (src/multigrid.f90:1864-1894)."""
    if _current is None:
        raise RuntimeError("multigrid_create has not been called")
    return _current.solve_vcycle(sol, rsd, a_poisson, rhs, maxiteration, tolerance, omega_sor)


def multigrid_destroy():
    """multigrid_destroy (src/multigrid.f90:1320-1334)."""
    global _current
    if _current is not None:
        _current.destroy()
        _current = None
