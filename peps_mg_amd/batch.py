"""Batched solves: many independent Poisson problems in shared launches.

Each problem has its own grid (a Subdomain of module geometry), number of
coarse levels and number of right-hand sides; every (problem, right-hand
side) pair is an item.  A solve runs the V-cycle of multigrid_solve_vcycle
(src/multigrid.f90:1907-1976) for all items at once, one launch per
operation and level, replayed from a captured hipGraph (pepsmg_batch_* in
include/pepsmg.h).  The reference has no batched interface; this module is
an addition of this implementation.
"""
import ctypes as C

import numpy as np

from . import _lib
from .geometry import grid_descriptor
from .multigrid import SolveResult


class BatchSolver:
    def __init__(self, sdms, nlevels, nrhs, ncycle=200, **options):
        L = _lib.lib()
        if not (len(sdms) == len(nlevels) == len(nrhs)) or not sdms:
            raise _lib.PepsmgError(_lib.ERR_ARG, "sdms, nlevels and nrhs must have the same non-zero length")
        self.sdms = list(sdms)
        self.nlevels = [int(x) for x in nlevels]
        self.grids = (_lib.Grid * len(sdms))(*[grid_descriptor(s) for s in sdms])
        self.nrhs = [int(r) for r in nrhs]
        self.ncycle = int(ncycle)
        h = C.c_void_p()
        _lib.check(L.pepsmg_batch_create(len(sdms), self.grids, (C.c_int32 * len(sdms))(*nlevels),
                                         (C.c_int32 * len(sdms))(*self.nrhs), self.ncycle, C.byref(h)))
        self.h = h
        self.size = L.pepsmg_batch_size(h)
        self.options = _lib.default_options()
        if options:
            self.set_options(**options)

    def item_shapes(self):
        """Padded (nz+2, ny+2, nx+2) shape of every item, in item order."""
        out = []
        for s, r in zip(self.sdms, self.nrhs):
            out += [(s.nz + 2, s.ny + 2, s.nx + 2)] * r
        return out

    def set_options(self, **kw):
        for k, v in kw.items():
            if not hasattr(self.options, k):
                raise KeyError(f"unknown option {k}")
            setattr(self.options, k, _lib.option_value(k, v))
        _lib.check(_lib.lib().pepsmg_batch_set_options(self.h, C.byref(self.options)))

    def set_tiles(self, tile_x=None, tile_y=None):
        """Fused-smoother tile shape per problem (lists of nprob; None or 0:
        the options' values)."""
        n = len(self.sdms)
        ax = (C.c_int32 * n)(*tile_x) if tile_x is not None else None
        ay = (C.c_int32 * n)(*tile_y) if tile_y is not None else None
        _lib.check(_lib.lib().pepsmg_batch_set_tiles(self.h, ax, ay))

    def solve_device(self, sols, rhss, maxiteration, tolerance, omega_sor, use_graph=True, stream=None):
        """sols / rhss: one padded float64 device array per item (integer
        addresses or objects with data_ptr()); sols are overwritten."""
        n = self.size
        if len(sols) != n or len(rhss) != n:
            raise _lib.PepsmgError(_lib.ERR_ARG, f"expected {n} solution and right-hand-side arrays")
        addr = lambda a: a if isinstance(a, int) else a.data_ptr()
        ps = (C.c_void_p * n)(*[addr(a) for a in sols])
        pb = (C.c_void_p * n)(*[addr(a) for a in rhss])
        cap = self.ncycle + 2
        hist = np.zeros((n, cap))
        cit = np.zeros((n, cap), dtype=np.int32)
        stats = (_lib.Stats * n)()
        for i in range(n):
            stats[i].history_cap = cap
            stats[i].history = hist[i].ctypes.data_as(C.POINTER(C.c_double))
            stats[i].coarse_iters = cit[i].ctypes.data_as(C.POINTER(C.c_int32))
        _lib.check(_lib.lib().pepsmg_batch_solve_device(self.h, ps, pb, maxiteration, tolerance, omega_sor,
                                                        1 if use_graph else 0, C.c_void_p(stream) if stream else None,
                                                        stats))
        return [SolveResult(stats[i], hist[i], cit[i]) for i in range(n)]

    def destroy(self):
        if getattr(self, "h", None):
            _lib.lib().pepsmg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
