"""Differentiable Poisson solve: u = A^-1 b as a torch.autograd.Function.

A is the reference's cell-centred 7-point operator (matrix_heptadiagonal,
src/matrix.f90:78-88, homogeneous Dirichlet), solved on the GPU by the
multigrid V-cycle of libpepsmg (multigrid_solve_vcycle,
src/multigrid.f90:1864-1976) to a relative residual `tolerance`.

Backward.  For a loss L(u) with upstream gradient g = dL/du,
dL/db = A^-T g.  The finite-volume operator satisfies V A = (V A)^T with
V = diag(dxm_i dym_j dzm_k) the cell volumes (the coupling of cells i and
i+1 is dy dz / dxg_{i+1} seen from either side), so A^T = V A V^-1 and
A^-T g = V A^-1 (V^-1 g): the backward is a second multigrid solve with the
same hierarchy.  On grids uniform along every axis V is a multiple of the
identity and the scaling is skipped.

Tensors are interior fields of shape (nz, ny, nx), float64, on the GPU; the
solver stages them in zero-ghost padded buffers it owns.  The reference has
no derivative code; this module is an addition of this implementation.
"""
import numpy as np
import torch

from . import _lib
from .geometry import geometry_domain_create, geometry_subdomain_create
from .multigrid import Multigrid


class PoissonSolver:
    """Owns a multigrid hierarchy and padded device buffers for one grid.

    Solves A u = b from u0 = 0 to sqrt(|b - A u|^2 / |b|^2) < tolerance with
    V(maxiteration, maxiteration) cycles (precision 0: FP64 cycles; 1: FP32
    cycles inside FP64 refinement).
    """

    def __init__(self, sdm, nlevel, tolerance=1e-10, maxiteration=2, omega=1.0, ncycle=200, device="cuda",
                 **options):
        if not torch.cuda.is_available():
            raise _lib.PepsmgError(_lib.ERR_HIP, "no GPU: the Poisson solve runs only through libpepsmg on HIP")
        self.sdm = sdm
        self.mg = Multigrid(sdm, nlevel, ncycle, **options)
        self.tol, self.nu, self.omega = float(tolerance), int(maxiteration), float(omega)
        self.device = torch.device(device)
        shape = (sdm.nz + 2, sdm.ny + 2, sdm.nx + 2)
        self._x = torch.zeros(shape, dtype=torch.float64, device=self.device)
        self._b = torch.zeros(shape, dtype=torch.float64, device=self.device)
        self.last = None
        uni = all(np.all(d[1:-1] == d[1]) for d in (sdm.dxm, sdm.dym, sdm.dzm))
        if uni:
            self._vol = None
        else:
            v = (sdm.dzm[1:-1, None, None] * sdm.dym[None, 1:-1, None] * sdm.dxm[None, None, 1:-1])
            self._vol = torch.from_numpy(np.ascontiguousarray(v)).to(self.device)

    @classmethod
    def uniform(cls, n, nlevel, lengths=(1.0, 1.0, 1.0), **kw):
        gd = geometry_domain_create(n, n, n, 0.0, 0.0, 0.0, *lengths)
        return cls(geometry_subdomain_create(gd), nlevel, **kw)

    @property
    def shape(self):
        return (self.sdm.nz, self.sdm.ny, self.sdm.nx)

    def solve(self, b):
        """u = A^-1 b (interior tensors)."""
        if b.shape != self.shape or b.dtype != torch.float64 or b.device.type != "cuda":
            raise _lib.PepsmgError(_lib.ERR_ARG, f"b must be float64 {self.shape} on the GPU")
        if not bool(torch.any(b != 0)):
            # the relative-residual stopping test is undefined for b = 0 (and u = 0 exactly)
            return torch.zeros_like(b)
        self._b[1:-1, 1:-1, 1:-1].copy_(b)
        self._x.zero_()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        self.last = self.mg.solve_vcycle_device(self._x.data_ptr(), 0, self._b.data_ptr(), self.nu, self.tol,
                                                self.omega, stream=stream)
        if not self.last.converged:
            raise _lib.PepsmgError(_lib.ERR_NOTCONV, f"multigrid stopped at rho={self.last.rho:.3e}")
        return self._x[1:-1, 1:-1, 1:-1].clone()

    def solve_adjoint(self, g):
        """A^-T g = V A^-1 (V^-1 g)."""
        if self._vol is None:
            return self.solve(g)
        return self._vol * self.solve(g / self._vol)

    def destroy(self):
        self.mg.destroy()


class PoissonSolve(torch.autograd.Function):
    """u = A^-1 b; backward: grad_b = A^-T grad_u (a second multigrid solve)."""

    @staticmethod
    def forward(ctx, b, solver):
        ctx.solver = solver
        return solver.solve(b.detach())

    @staticmethod
    def backward(ctx, grad_u):
        return ctx.solver.solve_adjoint(grad_u.contiguous()), None


def poisson_solve(b, solver):
    return PoissonSolve.apply(b, solver)
