"""Differentiable solve (peps_mg_amd.adjoint): gradients of u = A^-1 b.

The CPU reference gradient is built from an independent sparse matrix
(tests/refmath.py Kronecker-sum operator) with scipy's direct solver:
dL/db = A^-T dL/du.  Tolerances (written in each test) follow from the
solver tolerance: relative residual 1e-12 gives about 1e-10 relative error
in u on these grids; central differences of a quadratic loss have no
truncation error, so their accuracy is rounding of the loss / (2 eps).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla
import torch

import refmath as RM
from peps_mg_amd import _lib, geometry


def grid(n, stretch, lz=1.0):
    gd = geometry.geometry_domain_create(n, n, n, 0.0, 0.0, 0.0, 1.0, 1.0, lz, stretch, stretch, stretch)
    return geometry.geometry_subdomain_create(gd)


def ref_matrix(sd):
    return RM.laplacian(sd.dxm[1:-1], sd.dym[1:-1], sd.dzm[1:-1]).tocsc()


@pytest.mark.parametrize("stretch", [1.0, 1.2])
def test_volume_weighting_symmetrises_operator(stretch):
    """V A is symmetric (V = cell volumes), the identity behind
    A^-T g = V A^-1 V^-1 g; A itself is symmetric only on uniform axes."""
    sd = grid(8, stretch, lz=2.0)
    A = ref_matrix(sd)
    v = (sd.dzm[1:-1, None, None] * sd.dym[None, 1:-1, None] * sd.dxm[None, None, 1:-1]).ravel()
    VA = (A.T.multiply(v).T).toarray()          # rows scaled: diag(v) @ A
    assert np.allclose(VA, VA.T, rtol=1e-13, atol=1e-9)
    if stretch == 1.0:
        assert np.allclose(A.toarray(), A.toarray().T, rtol=1e-13)
    else:
        assert not np.allclose(A.toarray(), A.toarray().T, rtol=1e-6)
    # the adjoint formula against a dense transpose solve
    rng = np.random.default_rng(3)
    g = rng.normal(size=v.size)
    lam_formula = v * spla.spsolve(A, g / v)
    lam_dense = np.linalg.solve(A.toarray().T, g)
    assert np.allclose(lam_formula, lam_dense, rtol=1e-10, atol=1e-12)


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU error path")
def test_solver_requires_gpu():
    from peps_mg_amd.adjoint import PoissonSolver
    with pytest.raises(_lib.PepsmgError):
        PoissonSolver.uniform(8, 2)


@pytest.mark.gpu
@pytest.mark.parametrize("n,stretch,lz,precision", [(16, 1.0, 1.0, 0), (16, 1.2, 1.0, 0), (16, 1.0, 2.0, 0),
                                                     (16, 1.2, 1.0, 1)])
def test_gradient_matches_cpu_reference(n, stretch, lz, precision):
    from peps_mg_amd.adjoint import PoissonSolver, poisson_solve
    sd = grid(n, stretch, lz)
    S = PoissonSolver(sd, 2, tolerance=1e-12, precision=precision)
    rng = np.random.default_rng(7 + precision)
    f = rng.uniform(-1, 1, (n, n, n))
    uobs = rng.normal(scale=1e-2, size=(n, n, n))
    A = ref_matrix(sd)
    u_ref = spla.spsolve(A, f.ravel())
    g_ref = spla.spsolve(A.T.tocsc(), u_ref - uobs.ravel())

    b = torch.tensor(f, device="cuda", requires_grad=True)
    u = poisson_solve(b, S)
    loss = 0.5 * ((u - torch.tensor(uobs, device="cuda")) ** 2).sum()
    loss.backward()
    u_np = u.detach().cpu().numpy().ravel()
    g_np = b.grad.cpu().numpy().ravel()
    # solver tolerance 1e-12 -> relative error of u and of the gradient well below 1e-8
    assert np.linalg.norm(u_np - u_ref) <= 1e-8 * np.linalg.norm(u_ref)
    assert np.linalg.norm(g_np - g_ref) <= 1e-8 * np.linalg.norm(g_ref)
    S.destroy()


@pytest.mark.gpu
def test_gradient_central_differences():
    """32 random directions, eps = 1e-6: |FD - grad.d| <= 1e-5 |grad| |d|."""
    from peps_mg_amd.adjoint import PoissonSolver, poisson_solve
    n = 16
    sd = grid(n, 1.2)
    S = PoissonSolver(sd, 2, tolerance=1e-13)
    rng = np.random.default_rng(11)
    uobs = torch.tensor(rng.normal(scale=1e-2, size=(n, n, n)), device="cuda")
    f = torch.tensor(rng.uniform(-1, 1, (n, n, n)), device="cuda", requires_grad=True)

    def loss(b):
        return 0.5 * ((poisson_solve(b, S) - uobs) ** 2).sum()

    loss(f).backward()
    g = f.grad.detach()
    eps = 1e-6
    with torch.no_grad():
        for _ in range(32):
            d = torch.tensor(rng.normal(size=(n, n, n)), device="cuda")
            d /= d.norm()
            fd = (loss(f + eps * d) - loss(f - eps * d)).item() / (2 * eps)
            an = (g * d).sum().item()
            assert abs(fd - an) <= 1e-5 * g.norm().item(), (fd, an)
    S.destroy()


@pytest.mark.gpu
@pytest.mark.parametrize("stretch", [1.0, 1.25])
def test_adjoint_dot_product_identity(stretch):
    """<A^-1 b, g> = <b, A^-T g> on a 32^3 grid, through the GPU solves only."""
    from peps_mg_amd.adjoint import PoissonSolver
    n = 32
    S = PoissonSolver(grid(n, stretch), 3, tolerance=1e-13)
    gen = torch.Generator(device="cuda").manual_seed(5)
    b = torch.rand((n, n, n), generator=gen, device="cuda", dtype=torch.float64) - 0.5
    g = torch.rand((n, n, n), generator=gen, device="cuda", dtype=torch.float64) - 0.5
    lhs = (S.solve(b) * g).sum().item()
    rhs = (b * S.solve_adjoint(g)).sum().item()
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    S.destroy()


def inverse_source_problem(n, seed=20210501):
    """The adjoint workload's data: f* = six Gaussians (centres U[0.2,0.8]^3,
    widths U[0.03,0.1], amplitudes U[-1,1]), N(0, 1e-3) noise, a random
    point f0 of U[-1,1]."""
    rng = np.random.default_rng(seed)
    c = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    fstar = np.zeros((n, n, n))
    for _ in range(6):
        cx, cy, cz = rng.uniform(0.2, 0.8, 3)
        w = rng.uniform(0.03, 0.1)
        fstar += rng.uniform(-1, 1) * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2) / (2 * w * w))
    return fstar, rng.normal(scale=1e-3, size=(n, n, n)), rng.uniform(-1, 1, (n, n, n)), rng


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_adjoint_128_inverse_source(precision):
    """The inverse-source workload at 128^3 (loss 0.5 |u(f) - u_obs|^2 +
    1e-4 |grad f|^2, A u = -f): the gradient of the data term, which is the
    adjoint solve's output, against a CPU FP64 reference built from the
    oracle's multigrid solves (forward and adjoint, to 1e-13; A is symmetric
    on this uniform grid); and central differences (eps 1e-6) of the whole
    loss over 32 random unit directions.  The GPU solves run to 1e-12, which
    leaves about 1e-11 relative error in u and in the gradient; the bound
    is 1e-9."""
    from oracle import oracle as O
    from peps_mg_amd.adjoint import PoissonSolver, poisson_solve
    n, levels = 128, 5
    h = 1.0 / n
    fstar, noise, f0, rng = inverse_source_problem(n)
    g = O.Grid(n, n, n)
    A = g.matrix()

    def cpu_solve(rhs_interior):
        b = O.padded(n, n, n)
        b[1:-1, 1:-1, 1:-1] = rhs_interior
        x = O.padded(n, n, n)
        O.Multigrid(g, levels, 200).solve_vcycle(x, O.padded(n, n, n), A, b, 2, 1e-13, 1.0)
        return x[1:-1, 1:-1, 1:-1].copy()

    uobs = cpu_solve(-fstar) + noise
    grad_ref = -cpu_solve(cpu_solve(-f0) - uobs)

    S = PoissonSolver.uniform(n, levels, tolerance=1e-12, precision=precision)
    uobs_t = torch.tensor(uobs, device="cuda")
    f = torch.tensor(f0, device="cuda", requires_grad=True)
    (0.5 * ((poisson_solve(-f, S) - uobs_t) ** 2).sum()).backward()
    gd = f.grad.detach().cpu().numpy()
    err = np.linalg.norm(gd - grad_ref) / np.linalg.norm(grad_ref)
    assert err <= 1e-9, err

    def loss(x):
        reg = sum(((torch.diff(x, dim=d) / h) ** 2).sum() for d in range(3))
        return 0.5 * ((poisson_solve(-x, S) - uobs_t) ** 2).sum() + 1e-4 * reg

    f.grad = None
    loss(f).backward()
    gr = f.grad.detach()
    eps = 1e-6
    with torch.no_grad():
        for _ in range(32):
            d = torch.tensor(rng.normal(size=(n, n, n)), device="cuda")
            d /= d.norm()
            fd = (loss(f + eps * d) - loss(f - eps * d)).item() / (2 * eps)
            assert abs(fd - (gr * d).sum().item()) <= 1e-5 * gr.norm().item()
    S.destroy()
