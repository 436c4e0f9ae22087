"""GPU parity: libpepsmg.so (through its C-ABI) against the oracle.

FP64 smoother, residual, transfer and V-cycle values are bit-identical to
the oracle's (same expressions, no FMA contraction; the coarsest solve sums
its norms in the reference's order).  Dot products use a tree order, so
norms and the per-cycle relative residuals agree to rounding (|drho| <= 1e-12).
"""
import ctypes as C

import numpy as np
import pytest

import refmath as RM
from oracle import oracle as O
from peps_mg_amd import _lib, geometry, matrix
from peps_mg_amd.cg_poisson_matrix import cg_iterator_poisson_matrix, cg_solver_poisson_matrix
from peps_mg_amd.multigrid import Multigrid
from peps_mg_amd.poisson_matrix_operator import mv_mul_poisson_matrix, vv_dot_3d_matrix
from peps_mg_amd.rbgs_poisson_matrix import (rbgs_iterator_fp32, rbgs_iterator_poisson_matrix,
                                             rbgs_solver_poisson_matrix)

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(99)


def sdm_from(g):
    """Subdomain carrying exactly the oracle grid's metrics."""
    s = geometry.Subdomain(g.nx, g.ny, g.nz, 0, 0, 0, g.sox, g.soy, g.soz, g.dxm, g.dym, g.dzm, g.dxg, g.dyg,
                           g.dzg, g.xg, g.yg, g.zg)
    return s


def rand_padded(nx, ny, nz):
    a = O.padded(nx, ny, nz)
    a[1:-1, 1:-1, 1:-1] = RNG.uniform(-1, 1, (nz, ny, nx))
    return a


CASES = [((16, 16, 16), (1, 1, 1), (1, 1, 1)), ((12, 10, 9), (1, 2, 1), (1, 1, 1)),
         ((16, 8, 16), (1, 1, 1), (1.15, 1.0, 1.2)), ((5, 7, 3), (1, 1, 1), (1, 1, 1)),
         # one cell along an axis: a tile holds one cell, the pair (1, 2) reaches past nx, a
         # z-chunk holds one plane
         ((1, 1, 1), (1, 1, 1), (1, 1, 1)), ((1, 6, 3), (1, 1, 1), (1, 1, 1)),
         ((7, 1, 2), (1, 1, 1), (1, 1, 1)), ((2, 3, 1), (1, 1, 1), (1, 1, 1))]


@pytest.mark.parametrize("dims,lens,stretch", CASES)
def test_mv_mul_bitwise(dims, lens, stretch):
    g = O.Grid(*dims, l=lens, a=stretch)
    A = g.matrix()
    x = rand_padded(*dims)
    y = O.padded(*dims)
    y[0, 0, 0] = 5.0                                    # ghosts are overwritten with 0
    mv_mul_poisson_matrix(y, matrix.MatrixHeptadiagonal(0, A), x, sdm_from(g))
    assert np.array_equal(y, O.mv_mul(A, x))


def test_vv_dot():
    x, y = rand_padded(33, 17, 9), rand_padded(33, 17, 9)
    v = vv_dot_3d_matrix(x, y, 33, 17, 9)
    assert abs(v - O.vv_dot(x, y)) <= 1e-13 * abs(O.vv_dot(np.abs(x), np.abs(y)))


@pytest.mark.parametrize("dims,lens,stretch", CASES)
@pytest.mark.parametrize("omega,nu", [(1.0, 1), (1.6, 3)])
def test_rbgs_iterator_bitwise(dims, lens, stretch, omega, nu):
    g = O.Grid(*dims, l=lens, a=stretch)
    A = g.matrix()
    x, b = rand_padded(*dims), rand_padded(*dims)
    xg = x.copy()
    rbgs_iterator_poisson_matrix(xg, matrix.MatrixHeptadiagonal(0, A), b, sdm_from(g), nu, omega)
    O.rbgs_iterator(x, A, b, nu, omega)
    assert np.array_equal(xg, x)


@pytest.mark.parametrize("nu", [2, 4])
def test_rbgs_iterator_alternating_bitwise(nu):
    """Consecutive iterations march their z-chunks upward, then downward
    (k_rb_fused_down), on the 64 x 32 tiles of levels with 256 rows or more:
    partial tiles in x and y, and an odd nz, so the top chunk ends on an
    even plane and the downward march starts one plane above the domain.
    Only x and y are stretched: the stretched-axis formula sets cells 1 ..
    n/2 and their mirrors (src/geometry.f90:190-197), so an odd count leaves
    the middle cell's width unset (0 in the oracle: infinite coefficients)."""
    dims = (70, 260, 37)
    g = O.Grid(*dims, a=(1.1, 1.05, 1.0))
    A = g.matrix()
    x, b = rand_padded(*dims), rand_padded(*dims)
    xg = x.copy()
    rbgs_iterator_poisson_matrix(xg, matrix.MatrixHeptadiagonal(0, A), b, sdm_from(g), nu, 1.6)
    O.rbgs_iterator(x, A, b, nu, 1.6)
    assert np.array_equal(xg, x)


FP32_CASES = [((70, 260, 37), (1.1, 1.05, 1.0)), ((128, 128, 33), (1.0, 1.0, 1.0)),
              ((64, 64, 64), (1.15, 1.0, 1.2)), ((20, 12, 9), (1.0, 1.0, 1.0)), ((1, 6, 3), (1.0, 1.0, 1.0))]


@pytest.mark.parametrize("dims,stretch", FP32_CASES,
                         ids=["128x32-partial", "128x32-whole", "64x8", "64x8-ragged", "one-column"])
@pytest.mark.parametrize("omega,nu", [(1.0, 1), (1.35, 2), (1.6, 3)])
def test_rbgs_iterator_fp32_against_oracle(dims, stretch, omega, nu):
    """The FP32 smoother of the mixed-precision path, launched as the solver
    launches it (128 x 32 tiles on levels of 256 rows, or of 128 rows of
    whole tiles; 64 x 8 below; alternating upward and downward marches),
    against the oracle's FP64 iteration from the same FP32-representable
    inputs.  A numpy FP32 emulation of the reference loop lands within 1.5e-6
    (relative to max|x|) of the FP64 iteration on these grids for up to 6
    iterations; the bound 1e-6 (1 + nu) leaves a margin of about 7, while a
    wrong neighbour, colour or coefficient shows at 1e-2 or more.  The outer
    FP64 refinement of a mixed solve would hide such an error; this test
    does not."""
    g = O.Grid(*dims, a=stretch)
    A = g.matrix()
    x = rand_padded(*dims).astype(np.float32).astype(np.float64)
    b = rand_padded(*dims).astype(np.float32).astype(np.float64)
    x32 = x.copy()
    rbgs_iterator_fp32(x32, b, sdm_from(g), nu, omega)
    O.rbgs_iterator(x, A, b, nu, omega)
    assert np.array_equal(x32, x32.astype(np.float32).astype(np.float64))   # FP32 values, widened
    assert not x32[0].any() and not x32[-1].any() and not x32[:, :, 0].any() and not x32[:, :, -1].any()
    err = np.abs(x32 - x).max() / np.abs(x).max()
    assert err < 1e-6 * (1 + nu), err


def test_rbgs_solver_matches():
    dims = (8, 8, 8)
    g = O.Grid(*dims)
    A = g.matrix()
    b = rand_padded(*dims)
    xg, xo = O.padded(*dims), O.padded(*dims)
    it, rel = rbgs_solver_poisson_matrix(xg, None, b, sdm_from(g), 1000, 1e-10, 1.0)
    ito, relo = O.rbgs_solver(xo, A, b, 1000, 1e-10, 1.0)
    assert it == ito
    assert np.array_equal(xg, xo)
    assert abs(rel - relo) <= 1e-14


def test_cg_iterator_and_solver():
    dims = (12, 10, 8)
    g = O.Grid(*dims)
    A = g.matrix()
    x, b = rand_padded(*dims), rand_padded(*dims)
    xg = x.copy()
    cg_iterator_poisson_matrix(xg, None, b, sdm_from(g), 6)
    O.cg_iterator(x, A, b, 6)
    np.testing.assert_allclose(xg, x, rtol=1e-11, atol=1e-13 * np.abs(x).max())
    xg, xo = O.padded(*dims), O.padded(*dims)
    it, rel = cg_solver_poisson_matrix(xg, None, b, sdm_from(g), 400, 1e-9)
    ito, relo = O.cg_solver(xo, A, b, 400, 1e-9)
    assert abs(it - ito) <= 1 and rel < 1e-9
    np.testing.assert_allclose(xg, xo, rtol=1e-8, atol=1e-9 * np.abs(xo).max())


VCASES = [  # dims, lens, stretch, nlevel, nu, omega
    ((32, 32, 32), (1, 1, 1), (1, 1, 1), 3, 2, 1.0),
    ((64, 64, 64), (1, 1, 1), (1, 1, 1), 4, 2, 1.0),          # config 0 shape
    ((48, 48, 48), (1, 1, 1), (1, 1, 1), 4, 2, 1.0),          # coarsest 3^3
    ((32, 32, 32), (1, 1, 2), (1, 1, 1), 3, 2, 1.0),          # hx:hy:hz = 1:1:2
    ((32, 16, 8), (2, 1, 0.5), (1, 1, 1), 4, 1, 1.2),         # axes stop coarsening at different levels
    ((32, 32, 32), (1, 1, 1), (1.1, 1.0, 1.05), 3, 2, 1.0),   # stretched
    ((16, 16, 16), (1, 1, 1), (1, 1, 1), 4, 2, 1.0),          # coarsest 1^3, direct
]


def run_both(dims, lens, stretch, nlevel, nu, omega, ncycle=60, smoother=0, seed=20210501, **opts):
    g = O.Grid(*dims, l=lens, a=stretch)
    A = g.matrix()
    b = -O.uniform_field(*dims, seed)
    mo = O.Multigrid(g, nlevel, ncycle, smoother=smoother)
    xo, ro = O.padded(*dims), O.padded(*dims)
    nc, hist, cit, res0 = mo.solve_vcycle(xo, ro, A, b, nu, 1e-10, omega)
    mg = Multigrid(sdm_from(g), nlevel, ncycle, smoother=smoother, **opts)
    xg, rg = O.padded(*dims), O.padded(*dims)
    res = mg.solve_vcycle(xg, rg, matrix.MatrixHeptadiagonal(0, A), b, nu, 1e-10, omega)
    return (xo, ro, nc, hist, cit, res0), (xg, rg, res)


SMOOTHER_PATHS = [dict(fused_smoother=1, tile_y=8), dict(fused_smoother=1, tile_y=16), dict(fused_smoother=2, tile_y=8),
                  dict(fused_smoother=2, tile_y=16), dict(fused_smoother=1, tile_y=32), dict(fused_smoother=0),
                  dict(fused_smoother=1, coarse_switch=0), dict(fused_smoother=1, coarse_switch=100000)]


@pytest.mark.parametrize("opts", SMOOTHER_PATHS, ids=["fused8", "fused16", "fused8-pf2", "fused16-pf2", "fused32",
                                                      "colour", "nopersist", "persist-big"])
@pytest.mark.parametrize("dims,lens,stretch,nlevel,nu,omega", VCASES)
def test_vcycle_fp64_bitwise(dims, lens, stretch, nlevel, nu, omega, opts):
    (xo, ro, nc, hist, cit, res0), (xg, rg, res) = run_both(dims, lens, stretch, nlevel, nu, omega, **opts)
    assert res.cycles == nc and res.converged
    assert np.max(np.abs(res.history - hist)) <= 1e-12
    assert abs(res.res0 - res0) <= 1e-13 * res0
    assert np.array_equal(res.coarse_iters, cit)
    assert np.array_equal(xg, xo), np.abs(xg - xo).max()
    assert np.array_equal(rg, ro)


def test_stretched_tall_level_fp64_bitwise_and_mixed():
    """Every axis stretched on a grid whose finest level has 512 rows: the
    smoother's 64 x 32 tiles (partial in x on the next level) and the
    refinement kernels' 32-row tiles, with non-uniform weights in every
    kernel.  The lengths follow the cell counts and each stretch ratio keeps
    the widest cell of an axis under twice the narrowest, so that the
    point-smoothed V-cycle converges (in the unit cube, or with a ratio of
    1.05 over 256 cells, the reference's cycle stalls).  FP64 equals the
    oracle bit for bit; mixed precision lands within 1e-10 (max, relative)
    of the oracle's FP64 solution, both solved to 1e-12 (the right-hand side
    is random)."""
    dims, lens, stretch, nl = (64, 512, 32), (0.125, 1.0, 0.0625), (1.03, 1.003, 1.04), 5
    (xo, ro, nc, hist, cit, res0), (xg, rg, res) = run_both(dims, lens, stretch, nl, 2, 1.0)
    assert res.cycles == nc and res.converged
    assert np.max(np.abs(res.history - hist)) <= 1e-12
    assert np.array_equal(res.coarse_iters, cit)
    assert np.array_equal(xg, xo) and np.array_equal(rg, ro)
    g = O.Grid(*dims, l=lens, a=stretch)
    A = g.matrix()
    b = -O.uniform_field(*dims, 20210501)
    xo, ro = O.padded(*dims), O.padded(*dims)
    assert O.Multigrid(g, nl, 80).solve_vcycle(xo, ro, A, b, 2, 1e-12, 1.0)[0] < 80
    mg = Multigrid(sdm_from(g), nl, 80, precision=1)
    xm, rm = O.padded(*dims), O.padded(*dims)
    res = mg.solve_vcycle(xm, rm, None, b, 2, 1e-12, 1.0)
    assert res.converged and res.rho < 1e-12
    err = np.abs(xm - xo).max() / np.abs(xo).max()
    assert err <= 1e-10, err


def test_vcycle_cg_smoother():
    (xo, ro, nc, hist, cit, res0), (xg, rg, res) = run_both((32, 32, 32), (1, 1, 1), (1, 1, 1), 3, 4, 1.0,
                                                            smoother=1)
    assert res.cycles == nc and res.converged
    assert np.max(np.abs(res.history - hist)) <= 1e-12
    assert np.max(np.abs(xg - xo)) <= 1e-10 * np.abs(xo).max()


def workload_rhs(g, seed):
    """b = -(U[-1,1] + 3 pi^2 sin(pi x) sin(pi y) sin(pi z)), the workload's
    right-hand side (U: the SplitMix64 field of the seed)."""
    b = O.uniform_field(g.nx, g.ny, g.nz, seed)
    s = (np.sin(np.pi * g.xg)[None, None, 1:-1] * np.sin(np.pi * g.yg)[None, 1:-1, None] *
         np.sin(np.pi * g.zg)[1:-1, None, None])
    b[1:-1, 1:-1, 1:-1] += 3 * np.pi ** 2 * s
    return -b


@pytest.mark.parametrize("inner", [dict(), dict(inner_nu=1, inner_omega=0.0, inner_cycles=4),
                                   dict(inner_nu=2, inner_nu_post=1), dict(inner_nu=1, inner_nu_post=3, inner_cycles=3)],
                         ids=["default", "nu1-omega1", "v21", "v13"])
@pytest.mark.parametrize("tx,ty", [(0, 0), (64, 0), (128, 32), (64, 32), (64, 16)])
@pytest.mark.parametrize("dims,lens", [((64, 64, 64), (1, 1, 1)), ((32, 32, 32), (1, 1, 2))])
@pytest.mark.parametrize("rhs,tol", [("workload", 1e-10), ("random", 1e-12)])
def test_mixed_precision_vs_fp64_oracle(dims, lens, tx, ty, inner, rhs, tol):
    """Mixed precision within 1e-10 (max, relative) of the oracle's FP64
    solution at the same tolerance.  A solution with relative residual below
    tol differs from the exact one by at most 4e-11 (relative) on these grids
    for the workload's right-hand side at tol 1e-10, and for a purely random
    one at tol 1e-12 (at 1e-10 a random right-hand side's solution is only
    ~1e-9 accurate, so that pair is tested at 1e-12); both figures measured
    with the oracle against 1e-14 solves."""
    g = O.Grid(*dims, l=lens)
    A = g.matrix()
    b = workload_rhs(g, 7) if rhs == "workload" else -O.uniform_field(*dims, 7)
    nl = int(np.log2(dims[0] // 4))
    mo = O.Multigrid(g, nl, 80)
    xo, ro = O.padded(*dims), O.padded(*dims)
    mo.solve_vcycle(xo, ro, A, b, 2, tol, 1.0)
    mg = Multigrid(sdm_from(g), nl, 80, precision=1, tile_x=tx, tile_y=ty, **inner)
    xg, rg = O.padded(*dims), O.padded(*dims)
    res = mg.solve_vcycle(xg, rg, None, b, 2, tol, 1.0)
    assert res.converged and res.rho < tol
    # the true FP64 residual of the returned solution
    K = RM.laplacian(g.dxm[1:-1], g.dym[1:-1], g.dzm[1:-1])
    r = RM.flat(b) - K @ RM.flat(xg)
    assert np.linalg.norm(r) / np.sqrt(res.res0) < 1.01 * tol
    err = np.abs(xg - xo).max() / np.abs(xo).max()
    assert err <= 1e-10, err


def _refinement_sweep(dims, lens, nlevel, tols):
    """Mixed-precision solves of one grid over the tolerances `tols`; checks
    sol, rsd and rho of each against an independent FP64 evaluation of
    b - A x (scipy sparse Laplacian) and returns the set of refinement step
    counts seen."""
    import torch
    g = O.Grid(*dims, l=lens)
    b = -O.uniform_field(*dims, 5)
    K = RM.laplacian(g.dxm[1:-1], g.dym[1:-1], g.dzm[1:-1])
    mg = Multigrid(sdm_from(g), nlevel, 80, precision=1)
    bmax = np.abs(b).max()
    steps = set()
    for tol in tols:
        xg, rg = O.padded(*dims), O.padded(*dims)
        res = mg.solve_vcycle(xg, rg, None, b, 2, tol, 1.0)
        assert res.converged and res.rho < tol
        r = RM.flat(b) - K @ RM.flat(xg)
        assert abs(np.linalg.norm(r) / np.sqrt(res.res0) - res.rho) <= 1e-3 * res.rho + 1e-15
        assert np.abs(RM.flat(rg) - r).max() <= 1e-11 * bmax
        assert not rg[0].any() and not rg[-1].any() and not rg[:, 0].any() and not rg[:, :, -1].any()
        steps.add(res.refine_steps)
        d_b = torch.from_numpy(b).cuda()
        d_x = torch.zeros_like(d_b)
        res_d = mg.solve_vcycle_device(d_x.data_ptr(), d_b.data_ptr(), d_b.data_ptr(), 2, tol, 1.0)
        torch.cuda.synchronize()
        assert res_d.refine_steps == res.refine_steps and res_d.rho == res.rho
        assert np.array_equal(d_x.cpu().numpy(), xg) and np.array_equal(d_b.cpu().numpy(), rg)
    mg.destroy()
    return steps


def test_mixed_refinement_end_states():
    """Mixed precision over a range of tolerances.  The refinement step
    predicted to be the last stores x + e only into sol; when it misses the
    tolerance, x and the FP32 residual are restored from sol and the solve
    goes on.  Across the sweep both outcomes occur; in every case the
    returned sol, rsd (host and device entry points, rsd aliasing rhs on the
    device) and rho agree with an independent FP64 evaluation of b - A x."""
    steps = _refinement_sweep((48, 48, 48), (1, 1, 1), 3, np.geomspace(1e-6, 3e-13, 15))
    assert len(steps) >= 3, steps   # the sweep spans several step counts


@pytest.mark.parametrize("dims", [(64, 512, 32), (80, 528, 24)], ids=["whole-tiles", "partial-tiles"])
def test_mixed_refinement_tall_level(dims):
    """The same checks on levels of 512 rows or more, where the refinement
    kernels march 32-row tiles instead of 16 (80 x 528: the last column and
    row tiles are partial)."""
    nx, ny, nz = dims
    steps = _refinement_sweep(dims, (nx / 512, ny / 512, nz / 512), 3, np.geomspace(1e-5, 1e-12, 5))
    assert len(steps) >= 2, steps


@pytest.mark.parametrize("precision,where", [(0, (0, 3, 3)), (1, (0, 3, 3)), (1, (3, 9, 3)), (1, (3, 3, 9)),
                                             (1, (9, 9, 9))])
def test_ghost_cells_rejected(precision, where):
    dims = (8, 8, 8)
    g = O.Grid(*dims)
    mg = Multigrid(sdm_from(g), 2, 10, precision=precision)
    x, r = O.padded(*dims), O.padded(*dims)
    x[where] = 1.0
    with pytest.raises(_lib.PepsmgError) as e:
        mg.solve_vcycle(x, r, None, rand_padded(*dims), 2, 1e-10, 1.0)
    assert e.value.code == _lib.ERR_GHOST


def test_matrix_mismatch_rejected():
    dims = (8, 8, 8)
    g = O.Grid(*dims)
    A = g.matrix()
    A[2, 2, 2, 3] *= 1.0000001
    mg = Multigrid(sdm_from(g), 2, 10)
    x, r = O.padded(*dims), O.padded(*dims)
    with pytest.raises(_lib.PepsmgError) as e:
        mg.solve_vcycle(x, r, matrix.MatrixHeptadiagonal(0, A), rand_padded(*dims), 2, 1e-10, 1.0)
    assert e.value.code == _lib.ERR_MATRIX


@pytest.mark.parametrize("nl,aggr,cga", [(1, 0, 0), (2, 5, 0), (2, 0, 2), (9, 0, 0), (5, 2, 0)])
def test_create_errors(nl, aggr, cga):
    g = O.Grid(16, 16, 16)
    with pytest.raises(_lib.PepsmgError) as e:
        Multigrid(sdm_from(g), nl, 10, aggr, cga)
    assert e.value.code == _lib.ERR_LEVELS


def test_degenerate_grid_rejected_by_operators():
    """Standalone operators check the metrics too: a stretched z axis with an
    odd cell count has an unset (zero) middle width."""
    g = O.Grid(8, 8, 9, a=(1.0, 1.0, 1.2))
    assert g.dzm[5] == 0.0
    x, b = rand_padded(8, 8, 9), rand_padded(8, 8, 9)
    with pytest.raises(_lib.PepsmgError) as e:
        rbgs_iterator_poisson_matrix(x, None, b, sdm_from(g), 1, 1.0)
    assert e.value.code == _lib.ERR_ARG and "z axis" in str(e.value)


def test_device_api_and_rhs_generator():
    import torch
    n = 32
    g = O.Grid(n, n, n)
    d_b = torch.zeros((n + 2, n + 2, n + 2), dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().pepsmg_fill_rhs_device(n, n, n, C.c_void_p(d_b.data_ptr()), 20210501, 1.0, 0.0,
                                                 None, None, None, None))
    torch.cuda.synchronize()
    assert np.array_equal(d_b.cpu().numpy(), O.uniform_field(n, n, n, 20210501))
    b = d_b.cpu().numpy()
    mg = Multigrid(sdm_from(g), 3, 40)
    d_x = torch.zeros_like(d_b)
    d_r = torch.zeros_like(d_b)
    res_d = mg.solve_vcycle_device(d_x.data_ptr(), d_r.data_ptr(), d_b.data_ptr(), 2, 1e-10, 1.0,
                                   torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    xh, rh = O.padded(n, n, n), O.padded(n, n, n)
    res_h = mg.solve_vcycle(xh, rh, None, b, 2, 1e-10, 1.0)
    assert res_d.cycles == res_h.cycles
    assert np.array_equal(d_x.cpu().numpy(), xh) and np.array_equal(d_r.cpu().numpy(), rh)


@pytest.mark.parametrize("precision", [0, 1])
def test_full_size_256_properties(precision):
    """256^3: convergence and an independent FP64 residual of the returned
    solution computed with torch on the GPU."""
    import torch
    n = 256
    gdm = geometry.geometry_domain_create(n, n, n, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0)
    sdm = geometry.geometry_subdomain_create(gdm)
    d_b = torch.zeros((n + 2,) * 3, dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().pepsmg_fill_rhs_device(n, n, n, C.c_void_p(d_b.data_ptr()), 20210501, -1.0,
                                                 -3 * np.pi ** 2, _lib.dptr(sdm.xg), _lib.dptr(sdm.yg),
                                                 _lib.dptr(sdm.zg), None))
    mg = Multigrid(sdm, 6, 100, precision=precision)
    d_x = torch.zeros_like(d_b)
    res = mg.solve_vcycle_device(d_x.data_ptr(), 0, d_b.data_ptr(), 2, 1e-10, 1.0)
    torch.cuda.synchronize()
    assert res.converged and res.rho < 1e-10
    h = 1.0 / n
    x = d_x
    c = 1.0 / (h * h)
    ax = -6 * c * x[1:-1, 1:-1, 1:-1]
    ax = ax + c * (x[1:-1, 1:-1, :-2] + x[1:-1, 1:-1, 2:] + x[1:-1, :-2, 1:-1] + x[1:-1, 2:, 1:-1] +
                   x[:-2, 1:-1, 1:-1] + x[2:, 1:-1, 1:-1])
    # boundary cells: face coefficient 2/h^2 enters the diagonal only
    ax[:, :, 0] -= c * x[1:-1, 1:-1, 1]
    ax[:, :, -1] -= c * x[1:-1, 1:-1, -2]
    ax[:, 0, :] -= c * x[1:-1, 1, 1:-1]
    ax[:, -1, :] -= c * x[1:-1, -2, 1:-1]
    ax[0, :, :] -= c * x[1, 1:-1, 1:-1]
    ax[-1, :, :] -= c * x[-2, 1:-1, 1:-1]
    r = d_b[1:-1, 1:-1, 1:-1] - ax
    rho = float(torch.sqrt((r * r).sum() / res.res0))
    assert rho < 1.05e-10 and abs(rho - res.rho) < 1e-12


@pytest.mark.parametrize("precision", [0, 1])
def test_device_solve_orders_after_default_stream_work(precision):
    """A device solve with stream NULL (the library's own stream) waits for
    work the caller queued on the default stream: here a long matrix product
    followed by zeroing sol; the result equals a solve from a synchronized
    zero start, bit for bit, on every repetition."""
    import torch
    n = 64
    g = O.Grid(n, n, n)
    sdm = sdm_from(g)
    mg = Multigrid(sdm, 4, 60, precision=precision)
    b = torch.from_numpy(-O.uniform_field(n, n, n, 11)).cuda()
    ref = torch.zeros_like(b)
    torch.cuda.synchronize()
    mg.solve_vcycle_device(ref.data_ptr(), 0, b.data_ptr(), 2, 1e-10, 1.0)
    big = torch.randn(6144, 6144, device="cuda")
    x = torch.zeros_like(b)
    for _ in range(3):
        x.fill_(1.0)
        torch.cuda.synchronize()
        y = big @ big                    # keeps the default stream busy for milliseconds
        x.zero_()
        mg.solve_vcycle_device(x.data_ptr(), 0, b.data_ptr(), 2, 1e-10, 1.0)
        assert torch.equal(x, ref)
        del y
    mg.destroy()


@pytest.mark.parametrize("dims,nlevel", [((256, 256, 128), 4), ((512, 256, 64), 6)])
def test_fp64_decompositions_agree_large(dims, nlevel):
    """On grids where the smoother's one-wave decomposition and the per-level
    tile choice apply, the FP64 solve gives the same solution bit for bit as
    fixed tiles and as the two-sweep colour smoother (every tiling computes
    the same values; only the block layout differs)."""
    import torch
    nx, ny, nz = dims
    g = O.Grid(nx, ny, nz, l=(nx / 256, ny / 256, nz / 256))   # cubic cells
    sdm = sdm_from(g)
    b = torch.from_numpy(-O.uniform_field(nx, ny, nz, 5)).cuda()
    out = []
    for opts in (dict(), dict(tile_y=8), dict(tile_y=32), dict(fused_smoother=0)):
        mg = Multigrid(sdm, nlevel, 60, **opts)
        x = torch.zeros_like(b)
        r = mg.solve_vcycle_device(x.data_ptr(), 0, b.data_ptr(), 2, 1e-10, 1.0)
        torch.cuda.synchronize()
        assert r.converged
        out.append((x, r))
        mg.destroy()
    for x, r in out[1:]:
        assert r.cycles == out[0][1].cycles
        np.testing.assert_allclose(r.history, out[0][1].history, rtol=1e-12, atol=0)
        assert torch.equal(x, out[0][0])


# --- behaviours of the reference's cycle (src/multigrid.f90:1907-1976) -----

def rand_guess(dims, seed):
    """A non-zero initial guess with zero ghosts (homogeneous Dirichlet)."""
    rng = np.random.default_rng(seed)
    x = O.padded(*dims)
    x[1:-1, 1:-1, 1:-1] = rng.uniform(-0.05, 0.05, dims[::-1])
    return x


def solve_pair(dims, lens, nlevel, ncycle, nu, omega, x0, b, tol=1e-10, **opts):
    """The same solve on the oracle and the GPU (host entry point) from x0."""
    g = O.Grid(*dims, l=lens)
    A = g.matrix()
    mo = O.Multigrid(g, nlevel, ncycle)
    xo, ro = x0.copy(), O.padded(*dims)
    nc, hist, cit, res0 = mo.solve_vcycle(xo, ro, A, b, nu, tol, omega)
    mg = Multigrid(sdm_from(g), nlevel, ncycle, **opts)
    xg, rg = x0.copy(), O.padded(*dims)
    res = mg.solve_vcycle(xg, rg, matrix.MatrixHeptadiagonal(0, A), b, nu, tol, omega)
    mg.destroy()
    return (xo, ro, nc, hist, cit, res0), (xg, rg, res)


FP64_PATHS = [dict(), dict(fused_smoother=0), dict(coarse_switch=0)]
FP64_IDS = ["fused", "colour", "per-level-coarse"]


@pytest.mark.parametrize("opts", FP64_PATHS, ids=FP64_IDS)
@pytest.mark.parametrize("dims,lens,nlevel", [((64, 64, 64), (1, 1, 1), 4), ((32, 32, 32), (1, 1, 2), 3)])
def test_nonzero_initial_guess_fp64_bitwise(dims, lens, nlevel, opts):
    """rho is relative to |b - A x0| of the incoming sol (:1924-1925); with
    the fused smoother that norm comes out of the first smoothing pass."""
    g = O.Grid(*dims, l=lens)
    x0 = rand_guess(dims, 4)
    b = workload_rhs(g, 3)
    (xo, ro, nc, hist, cit, res0), (xg, rg, res) = solve_pair(dims, lens, nlevel, 80, 2, 1.0, x0, b, **opts)
    assert res.converged and res.cycles == nc
    assert abs(res.res0 - res0) <= 1e-13 * res0
    assert np.max(np.abs(res.history - hist)) <= 1e-12
    assert np.array_equal(res.coarse_iters, cit)
    assert np.array_equal(xg, xo) and np.array_equal(rg, ro)


@pytest.mark.parametrize("start,tol", [("warm", 1e-10), ("random", 1e-12)])
@pytest.mark.parametrize("dims,lens,nlevel", [((64, 64, 64), (1, 1, 1), 4), ((32, 32, 32), (1, 1, 2), 3)])
def test_nonzero_initial_guess_mixed(dims, lens, nlevel, start, tol):
    """The first refinement step reads the non-zero sol from the caller's
    array (k_refine_start); the result is within 1e-10 of the oracle's FP64
    solve from the same x0, and res0 is the same norm.  Warm start: x0 is the
    oracle's iterate after two V-cycles (a resumed solve).  Random x0: |A x0|
    dwarfs |b|, so a 1e-10 residual relative to it leaves even the oracle's
    FP64 answer ~1e-9 from the exact one; that pair is compared at 1e-12."""
    g = O.Grid(*dims, l=lens)
    b = workload_rhs(g, 3)
    if start == "warm":
        x0 = O.padded(*dims)
        O.Multigrid(g, nlevel, 2).solve_vcycle(x0, O.padded(*dims), g.matrix(), b, 2, 1e-30, 1.0)
    else:
        x0 = rand_guess(dims, 5)
    (xo, ro, nc, hist, cit, res0), (xg, rg, res) = solve_pair(dims, lens, nlevel, 80, 2, 1.0, x0, b, tol=tol,
                                                              precision=1)
    assert res.converged and res.rho < tol
    assert abs(res.res0 - res0) <= 1e-12 * res0
    assert np.abs(xg - xo).max() <= 1e-10 * np.abs(xo).max()


@pytest.mark.parametrize("opts", FP64_PATHS, ids=FP64_IDS)
@pytest.mark.parametrize("ncycle", [1, 3])
def test_cycle_cap_fp64_bitwise(ncycle, opts):
    """number_of_vcycles below what the tolerance needs: the loop ends after
    ncycle cycles (:1928) without meeting rho < tol; converged = 0 and the
    state equals the oracle's after the same cycles."""
    dims = (64, 64, 64)
    g = O.Grid(*dims)
    (xo, ro, nc, hist, cit, res0), (xg, rg, res) = solve_pair(dims, (1, 1, 1), 4, ncycle, 2, 1.0, O.padded(*dims),
                                                              workload_rhs(g, 8), **opts)
    assert nc == ncycle and res.cycles == ncycle and not res.converged and res.rho > 1e-10
    assert np.max(np.abs(res.history - hist)) <= 1e-12
    assert np.array_equal(res.coarse_iters, cit)
    assert np.array_equal(xg, xo) and np.array_equal(rg, ro)


@pytest.mark.parametrize("ncycle,inner", [(3, 2), (5, 2), (4, 3)])
def test_cycle_cap_mixed(ncycle, inner):
    """In mixed precision number_of_vcycles caps the FP32 V-cycles of the
    whole solve: the last refinement step runs only the cycles that are
    left, so exactly ncycle run; the reported rho is the true FP64 relative
    residual of the returned sol."""
    dims = (48, 48, 48)
    g = O.Grid(*dims)
    b = workload_rhs(g, 9)
    mg = Multigrid(sdm_from(g), 4, ncycle, precision=1, inner_cycles=inner)
    xg, rg = O.padded(*dims), O.padded(*dims)
    res = mg.solve_vcycle(xg, rg, None, b, 2, 1e-10, 1.0)
    assert res.cycles == ncycle and not res.converged
    assert res.refine_steps == -(-ncycle // inner)
    K = RM.laplacian(g.dxm[1:-1], g.dym[1:-1], g.dzm[1:-1])
    r = RM.flat(b) - K @ RM.flat(xg)
    assert abs(np.linalg.norm(r) / np.sqrt(res.res0) - res.rho) <= 1e-6 * res.rho
    assert np.abs(RM.flat(rg) - r).max() <= 1e-11 * np.abs(b).max()
    mg.destroy()


@pytest.mark.parametrize("dims,nlevel", [((32, 32, 32), 3), ((16, 16, 16), 4)])
def test_maxiteration_zero_bitwise(dims, nlevel):
    """maxiteration = 0: no smoothing, the cycle is the coarse correction
    alone (the fused smoother's norm fold needs nu > 0, so the norms come
    from separate residual passes)."""
    g = O.Grid(*dims)
    (xo, ro, nc, hist, cit, res0), (xg, rg, res) = solve_pair(dims, (1, 1, 1), nlevel, 6, 0, 1.0, O.padded(*dims),
                                                              workload_rhs(g, 2))
    assert res.cycles == nc
    assert abs(res.res0 - res0) <= 1e-13 * res0
    assert np.max(np.abs(res.history - hist)) <= 1e-12
    assert np.array_equal(res.coarse_iters, cit)
    assert np.array_equal(xg, xo) and np.array_equal(rg, ro)


@pytest.mark.parametrize("opts", [dict(), dict(fused_smoother=0), dict(precision=1)], ids=["fused", "colour", "mixed"])
def test_zero_rhs_returns_zero(opts):
    """b = 0 from x0 = 0: x0 is exact.  The reference would divide 0 by 0
    and run every cycle; here the solve reports converged with rho = 0 and
    returns sol = rsd = 0."""
    import torch
    dims = (32, 32, 32)
    g = O.Grid(*dims)
    mg = Multigrid(sdm_from(g), 3, 50, **opts)
    x, r = O.padded(*dims), O.padded(*dims)
    res = mg.solve_vcycle(x, r, None, O.padded(*dims), 2, 1e-10, 1.0)
    assert res.converged and res.rho == 0.0 and res.res0 == 0.0
    assert not x.any() and not r.any()
    d_b = torch.zeros((34, 34, 34), dtype=torch.float64, device="cuda")
    d_x = torch.zeros_like(d_b)
    res = mg.solve_vcycle_device(d_x.data_ptr(), 0, d_b.data_ptr(), 2, 1e-10, 1.0)
    assert res.converged and res.rho == 0.0 and not d_x.any().item()
    mg.destroy()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("where,value", [("rhs", np.nan), ("rhs", np.inf), ("sol", np.nan)])
def test_nonfinite_input_rejected(precision, where, value):
    """A NaN or infinity in rhs or in the interior of sol makes |b - A x0|
    non-finite: PEPSMG_ERR_ARG, and (device entry point) sol is left as it
    was."""
    import torch
    dims = (32, 32, 32)
    g = O.Grid(*dims)
    mg = Multigrid(sdm_from(g), 3, 50, precision=precision)
    b = workload_rhs(g, 1)
    x = O.padded(*dims)
    (b if where == "rhs" else x)[9, 10, 11] = value
    with pytest.raises(_lib.PepsmgError) as e:
        mg.solve_vcycle(x.copy(), O.padded(*dims), None, b, 2, 1e-10, 1.0)
    assert e.value.code == _lib.ERR_ARG
    d_b = torch.from_numpy(b).cuda()
    d_x = torch.from_numpy(x).cuda()
    with pytest.raises(_lib.PepsmgError) as e:
        mg.solve_vcycle_device(d_x.data_ptr(), 0, d_b.data_ptr(), 2, 1e-10, 1.0)
    assert e.value.code == _lib.ERR_ARG
    got = d_x.cpu().numpy()
    assert np.array_equal(np.isnan(got), np.isnan(x)) and np.array_equal(got[~np.isnan(x)], x[~np.isnan(x)])
    # the context is usable afterwards
    x = O.padded(*dims)
    res = mg.solve_vcycle(x, O.padded(*dims), None, workload_rhs(g, 1), 2, 1e-10, 1.0)
    assert res.converged
    mg.destroy()


def test_solve_rejects_wrong_array_shapes():
    """The host entry point copies (nx+2)(ny+2)(nz+2) values through each
    pointer: interior-only or transposed arrays are refused before any copy."""
    dims = (16, 12, 8)
    g = O.Grid(*dims)
    mg = Multigrid(sdm_from(g), 2, 10)
    good = O.padded(*dims)
    for bad in (np.zeros((8, 12, 16)), np.zeros((18, 14, 10)), np.zeros((10, 14, 18), dtype=np.float32),
                np.zeros((10, 14, 36))[:, :, ::2]):
        with pytest.raises(_lib.PepsmgError) as e:
            mg.solve_vcycle(bad, good.copy(), None, good, 2, 1e-10, 1.0)
        assert e.value.code == _lib.ERR_ARG
    with pytest.raises(_lib.PepsmgError):
        mg.solve_vcycle(good.copy(), good.copy(), matrix.MatrixHeptadiagonal(0, np.zeros((8, 12, 16, 6))), good,
                        2, 1e-10, 1.0)
    mg.destroy()


@pytest.mark.parametrize("precision", [0, 1])
def test_kernel_timing_hook(precision):
    """pepsmg_kernel_timing brackets the fine-level plain smoother launches
    of the following solves with events; timed solves give the same result
    bit for bit, and the kernel named is the one bench_smoother launches."""
    import torch
    n = 64
    g = O.Grid(n, n, n)
    mg = Multigrid(sdm_from(g), 4, 60, precision=precision)
    b = torch.from_numpy(workload_rhs(g, 4)).cuda()
    x0, x1 = torch.zeros_like(b), torch.zeros_like(b)
    r0 = mg.solve_vcycle_device(x0.data_ptr(), 0, b.data_ptr(), 2, 1e-10, 1.0)
    L = _lib.lib()
    _lib.check(L.pepsmg_kernel_timing(mg.h, 1))
    r1 = mg.solve_vcycle_device(x1.data_ptr(), 0, b.data_ptr(), 2, 1e-10, 1.0)
    _lib.check(L.pepsmg_kernel_timing(mg.h, 0))
    tot, cnt, name = C.c_double(0), C.c_int32(0), C.c_char_p()
    _lib.check(L.pepsmg_kernel_timing_read(mg.h, C.byref(tot), C.byref(cnt), C.byref(name)))
    timed_name = name.value
    assert torch.equal(x0, x1) and r0.cycles == r1.cycles
    # FP64 V(2,2): per cycle the pre-smoothing's second iteration and the
    # post-smoothing's (the first pre-smoothing iteration forms the norm and
    # the last one of the final cycle may store into sol: other modes)
    assert cnt.value >= (2 * r1.cycles if precision == 0 else r1.cycles) and tot.value > 0
    avg, nb, kn = C.c_double(0), C.c_double(0), C.c_char_p()
    _lib.check(L.pepsmg_bench_smoother(mg.h, precision, 2, 1.0, C.byref(avg), C.byref(nb), C.byref(kn)))
    assert timed_name == kn.value
    mg.destroy()


@pytest.mark.parametrize("nu", [2, 1])
def test_fp64_caller_array_ends_bitwise(nu):
    """FP64 solves on device arrays read sol and rhs in their first
    smoothing iteration and store the last iteration of a cycle predicted
    (or bound by the cycle cap) to end the solve straight into sol; over a
    range of tolerances and caps the prediction is right, wrong, or never
    made, and every sol and rsd equals the oracle's bit for bit (rsd
    aliasing rhs as the entry point allows)."""
    import torch
    dims = (48, 40, 32)
    g = O.Grid(*dims, l=(1.5, 1.25, 1.0))
    A = g.matrix()
    b = workload_rhs(g, 12)
    for ncycle, tol in [(60, 1e-6), (60, 3e-8), (60, 1e-9), (60, 1e-10), (60, 1e-12), (4, 1e-12), (2, 1e-12)]:
        mo = O.Multigrid(g, 3, ncycle)
        xo, ro = O.padded(*dims), O.padded(*dims)
        nc, hist, cit, res0 = mo.solve_vcycle(xo, ro, A, b, nu, tol, 1.0)
        mg = Multigrid(sdm_from(g), 3, ncycle)
        d_b = torch.from_numpy(b).cuda()
        d_x = torch.zeros_like(d_b)
        res = mg.solve_vcycle_device(d_x.data_ptr(), d_b.data_ptr(), d_b.data_ptr(), nu, tol, 1.0)
        torch.cuda.synchronize()
        assert res.cycles == nc, (ncycle, tol, res.cycles, nc)
        assert np.max(np.abs(res.history - hist)) <= 1e-12
        assert np.array_equal(d_x.cpu().numpy(), xo), (ncycle, tol)
        assert np.array_equal(d_b.cpu().numpy(), ro), (ncycle, tol)
        mg.destroy()


@pytest.mark.parametrize("dims", [(64, 512, 32), (80, 528, 24)], ids=["whole-tiles", "partial-tiles"])
def test_cg_smoother_tall_level(dims):
    """A CG(4) V-cycle solve on a level of 512 rows or more, the second with
    partial tiles in x and y, agrees with the oracle (cycle count, rho to
    1e-12, x to 1e-10; the dot products are summed in another order)."""
    nx, ny, nz = dims
    (xo, ro, nc, hist, cit, res0), (xg, rg, res) = run_both(dims, (nx / 512, ny / 512, nz / 512), (1, 1, 1), 3, 4,
                                                            1.0, smoother=1)
    assert res.converged and res.cycles == nc
    assert np.max(np.abs(res.history - hist)) <= 1e-12
    assert np.abs(xg - xo).max() <= 1e-10 * np.abs(xo).max()
