"""Pins the oracle (oracle/pepsmg_oracle.c) against closed forms, independent
matrix constructions (tests/refmath.py), dense solves and invariants.

The reference ships no tests, golden vectors or fixtures, and its Fortran
sources need MPI and source edits to compile here, so every oracle function
is pinned by the mathematics instead.
"""
import math

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import refmath as RM
from oracle import oracle as O

RNG = np.random.default_rng(1234)


def rand_padded(nx, ny, nz, rng=RNG):
    a = O.padded(nx, ny, nz)
    a[1:-1, 1:-1, 1:-1] = rng.uniform(-1, 1, (nz, ny, nx))
    return a


# ---------------------------------------------------------------- metrics
def test_uniform_metrics_closed_form():
    dm, dg, g = O.axis_create(8, 0.25, 2.0, 1.0)
    h = 2.0 / 8
    assert dm[0] == 0.0 and dm[-1] == 0.0
    assert np.all(dm[1:-1] == h)
    assert dg[1] == h / 2 and dg[-1] == h / 2 and np.all(dg[2:-1] == h)
    np.testing.assert_allclose(g[1:-1], 0.25 + (np.arange(8) + 0.5) * h, rtol=0, atol=1e-15)


def test_stretched_metrics_sum_and_ratio():
    n, l, a = 16, 3.0, 1.1
    dm, dg, g = O.axis_create(n, 0.0, l, a)
    w = dm[1:-1]
    assert math.isclose(w.sum(), l, rel_tol=1e-13)
    np.testing.assert_allclose(w[1:n // 2] / w[:n // 2 - 1], a, rtol=1e-13)
    np.testing.assert_allclose(w, w[::-1], rtol=0, atol=0)
    np.testing.assert_allclose(w, RM.widths_stretched(n, l, a), rtol=1e-13)


def test_coefficients_closed_form():
    g = O.Grid(4, 6, 8, l=(1.0, 3.0, 2.0))
    A = g.matrix()                        # (nz, ny, nx, 7)
    hx, hy, hz = 1.0 / 4, 3.0 / 6, 2.0 / 8
    c = A[3, 2, 1]                        # interior cell
    np.testing.assert_allclose(c[1:], [1 / hx**2] * 2 + [1 / hy**2] * 2 + [1 / hz**2] * 2, rtol=1e-15)
    assert c[0] == -(c[1] + c[2] + c[3] + c[4] + c[5] + c[6])
    c = A[0, 0, 0]                        # corner: face coefficient 2/h^2
    np.testing.assert_allclose(c[[1, 3, 5]], [2 / hx**2, 2 / hy**2, 2 / hz**2], rtol=1e-15)


# ---------------------------------------------------------------- operator
@pytest.mark.parametrize("stretch", [(1.0, 1.0, 1.0), (1.15, 1.0, 1.3)])
def test_mv_mul_matches_kronecker_laplacian(stretch):
    nx, ny, nz = 6, 8, 10
    lens = (1.0, 2.0, 0.5)
    g = O.Grid(nx, ny, nz, l=lens, a=stretch)
    A = g.matrix()
    x = rand_padded(nx, ny, nz)
    y = O.mv_mul(A, x)
    wx, wy, wz = g.dxm[1:-1], g.dym[1:-1], g.dzm[1:-1]
    K = RM.laplacian(wx, wy, wz)
    np.testing.assert_allclose(RM.flat(y), K @ RM.flat(x), rtol=1e-11, atol=1e-9)
    # mv_mul zero-fills the padded array: ghosts stay 0
    assert np.all(y[0] == 0) and np.all(y[:, 0] == 0) and np.all(y[:, :, -1] == 0)


def test_operator_symmetric_negative_definite():
    g = O.Grid(5, 6, 7, l=(1.0, 1.0, 2.0))
    A = g.matrix()
    x, y = rand_padded(5, 6, 7), rand_padded(5, 6, 7)
    lhs = O.vv_dot(O.mv_mul(A, x), y)
    rhs = O.vv_dot(x, O.mv_mul(A, y))
    assert math.isclose(lhs, rhs, rel_tol=1e-12)
    assert O.vv_dot(O.mv_mul(A, x), x) < 0


def test_manufactured_second_order():
    errs = []
    for n in (16, 32):
        g = O.Grid(n, n, n)
        A = g.matrix()
        xc = g.xg
        s = np.sin(np.pi * xc)
        u = O.padded(n, n, n)
        u[1:-1, 1:-1, 1:-1] = (s[None, None, 1:-1] * s[None, 1:-1, None] * s[1:-1, None, None])
        au = O.mv_mul(A, u)
        # A = +laplacian, so A u = -3 pi^2 u for the manufactured mode
        err = np.abs(RM.interior(au) + 3 * np.pi**2 * RM.interior(u)).max()
        errs.append(err)
    ratio = errs[0] / errs[1]
    assert 3.0 < ratio < 5.0, ratio


def test_vv_dot_matches_fsum():
    x, y = rand_padded(7, 5, 9), rand_padded(7, 5, 9)
    x[0, 0, 0] = 123.0                    # ghosts are not part of the product
    ref = math.fsum((RM.interior(x) * RM.interior(y)).reshape(-1))
    assert math.isclose(O.vv_dot(x, y), ref, rel_tol=1e-14)


def test_residual():
    g = O.Grid(6, 5, 4)
    A = g.matrix()
    x, b = rand_padded(6, 5, 4), rand_padded(6, 5, 4)
    r = O.residual(A, x, b)
    K = RM.laplacian(g.dxm[1:-1], g.dym[1:-1], g.dzm[1:-1])
    np.testing.assert_allclose(RM.flat(r), RM.flat(b) - K @ RM.flat(x), rtol=1e-11, atol=1e-9)


# ---------------------------------------------------------------- smoothers
@pytest.mark.parametrize("omega,dims,iters", [(1.0, (6, 6, 6), 1), (1.6, (5, 6, 7), 2),
                                              (0.8, (8, 4, 3), 3)])
def test_rbgs_iterator_matches_sor_splitting(omega, dims, iters):
    nx, ny, nz = dims
    g = O.Grid(nx, ny, nz, l=(1.0, 1.5, 2.0))
    A = g.matrix()
    x, b = rand_padded(nx, ny, nz), rand_padded(nx, ny, nz)
    K = RM.laplacian(g.dxm[1:-1], g.dym[1:-1], g.dzm[1:-1])
    ref = RM.sor_rb(K, RM.flat(b), RM.flat(x), RM.red_mask(nx, ny, nz), omega, iters)
    O.rbgs_iterator(x, A, b, iters, omega)
    np.testing.assert_allclose(RM.flat(x), ref, rtol=1e-12, atol=1e-13)


def test_rbgs_colour_invariants():
    """omega=1: after the first colour the first-colour residual vanishes and
    the other colour is untouched; the colour is 1-based i+j+k odd."""
    n = 6
    g = O.Grid(n, n, n)
    A = g.matrix()
    x, b = rand_padded(n, n, n), rand_padded(n, n, n)
    red = RM.red_mask(n, n, n)
    K = RM.laplacian(g.dxm[1:-1], g.dym[1:-1], g.dzm[1:-1])
    x1 = RM.flat(x).copy()
    x_after = RM.sor_rb(K, RM.flat(b), x1, red, 1.0, 1)
    # the splitting reference itself: black-colour residual is zero after one iteration
    r = RM.flat(b) - K @ x_after
    assert np.abs(r[~red]).max() < 1e-10 * np.abs(RM.flat(b)).max() * 100
    # oracle: run one iteration and compare residual pattern
    O.rbgs_iterator(x, A, b, 1, 1.0)
    r = RM.flat(O.residual(A, x, b))
    assert np.abs(r[~red]).max() < 1e-9
    assert np.abs(r[red]).max() > 1e-3


def test_rbgs_solver_converges_and_stopping_rule():
    n = 4
    g = O.Grid(n, n, n)
    A = g.matrix()
    b = rand_padded(n, n, n)
    K = RM.laplacian(g.dxm[1:-1], g.dym[1:-1], g.dzm[1:-1])
    x = O.padded(n, n, n)
    it, rel = O.rbgs_solver(x, A, b, 1000, 1e-10, 1.0)
    exact = spla.spsolve(K.tocsc(), RM.flat(b))
    np.testing.assert_allclose(RM.flat(x), exact, rtol=1e-8, atol=1e-12)
    # it is the 0-based iteration index at which rel <= tol first held
    red = RM.red_mask(n, n, n)
    xr = np.zeros(n**3)
    r0 = np.dot(RM.flat(b), RM.flat(b))
    rels = []
    for _ in range(it + 1):
        xr = RM.sor_rb(K, RM.flat(b), xr, red, 1.0, 1)
        rr = RM.flat(b) - K @ xr
        rels.append(math.sqrt(np.dot(rr, rr) / r0))
    assert rels[-1] <= 1e-10 < rels[-2]
    assert math.isclose(rel, rels[-1], rel_tol=1e-6)


def test_rbgs_solver_cap_returns_maxiteration():
    n = 4
    g = O.Grid(n, n, n)
    A = g.matrix()
    b = rand_padded(n, n, n)
    x = O.padded(n, n, n)
    it, rel = O.rbgs_solver(x, A, b, 3, 1e-30, 1.0)
    assert it == 3 and rel > 0


def textbook_cg(K, b, x, iters):
    r = b - K @ x
    p = r.copy()
    for _ in range(iters):
        rr = r @ r
        Ap = K @ p
        a = rr / (Ap @ p)
        x = x + a * p
        r = r - a * Ap
        p = r + (r @ r) / rr * p
    return x


def test_cg_iterator_matches_textbook():
    nx, ny, nz = 6, 5, 7
    g = O.Grid(nx, ny, nz)
    A = g.matrix()
    x, b = rand_padded(nx, ny, nz), rand_padded(nx, ny, nz)
    K = RM.laplacian(g.dxm[1:-1], g.dym[1:-1], g.dzm[1:-1])
    ref = textbook_cg(K, RM.flat(b), RM.flat(x), 4)
    O.cg_iterator(x, A, b, 4)
    np.testing.assert_allclose(RM.flat(x), ref, rtol=1e-10, atol=1e-12)


def test_cg_finite_termination():
    """2x2x2 uniform grid: the operator has 4 distinct eigenvalues, so CG is
    exact after 4 iterations."""
    g = O.Grid(2, 2, 2)
    A = g.matrix()
    b = rand_padded(2, 2, 2)
    x = O.padded(2, 2, 2)
    O.cg_iterator(x, A, b, 4)
    K = RM.laplacian(g.dxm[1:-1], g.dym[1:-1], g.dzm[1:-1])
    np.testing.assert_allclose(RM.flat(x), np.linalg.solve(K.toarray(), RM.flat(b)), rtol=1e-10)


def test_cg_solver_stops_below_tol():
    g = O.Grid(6, 6, 6)
    A = g.matrix()
    b = rand_padded(6, 6, 6)
    x = O.padded(6, 6, 6)
    it, rel = O.cg_solver(x, A, b, 500, 1e-9)
    assert rel < 1e-9 and it < 500


# ---------------------------------------------------------------- hierarchy
def test_level_sizes_and_coarsest_levels():
    g = O.Grid(48, 48, 48)
    mg = O.Multigrid(g, 4, 10)
    assert [mg.level_dims(l) for l in range(5)] == [(48,) * 3, (24,) * 3, (12,) * 3, (6,) * 3, (3,) * 3]
    assert mg.lvc() == (4, 4, 4)
    # the reference's sample input shape (run/PARA_INPUT.inp): axes stop at 1
    g = O.Grid(64, 8, 32)
    mg = O.Multigrid(g, 6, 10, aggr=2)
    assert mg.level_dims(6) == (1, 1, 1)
    assert mg.lvc() == (6, 3, 5)
    assert mg.level_dims(4) == (4, 1, 2)


def test_level_metrics_closed_form():
    g = O.Grid(16, 8, 8, l=(2.0, 1.0, 1.0))
    mg = O.Multigrid(g, 3, 10)
    m = mg.level_metrics(2)
    assert np.all(m["dxm"][1:-1] == 0.5) and m["dxm"][0] == 0 and m["dxm"][-1] == 0
    assert m["dxg"][1] == 0.25 and np.all(m["dxg"][2:-1] == 0.5)
    assert np.all(m["dym"][1:-1] == 0.5)


@pytest.mark.parametrize("args,code", [((1, 10, 0, 0), -1), ((2, 10, 3, 0), -2), ((2, 10, 0, 1), -3),
                                       ((2, 10, 1, 0), -4), ((5, 10, 0, 0), -6), ((4, 10, 2, 0), -7),
                                       ((4, 10, 1, 4), -5)])
def test_create_error_codes(args, code):
    g = O.Grid(8, 8, 12)          # 8 -> 4 -> 2 -> 1 ; 12 -> 6 -> 3 (odd)
    with pytest.raises(ValueError, match=f"code {code}"):
        O.Multigrid(g, args[0], args[1], aggr=args[2], cga_level=args[3])


# ---------------------------------------------------------------- transfers
@pytest.mark.parametrize("dims,stretch", [((8, 4, 6), (1.0, 1.0, 1.0)), ((8, 8, 8), (1.2, 1.0, 1.1)),
                                          ((8, 3, 4), (1.0, 1.0, 1.0))])
def test_restriction_matches_kronecker(dims, stretch):
    nx, ny, nz = dims
    g = O.Grid(nx, ny, nz, a=stretch)
    mg = O.Multigrid(g, 1 + 1, 10) if min(nx, nz) >= 4 else None
    mg = O.Multigrid(g, 2, 10)
    r = rand_padded(nx, ny, nz)
    rc = mg.restriction(r, 0)
    co = [d % 2 == 0 for d in dims]
    R = RM.kron3(RM.restrict1d(g.dzm[1:-1], co[2]), RM.restrict1d(g.dym[1:-1], co[1]),
                 RM.restrict1d(g.dxm[1:-1], co[0]))
    np.testing.assert_allclose(RM.flat(rc), R @ RM.flat(r), rtol=1e-13, atol=1e-15)
    assert np.all(rc[0] == 0) and np.all(rc[:, :, -1] == 0)


def test_restriction_exact_on_constants_and_conserves_mass():
    g = O.Grid(8, 8, 8, a=(1.2, 1.1, 1.0))
    mg = O.Multigrid(g, 2, 10)
    ones = O.padded(8, 8, 8)
    ones[1:-1, 1:-1, 1:-1] = 3.5
    rc = mg.restriction(ones, 0)
    np.testing.assert_allclose(RM.interior(rc), 3.5, rtol=1e-15)
    r = rand_padded(8, 8, 8)
    rc = mg.restriction(r, 0)
    mc = mg.level_metrics(1)
    volf = np.einsum("k,j,i->kji", g.dzm[1:-1], g.dym[1:-1], g.dxm[1:-1])
    volc = np.einsum("k,j,i->kji", mc["dzm"][1:-1], mc["dym"][1:-1], mc["dxm"][1:-1])
    assert math.isclose((volf * RM.interior(r)).sum(), (volc * RM.interior(rc)).sum(),
                        rel_tol=1e-12, abs_tol=1e-14)


@pytest.mark.parametrize("dims,stretch", [((8, 4, 6), (1.0, 1.0, 1.0)), ((8, 8, 8), (1.2, 1.0, 1.1)),
                                          ((8, 3, 4), (1.0, 1.0, 1.0))])
def test_prolongation_matches_kronecker(dims, stretch):
    nx, ny, nz = dims
    g = O.Grid(nx, ny, nz, a=stretch)
    mg = O.Multigrid(g, 2, 10)
    cx, cy, cz = mg.level_dims(1)
    xc = rand_padded(cx, cy, cz)
    xf = mg.prolongation(xc, 0)
    co = [d % 2 == 0 for d in dims]
    P = RM.kron3(RM.prolong1d(g.dzm[1:-1], co[2]), RM.prolong1d(g.dym[1:-1], co[1]),
                 RM.prolong1d(g.dxm[1:-1], co[0]))
    np.testing.assert_allclose(RM.flat(xf), P @ RM.flat(xc), rtol=1e-12, atol=1e-14)


def test_prolongation_uniform_weights():
    """Interior fine cells get 3/4 of their parent and 1/4 of the next coarse
    cell per axis; boundary fine cells get 1/2 of the parent."""
    g = O.Grid(8, 8, 8)
    mg = O.Multigrid(g, 2, 10)
    xc = O.padded(4, 4, 4)
    xc[2, 2, 2] = 1.0                     # coarse (2,2,2), 1-based
    xf = mg.prolongation(xc, 0)
    # fine (3,3,3) is child 2*2-1 on each axis: 3/4 per axis
    assert xf[3, 3, 3] == 0.75 ** 3
    assert xf[2, 3, 3] == 0.25 * 0.75 ** 2  # fine 2 (child of coarse 1), neighbour weight 1/4
    xc = O.padded(4, 4, 4)
    xc[1, 1, 1] = 1.0
    xf = mg.prolongation(xc, 0)
    assert xf[1, 1, 1] == 0.5 ** 3        # boundary: linear to the zero face


# ---------------------------------------------------------------- V-cycle
def _problem(nx, ny, nz, lens=(1.0, 1.0, 1.0), stretch=(1.0, 1.0, 1.0), seed=7):
    g = O.Grid(nx, ny, nz, l=lens, a=stretch)
    A = g.matrix()
    b = O.uniform_field(nx, ny, nz, seed)
    return g, A, b


@pytest.mark.parametrize("dims,nlev,nu,omega,lens", [((8, 8, 8), 3, 0, 1.0, (1, 1, 1)),
                                                     ((8, 8, 8), 2, 2, 1.0, (1, 1, 1)),
                                                     ((16, 8, 8), 3, 1, 1.2, (1, 1, 2)),
                                                     ((12, 8, 8), 2, 2, 1.0, (1, 1, 1))])
def test_vcycle_matches_matrix_vcycle(dims, nlev, nu, omega, lens):
    nx, ny, nz = dims
    g, A, b = _problem(nx, ny, nz, lens=lens)
    ncyc = 6
    mg = O.Multigrid(g, nlev, ncyc)
    x = O.padded(nx, ny, nz)
    r = O.padded(nx, ny, nz)
    n, hist, cit, res0 = mg.solve_vcycle(x, r, A, b, nu, 1e-10, omega)
    MM = RM.MatrixMG(g.dxm[1:-1], g.dym[1:-1], g.dzm[1:-1], nlev)
    xr, hr = MM.vcycle(np.zeros(nx * ny * nz), RM.flat(b), nu, 1e-10, omega, ncyc)
    assert n == len(hr)
    np.testing.assert_allclose(hist, hr, rtol=1e-9, atol=1e-13)
    np.testing.assert_allclose(RM.flat(x), xr, rtol=1e-10, atol=1e-12 * np.abs(xr).max())
    # rsd holds the final fine residual
    np.testing.assert_allclose(RM.flat(r), RM.flat(b) - MM.A[0] @ RM.flat(x), rtol=1e-8,
                               atol=1e-9 * np.abs(RM.flat(b)).max())


def test_vcycle_converges_to_dense_solution():
    nx, ny, nz = 16, 16, 16
    g, A, b = _problem(nx, ny, nz)
    mg = O.Multigrid(g, 3, 50)
    x = O.padded(nx, ny, nz)
    r = O.padded(nx, ny, nz)
    n, hist, cit, _ = mg.solve_vcycle(x, r, A, b, 2, 1e-10, 1.0)
    assert hist[-1] < 1e-10 and np.all(np.diff(hist) < 0)
    assert hist[-1] / hist[-2] < 0.2          # multigrid-rate contraction
    K = RM.laplacian(g.dxm[1:-1], g.dym[1:-1], g.dzm[1:-1])
    exact = spla.spsolve(K.tocsc(), RM.flat(b))
    np.testing.assert_allclose(RM.flat(x), exact, rtol=1e-8, atol=1e-9 * np.abs(exact).max())


def test_vcycle_cg_smoother_converges():
    nx = 16
    g, A, b = _problem(nx, nx, nx)
    mg = O.Multigrid(g, 2, 60, smoother=1)
    x = O.padded(nx, nx, nx)
    r = O.padded(nx, nx, nx)
    n, hist, cit, _ = mg.solve_vcycle(x, r, A, b, 4, 1e-10, 1.0)
    assert hist[-1] < 1e-10


def test_uniform_field_range_and_determinism():
    a = O.uniform_field(8, 6, 4, 20210501)
    b = O.uniform_field(8, 6, 4, 20210501)
    c = O.uniform_field(8, 6, 4, 20210502)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    v = RM.interior(a)
    assert v.min() >= -1.0 and v.max() < 1.0 and abs(v.mean()) < 0.2
    assert np.all(a[0] == 0)
