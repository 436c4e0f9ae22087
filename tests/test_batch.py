"""Batched solves (pepsmg_batch_*): every item against its own single-grid
solve through the same C-ABI, and graph replay against direct launches.

In FP64 an item's batched solve runs the same smoothing and transfer
kernels in the same order as its single-grid solve, so solutions are
identical bit for bit; the residual norms are sums of the same squares in
another order (the single-grid solver takes them from its smoother), so
histories agree to rounding.  In mixed precision the batch always runs inner_cycles cycles
in the last refinement step (the single-grid solver may run fewer), so the
check there is convergence and agreement with the FP64 solution.
"""
import math

import numpy as np
import pytest

from peps_mg_amd import _lib, geometry

PROBLEMS = [(16, 2, 1), (24, 3, 2), (32, 3, 1), (20, 2, 1)]   # (edge, coarse levels, right-hand sides)


def cube(n):
    return geometry.geometry_subdomain_create(geometry.geometry_domain_create(n, n, n, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0))


def test_batch_symbols_in_table():
    for name in ["pepsmg_batch_create", "pepsmg_batch_set_options", "pepsmg_batch_size",
                 "pepsmg_batch_solve_device", "pepsmg_batch_destroy"]:
        assert name in _lib.SIGNATURES


def make_inputs(torch, problems, seed=3):
    gen = torch.Generator(device="cuda").manual_seed(seed)
    sols, rhss = [], []
    for n, _, r in problems:
        for _ in range(r):
            b = torch.zeros((n + 2,) * 3, dtype=torch.float64, device="cuda")
            b[1:-1, 1:-1, 1:-1] = torch.rand((n,) * 3, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
            rhss.append(b)
            sols.append(torch.zeros_like(b))
    return sols, rhss


def single_solves(torch, problems, rhss, tol=1e-10, **opts):
    from peps_mg_amd.multigrid import Multigrid
    out, res = [], []
    q = 0
    for n, nl, r in problems:
        mg = Multigrid(cube(n), nl, 100, **opts)
        for _ in range(r):
            x = torch.zeros_like(rhss[q])
            res.append(mg.solve_vcycle_device(x.data_ptr(), 0, rhss[q].data_ptr(), 2, tol, 1.0))
            out.append(x)
            q += 1
        mg.destroy()
    torch.cuda.synchronize()
    return out, res


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [True, False])
def test_batch_fp64_matches_single_bitwise(graph):
    import torch
    from peps_mg_amd.batch import BatchSolver
    sols, rhss = make_inputs(torch, PROBLEMS)
    B = BatchSolver([cube(n) for n, _, _ in PROBLEMS], [nl for _, nl, _ in PROBLEMS], [r for _, _, r in PROBLEMS],
                    ncycle=100)
    assert B.size == sum(r for _, _, r in PROBLEMS)
    res = B.solve_device(sols, rhss, 2, 1e-10, 1.0, use_graph=graph)
    torch.cuda.synchronize()
    ref, rref = single_solves(torch, PROBLEMS, rhss)
    for i in range(B.size):
        assert res[i].converged and res[i].cycles == rref[i].cycles, (i, res[i], rref[i])
        np.testing.assert_allclose(res[i].history, rref[i].history, rtol=1e-12, atol=0)
        assert torch.equal(sols[i], ref[i]), i
    B.destroy()


@pytest.mark.gpu
def test_batch_mixed_converges_to_fp64_solution():
    """Mixed-precision items within 1e-10 (max, relative) of the FP64
    single-grid solves.  The right-hand sides are purely random, whose
    solutions are only ~1e-9 accurate at a 1e-10 residual, so both sides
    solve to 1e-12 (then within 3e-11 of the exact solution)."""
    import torch
    from peps_mg_amd.batch import BatchSolver
    sols, rhss = make_inputs(torch, PROBLEMS, seed=4)
    B = BatchSolver([cube(n) for n, _, _ in PROBLEMS], [nl for _, nl, _ in PROBLEMS], [r for _, _, r in PROBLEMS],
                    ncycle=100, precision=1)
    res = B.solve_device(sols, rhss, 2, 1e-12, 1.0)
    torch.cuda.synchronize()
    ref, _ = single_solves(torch, PROBLEMS, rhss, tol=1e-12)
    for i in range(B.size):
        assert res[i].converged and res[i].rho < 1e-12 and res[i].refine_steps >= 1, (i, res[i])
        err = (sols[i] - ref[i]).abs().max().item() / ref[i].abs().max().item()
        assert err <= 1e-10, (i, err)
    B.destroy()


@pytest.mark.gpu
def test_batch_graph_replays_equal_direct_and_repeat():
    """Two graph solves in a row and a direct-launch solve give identical
    results (the graph holds no memset nodes; replays reuse the tables)."""
    import torch
    from peps_mg_amd.batch import BatchSolver
    probs = [(16, 2, 2), (32, 3, 1)]
    B = BatchSolver([cube(n) for n, _, _ in probs], [nl for _, nl, _ in probs], [r for _, _, r in probs],
                    ncycle=100, precision=1)
    outs = []
    for graph in (True, True, False):
        sols, rhss = make_inputs(torch, probs, seed=9)
        B.solve_device(sols, rhss, 2, 1e-10, 1.0, use_graph=graph)
        torch.cuda.synchronize()
        outs.append([s.clone() for s in sols])
    for k in (1, 2):
        for a, b in zip(outs[0], outs[k]):
            assert torch.equal(a, b)
    B.destroy()


@pytest.mark.gpu
def test_batch_rejects_bad_input():
    import torch
    from peps_mg_amd.batch import BatchSolver
    B = BatchSolver([cube(16)], [2], [1])
    sols, rhss = make_inputs(torch, [(16, 2, 1)])
    sols[0][0, 0, 0] = 1.0
    with pytest.raises(_lib.PepsmgError) as e:
        B.solve_device(sols, rhss, 2, 1e-10, 1.0)
    assert e.value.code == _lib.ERR_GHOST
    with pytest.raises(_lib.PepsmgError):
        B.set_options(smoother=1)
    B.destroy()
    with pytest.raises(_lib.PepsmgError):
        BatchSolver([cube(12)], [1], [1]).solve_device([0], [0], 2, 1e-10, 1.0)


@pytest.mark.gpu
def test_batch_per_problem_tiles_fp64_bitwise():
    """Problems with different tile shapes run in separate launches; FP64
    values stay those of the single-grid solves."""
    import torch
    from peps_mg_amd.batch import BatchSolver
    sols, rhss = make_inputs(torch, PROBLEMS, seed=5)
    B = BatchSolver([cube(n) for n, _, _ in PROBLEMS], [nl for _, nl, _ in PROBLEMS], [r for _, _, r in PROBLEMS],
                    ncycle=100)
    B.set_tiles([0, 64, 64, 0], [8, 16, 32, 0])
    res = B.solve_device(sols, rhss, 2, 1e-10, 1.0)
    torch.cuda.synchronize()
    ref, rref = single_solves(torch, PROBLEMS, rhss)
    for i in range(B.size):
        np.testing.assert_allclose(res[i].history, rref[i].history, rtol=1e-12, atol=0)
        assert torch.equal(sols[i], ref[i]), i
    with pytest.raises(_lib.PepsmgError):
        B.set_tiles([32, 0, 0, 0], None)
    B.destroy()


@pytest.mark.gpu
def test_autotune_tiles_and_batch(tmp_path, monkeypatch):
    import torch
    from peps_mg_amd import autotune
    from peps_mg_amd.batch import BatchSolver
    monkeypatch.setenv("PEPSMG_TUNE_CACHE", str(tmp_path / "tune.json"))
    t = autotune.tune_tiles(cube(32), 3, 1, reps=2)
    assert t["tile_x"] in (64, 128) and t["tile_y"] in (8, 16, 32) and t["smoother_ms"] > 0
    assert autotune.tune_tiles(cube(32), 3, 1) == t            # served from the cache
    probs = [(16, 2, 1), (32, 3, 2)]
    sols, rhss = make_inputs(torch, probs, seed=6)
    B = BatchSolver([cube(n) for n, _, _ in probs], [nl for _, nl, _ in probs], [r for _, _, r in probs],
                    ncycle=100, precision=1)
    out = autotune.tune_batch(B, sols, rhss, inner_cycles=(2, 3), repeats=0)
    assert out["inner_cycles"] in (2, 3) and "32x32x32" in out["tiles"]
    for x in sols:
        x.zero_()
    res = B.solve_device(sols, rhss, 2, 1e-10, 1.0)
    assert all(r.converged for r in res)
    B.destroy()


# The batched workload's size classes: unit cubes of edge 32, 48, 64, 96, 128
# and 192 (coarsest 4^3, or 3^3 for edges divisible by 3), one problem each,
# the 64^3 one with four right-hand sides.
CFG3 = [(32, 3, 1), (48, 4, 1), (64, 4, 4), (96, 5, 1), (128, 5, 1), (192, 6, 1)]


def cfg3_inputs(torch):
    """U[-1,1] right-hand sides from the oracle's SplitMix64 field, one seed
    per item; host copies for the oracle, device copies for the batch."""
    from oracle import oracle as O
    host, sols, rhss = [], [], []
    seed = 100
    for n, _, r in CFG3:
        for _ in range(r):
            b = -O.uniform_field(n, n, n, seed)
            seed += 1
            host.append(b)
            rhss.append(torch.from_numpy(b).cuda())
            sols.append(torch.zeros_like(rhss[-1]))
    return host, sols, rhss


def oracle_items(host, tol):
    from oracle import oracle as O
    out = []
    q = 0
    for n, nl, r in CFG3:
        g = O.Grid(n, n, n)
        A = g.matrix()
        for _ in range(r):
            mo = O.Multigrid(g, nl, 100)
            x, rs = O.padded(n, n, n), O.padded(n, n, n)
            nc, hist, cit, res0 = mo.solve_vcycle(x, rs, A, host[q], 2, tol, 1.0)
            out.append((x, nc, hist))
            q += 1
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [True, False])
def test_batch_cfg3_fp64_matches_oracle_bitwise(graph):
    """Every item of a batch spanning the size classes (4 right-hand sides
    on the 64^3 problem) equals the oracle's solve bit for bit, with the
    same cycle count and rho history to 1e-12."""
    import torch
    from peps_mg_amd.batch import BatchSolver
    host, sols, rhss = cfg3_inputs(torch)
    B = BatchSolver([cube(n) for n, _, _ in CFG3], [nl for _, nl, _ in CFG3], [r for _, _, r in CFG3], ncycle=100)
    res = B.solve_device(sols, rhss, 2, 1e-10, 1.0, use_graph=graph)
    torch.cuda.synchronize()
    ref = oracle_items(host, 1e-10)
    for i, (x, nc, hist) in enumerate(ref):
        assert res[i].converged and res[i].cycles == nc, (i, res[i], nc)
        assert np.max(np.abs(res[i].history - hist)) <= 1e-12, i
        assert np.array_equal(sols[i].cpu().numpy(), x), i
    B.destroy()


@pytest.mark.gpu
def test_batch_cfg3_mixed_within_1e10_of_oracle():
    """The same batch in mixed precision (FP32 V-cycles, FP64 refinement)
    with the autotuner's choice of tile per size class left at the defaults:
    every item within 1e-10 (max, relative) of the oracle's FP64 solution,
    both solved to 1e-12 (random right-hand sides; see
    test_batch_mixed_converges_to_fp64_solution)."""
    import torch
    from peps_mg_amd.batch import BatchSolver
    host, sols, rhss = cfg3_inputs(torch)
    B = BatchSolver([cube(n) for n, _, _ in CFG3], [nl for _, nl, _ in CFG3], [r for _, _, r in CFG3], ncycle=100,
                    precision=1)
    res = B.solve_device(sols, rhss, 2, 1e-12, 1.0, use_graph=True)
    torch.cuda.synchronize()
    ref = oracle_items(host, 1e-12)
    for i, (x, nc, hist) in enumerate(ref):
        assert res[i].converged and res[i].rho < 1e-12, (i, res[i])
        err = np.abs(sols[i].cpu().numpy() - x).max() / np.abs(x).max()
        assert err <= 1e-10, (i, err)
    B.destroy()
