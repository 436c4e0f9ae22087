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


def single_solves(torch, problems, rhss, **opts):
    from peps_mg_amd.multigrid import Multigrid
    out, res = [], []
    q = 0
    for n, nl, r in problems:
        mg = Multigrid(cube(n), nl, 100, **opts)
        for _ in range(r):
            x = torch.zeros_like(rhss[q])
            res.append(mg.solve_vcycle_device(x.data_ptr(), 0, rhss[q].data_ptr(), 2, 1e-10, 1.0))
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
    import torch
    from peps_mg_amd.batch import BatchSolver
    sols, rhss = make_inputs(torch, PROBLEMS, seed=4)
    B = BatchSolver([cube(n) for n, _, _ in PROBLEMS], [nl for _, nl, _ in PROBLEMS], [r for _, _, r in PROBLEMS],
                    ncycle=100, precision=1)
    res = B.solve_device(sols, rhss, 2, 1e-10, 1.0)
    torch.cuda.synchronize()
    ref, _ = single_solves(torch, PROBLEMS, rhss)
    for i in range(B.size):
        assert res[i].converged and res[i].rho < 1e-10 and res[i].refine_steps >= 1, (i, res[i])
        # both solutions satisfy the 1e-10 residual test; they agree to the
        # solution error that allows (relative 1e-8 here)
        err = (sols[i] - ref[i]).abs().max().item() / ref[i].abs().max().item()
        assert err < 1e-8, (i, err)
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
