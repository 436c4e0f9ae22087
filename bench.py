"""Benchmark: time to 1e-10 relative residual at 512^3, mixed precision vs FP64.

Workload: -lap(u) = f on the unit cube, 512^3 cells, homogeneous Dirichlet,
7-point stencil, b = -f with f = U[-1,1] (SplitMix64, seed 20210501) +
3 pi^2 sin(pi x) sin(pi y) sin(pi z); V(2,2) red/black Gauss-Seidel
cycles (omega = 1), 8-child volume-weighted restriction, trilinear
prolongation, 7 levels down to 4^3, coarsest solved to the tolerance.

A step is one complete solve from x = 0 to rho < 1e-10 with the inputs
already resident in HBM.  value = mean wall time of a mixed-precision step
(FP32 V-cycles inside FP64 iterative refinement); the FP64 solve is timed
the same way and reported beside it.  The roofline object is for the
fine-level smoother kernel, timed with HIP events on the solve stream.
cpu_baseline times the oracle (the C restatement of the reference loops,
OpenMP) on two V-cycles of the same 512^3 workload and extrapolates to the
FP64 cycle count.
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, default=512)
    p.add_argument("--lz", type=float, default=1.0, help="domain length in z (2.0: hx:hy:hz = 1:1:2)")
    p.add_argument("--inner", type=int, default=0, help="FP32 V-cycles per refinement step (0: default)")
    p.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    p.add_argument("--no-traffic", action="store_true",
                   help="skip the two rocprofv3 counter passes that measure the smoother's HBM traffic")
    p.add_argument("--smoother-reps", type=int, default=20)
    p.add_argument("--adjoint", action="store_true", help="run the adjoint inverse-source workload instead")
    p.add_argument("--adjoint-n", type=int, default=128)
    p.add_argument("--no-tune", action="store_true",
                   help="skip the autotuner (solve workload: time the default options only)")
    p.add_argument("--batch", action="store_true",
                   help="run the batched workload (48 cubes of mixed sizes, 8 with 4 right-hand sides) instead")
    p.add_argument("--variants", action="store_true",
                   help="run the FP64 GPU variants (RBGS or CG smoother with 4 iterations; coarsening to 4^3 in "
                        "per-level kernels or below 16^3 in the persistent kernel) and mixed precision at 256^3 "
                        "and 512^3 instead")
    p.add_argument("--kernels-only", action="store_true",
                   help="run only the fine-level smoother (FP64 and FP32) and a 1 GiB device copy, for "
                        "rocprofv3 counter passes; prints one JSON line with the event timings")
    return p.parse_args()


def kernels_only(a):
    """Smoother launches on the 512^3 workload plus a plain copy of known size,
    the calibration for rocprofv3 FETCH_SIZE / WRITE_SIZE."""
    import torch
    from peps_mg_amd import _lib, geometry
    from peps_mg_amd.multigrid import Multigrid
    n = a.n
    levels = int(round(math.log2(n // 4)))
    gdm = geometry.geometry_domain_create(n, n, n, 0.0, 0.0, 0.0, 1.0, 1.0, a.lz)
    sdm = geometry.geometry_subdomain_create(gdm)
    out = {}
    for prec in (0, 1):
        mg = Multigrid(sdm, levels, 1, precision=prec)
        avg = C.c_double(0)
        nb = C.c_double(0)
        kn = C.c_char_p()
        _lib.check(_lib.lib().pepsmg_bench_smoother(mg.h, prec, a.smoother_reps, 1.0, C.byref(avg), C.byref(nb),
                                                    C.byref(kn)))
        out[kn.value.decode()] = {"avg_ms": avg.value, "bytes": nb.value, "launches": 2 * a.smoother_reps}
        mg.destroy()
    src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    out["copy_1GiB"] = {"bytes_read": 1 << 30, "bytes_written": 1 << 30}
    print(json.dumps(out))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measure_traffic(a, kernel):
    """HBM bytes per launch of `kernel` from two rocprofv3 counter passes
    (FETCH_SIZE, then WRITE_SIZE, in KiB per dispatch) over this script's
    --kernels-only mode, each scaled by what the same counter reads for the
    1 GiB device copy that mode also runs (copy bytes / counter).  Each pass
    runs as a child process with its own time limit; None if a pass fails."""
    import csv
    import shutil
    import signal
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if not prof:
        return None, "rocprofv3 not found"
    out = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="pepsmg_pmc_")
        cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "pmc", "--", sys.executable,
               os.path.abspath(__file__), "--kernels-only", "--smoother-reps", "2", "--n", str(a.n), "--lz",
               str(a.lz)]
        log(f"[bench] counter pass {counter}")
        proc = subprocess.Popen(cmd, cwd=d, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                                start_new_session=True)
        try:
            rc = proc.wait(timeout=150)
        except subprocess.TimeoutExpired:
            os.killpg(proc.pid, signal.SIGKILL)
            proc.wait()
            return None, f"{counter} pass timed out"
        if rc != 0:
            return None, f"{counter} pass exited with {rc}"
        path = os.path.join(d, "pmc_counter_collection.csv")
        if not os.path.exists(path):
            return None, f"{counter} pass wrote no counters"
        vals, copy = [], []
        for r in csv.DictReader(open(path)):
            name = r["Kernel_Name"]
            if kernel.replace(" ", "") in name.replace(" ", ""):
                vals.append(float(r["Counter_Value"]))
            elif "copyBuffer" in name and float(r["Counter_Value"]) > 1e5:
                copy.append(float(r["Counter_Value"]))
        shutil.rmtree(d, ignore_errors=True)
        if not vals or not copy:
            return None, f"{counter}: kernel or calibration copy not found"
        scale = (1 << 30) / 1024.0 / (sum(copy) / len(copy))   # KiB the copy moves / KiB counted
        out[counter] = sum(vals) / len(vals) * 1024.0 * scale
        out[counter + "_scale"] = scale
    return out, None


def adjoint_workload(a):
    """Inverse source problem on a 128^3 grid: recover f from u_obs, where
    -lap(u) = f (A u = -f), u_obs = u(f*) + N(0, 1e-3) and f* is a sum of six
    Gaussians.  Loss 0.5 |u(f) - u_obs|^2 + 1e-4 |grad f|^2 (forward
    differences / h); gradients through the adjoint multigrid solve.  Reports
    the wall time of 50 L-BFGS iterations (torch.optim.LBFGS) in FP64 and
    mixed precision, and a central-difference check (eps 1e-6, 32 random
    unit directions) of the gradient."""
    import numpy as np
    import torch
    from peps_mg_amd.adjoint import PoissonSolver, poisson_solve
    n = a.adjoint_n
    levels = int(round(math.log2(n // 4)))
    rng = np.random.default_rng(20210501)
    h = 1.0 / n
    c = (np.arange(n) + 0.5) * h
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    fstar = np.zeros((n, n, n))
    for _ in range(6):
        cx, cy, cz = rng.uniform(0.2, 0.8, 3)
        w = rng.uniform(0.03, 0.1)
        amp = rng.uniform(-1, 1)
        fstar += amp * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2) / (2 * w * w))
    dev = torch.device("cuda:0")
    out = {}
    for prec in (0, 1):
        S = PoissonSolver.uniform(n, levels, tolerance=1e-10, precision=prec)
        ft = torch.tensor(fstar, device=dev)
        with torch.no_grad():
            uobs = S.solve(-ft) + torch.tensor(rng.normal(scale=1e-3, size=(n, n, n)), device=dev)

        def loss_fn(f):
            u = poisson_solve(-f, S)
            reg = sum(((torch.diff(f, dim=d) / h) ** 2).sum() for d in range(3))
            return 0.5 * ((u - uobs) ** 2).sum() + 1e-4 * reg

        # central-difference check of the gradient at a random point
        f0 = torch.tensor(rng.uniform(-1, 1, (n, n, n)), device=dev, requires_grad=True)
        loss_fn(f0).backward()
        g = f0.grad.detach()
        eps, worst = 1e-6, 0.0
        with torch.no_grad():
            for _ in range(32):
                d = torch.tensor(rng.normal(size=(n, n, n)), device=dev)
                d /= d.norm()
                fd = (loss_fn(f0 + eps * d) - loss_fn(f0 - eps * d)).item() / (2 * eps)
                worst = max(worst, abs(fd - (g * d).sum().item()) / g.norm().item())
        f = torch.zeros((n, n, n), dtype=torch.float64, device=dev, requires_grad=True)
        opt = torch.optim.LBFGS([f], lr=1.0, max_iter=50, max_eval=10 ** 6, history_size=10,
                                tolerance_grad=0.0, tolerance_change=0.0, line_search_fn="strong_wolfe")
        evals = [0]

        def closure():
            opt.zero_grad()
            L = loss_fn(f)
            L.backward()
            evals[0] += 1
            return L

        with torch.no_grad():
            l0 = loss_fn(f).item()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        opt.step(closure)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        with torch.no_grad():
            l1 = loss_fn(f).item()
            rel = (torch.linalg.vector_norm(f - ft) / torch.linalg.vector_norm(ft)).item()
        tag = "fp64" if prec == 0 else "mixed"
        log(f"[bench] adjoint {tag}: 50 L-BFGS iterations in {dt:.3f} s ({evals[0]} evaluations), "
            f"loss {l0:.4e} -> {l1:.4e}, FD worst {worst:.2e}")
        out[tag] = {"lbfgs50_s": round(dt, 4), "evaluations": evals[0], "loss0": l0, "loss50": l1,
                    "rel_error_vs_fstar": rel, "fd_worst_rel": worst}
        S.destroy()
    print(json.dumps({"metric": "adjoint inverse source, 50 L-BFGS iterations", "value": out["mixed"]["lbfgs50_s"],
                      "unit": "s", "higher_is_better": False, "n_gpus": 1, "data": "synthetic",
                      "config": {"workload": f"inverse source, {n}^3, 6 Gaussians, noise 1e-3, reg 1e-4 |grad f|^2"},
                      "fp64": out["fp64"], "mixed": out["mixed"]}))


def variants_workload(a):
    """Time to 1e-10 relative residual for the smoother and coarsest-level
    variants of the FP64 GPU drop-in (red/black Gauss-Seidel or CG with 4
    iterations per smoothing step; every level down to 4^3 in per-level
    kernels, or levels of at most 16^3 cells inside the persistent coarse
    kernel) and for mixed precision, at 256^3 and 512^3 (same right-hand side
    as the default workload).  value = the default FP64 solve at 512^3."""
    import torch
    from peps_mg_amd import _lib, geometry
    from peps_mg_amd.multigrid import Multigrid
    cases = [("rbgs_fp64_persist16", dict(), 2), ("rbgs_fp64_full_coarsening", dict(coarse_switch=0), 2),
             ("cg4_fp64_persist16", dict(smoother=1), 4), ("cg4_fp64_full_coarsening", dict(smoother=1, coarse_switch=0), 4),
             ("rbgs_mixed", dict(precision=1), 2), ("cg4_mixed", dict(smoother=1, precision=1), 4)]
    out = {}
    for n in (256, 512):
        levels = int(round(math.log2(n // 4)))
        sdm = geometry.geometry_subdomain_create(geometry.geometry_domain_create(n, n, n, 0.0, 0.0, 0.0, 1.0, 1.0,
                                                                                 a.lz))
        d_b = torch.zeros((n + 2,) * 3, dtype=torch.float64, device="cuda")
        _lib.check(_lib.lib().pepsmg_fill_rhs_device(n, n, n, C.c_void_p(d_b.data_ptr()), 20210501, -1.0,
                                                     -3 * math.pi ** 2, _lib.dptr(sdm.xg), _lib.dptr(sdm.yg),
                                                     _lib.dptr(sdm.zg), None))
        d_x = torch.zeros_like(d_b)
        stream = torch.cuda.current_stream().cuda_stream
        for label, kw, nu in cases:
            mg = Multigrid(sdm, levels, 200, **kw)
            best, res = float("inf"), None
            for _ in range(1 + max(1, a.steps // 2)):
                d_x.zero_()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                res = mg.solve_vcycle_device(d_x.data_ptr(), 0, d_b.data_ptr(), nu, 1e-10, 1.0, stream)
                torch.cuda.synchronize()
                best = min(best, time.perf_counter() - t0)
            out[f"{n}^3/{label}"] = {"s": round(best, 5), "cycles": res.cycles, "refine_steps": res.refine_steps,
                                     "converged": bool(res.converged), "rho": res.rho}
            log(f"[bench] {n}^3 {label}: {best * 1e3:.2f} ms, {res.cycles} cycles")
            mg.destroy()
        del d_x, d_b
        torch.cuda.empty_cache()
    print(json.dumps({"metric": "time to 1e-10 rel. residual, FP64 GPU variants and mixed precision",
                      "value": out["512^3/rbgs_fp64_persist16"]["s"], "unit": "s", "higher_is_better": False,
                      "n_gpus": 1, "data": "synthetic",
                      "config": {"workload": f"Poisson 7-point, 256^3 and 512^3 cells, unit cube x (1,1,{a.lz:g}), "
                                             "homogeneous Dirichlet, V-cycles to 4^3, tol 1e-10, x0=0"},
                      "variants": out}))


def batch_workload(a):
    """48 independent Dirichlet Poisson problems on unit cubes with edges
    drawn from {32, 48, 64, 96, 128, 192} (seeded), each with its own
    U[-1,1] right-hand side; 8 of them carry 4 right-hand sides (72 solves).
    Coarsest levels 4^3 (edges 32, 64, 128) or 3^3 (48, 96, 192).  The
    batched solver (FP32 V-cycles in FP64 refinement, every operation one
    launch over all solves, steps replayed from a hipGraph) is timed against
    looping the single-grid FP64 solver over the same 72 solves.  value =
    aggregate unknowns per second of the batched solve."""
    import numpy as np
    import torch
    from peps_mg_amd import geometry
    from peps_mg_amd.batch import BatchSolver
    from peps_mg_amd.multigrid import Multigrid
    rng = np.random.default_rng(20210501)
    edges = [int(e) for e in rng.choice([32, 48, 64, 96, 128, 192], size=48)]
    nrhs = [4 if p < 8 else 1 for p in range(48)]
    levels = [int(round(math.log2(e // (3 if e % 3 == 0 else 4)))) for e in edges]
    sdms = [geometry.geometry_subdomain_create(geometry.geometry_domain_create(e, e, e, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0))
            for e in edges]
    gen = torch.Generator(device="cuda").manual_seed(20210501)
    rhss, sols = [], []
    for e, r in zip(edges, nrhs):
        for _ in range(r):
            b = torch.zeros((e + 2,) * 3, dtype=torch.float64, device="cuda")
            b[1:-1, 1:-1, 1:-1] = torch.rand((e,) * 3, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
            rhss.append(b)
            sols.append(torch.zeros_like(b))
    unknowns = sum(e ** 3 * r for e, r in zip(edges, nrhs))
    log(f"[bench] batch: {len(rhss)} solves, {unknowns / 1e6:.1f} M unknowns, edges {sorted(set(edges))}")

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    out = {}
    tuned = {}
    for prec, tag in ((1, "mixed"), (0, "fp64")):
        B = BatchSolver(sdms, levels, nrhs, ncycle=100, precision=prec)
        if not a.no_tune:
            from peps_mg_amd.autotune import tune_batch
            tuned[tag] = tune_batch(B, sols, rhss, 2, 1e-10, 1.0)
            log(f"[bench] autotuned {tag}: {tuned[tag]}")
        res = []

        def run():
            for s in sols:
                s.zero_()
            res[:] = B.solve_device(sols, rhss, 2, 1e-10, 1.0, use_graph=True)

        t = timed(run, a.steps, a.warmup)
        ok = all(r.converged for r in res)
        out["batch_" + tag] = {"s": round(t, 5), "unknowns_per_s": round(unknowns / t, 1), "all_converged": ok,
                               "max_rho": max(r.rho for r in res)}
        log(f"[bench] batched {tag}: {t * 1e3:.2f} ms, converged={ok}")
        B.destroy()
    # loop of single-grid FP64 solves (the per-problem path)
    mgs = {}
    for e, nl in zip(edges, levels):
        if e not in mgs:
            mgs[e] = Multigrid(sdms[edges.index(e)], nl, 100)
    item_edge = [e for e, r in zip(edges, nrhs) for _ in range(r)]

    def loop():
        for s, b, e in zip(sols, rhss, item_edge):
            s.zero_()
            mgs[e].solve_vcycle_device(s.data_ptr(), 0, b.data_ptr(), 2, 1e-10, 1.0)

    t_loop = timed(loop, a.steps, a.warmup)
    log(f"[bench] looped single-grid fp64: {t_loop * 1e3:.2f} ms")
    value = unknowns / out["batch_mixed"]["s"]
    print(json.dumps({
        "metric": "batched solves: aggregate unknowns/s to 1e-10 rel. residual", "value": round(value, 1),
        "unit": "unknowns/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(out["batch_mixed"]["s"] * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "48 unit cubes, edges from {32,48,64,96,128,192} (seed 20210501), 8 with 4 RHS, "
                               "U[-1,1] RHS, V(2,2) RBGS, tol 1e-10", "solves": len(rhss), "unknowns": unknowns},
        "batch_mixed": out["batch_mixed"], "batch_fp64": out["batch_fp64"], "autotuned": tuned,
        "looped_single_fp64_s": round(t_loop, 5), "looped_single_fp64_unknowns_per_s": round(unknowns / t_loop, 1),
        "speedup_batch_mixed_vs_looped_fp64": round(t_loop / out["batch_mixed"]["s"], 3)}))


def cpu_baseline(n, lz, fp64_cycles):
    """Oracle (C restatement, OpenMP) on 2 V-cycles of the same workload."""
    import numpy as np
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    g = O.Grid(n, n, n, l=(1.0, 1.0, lz))
    A = g.matrix()
    b = O.uniform_field(n, n, n, 20210501)
    s = np.sin(np.pi * g.xg)[None, None, 1:-1] * np.sin(np.pi * g.yg)[None, 1:-1, None] * \
        np.sin(np.pi * g.zg)[1:-1, None, None]
    b[1:-1, 1:-1, 1:-1] = -(b[1:-1, 1:-1, 1:-1] + 3 * np.pi ** 2 * s)
    levels = int(round(math.log2(n // 4)))
    ncyc = 2
    mg = O.Multigrid(g, levels, ncyc)
    x = O.padded(n, n, n)
    r = O.padded(n, n, n)
    t0 = time.perf_counter()
    nc, hist, cit, res0 = mg.solve_vcycle(x, r, A, b, 2, 1e-30, 1.0)
    dt = time.perf_counter() - t0
    per_cycle = dt / nc
    return {"value": per_cycle * fp64_cycles, "unit": "s", "cores": threads, "kind": "port",
            "sample": f"oracle (C restatement of the reference loops, AoS coefficients, OpenMP {threads} threads) "
                      f"ran {nc} V(2,2) cycles at {n}^3 in {dt:.2f} s ({per_cycle:.2f} s/cycle); "
                      f"value = s/cycle x {fp64_cycles} cycles the FP64 solve needs"}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "the solver runs on one GPU; sharding a grid over several GPUs is not "
                                   "implemented", "n_gpus": a.gpus}))
        sys.exit(2)
    if a.kernels_only:
        kernels_only(a)
        return
    if a.adjoint:
        adjoint_workload(a)
        return
    if a.batch:
        batch_workload(a)
        return
    if a.variants:
        variants_workload(a)
        return
    import numpy as np
    import torch
    from peps_mg_amd import _lib, geometry
    from peps_mg_amd.multigrid import Multigrid

    n, lz = a.n, a.lz
    levels = int(round(math.log2(n // 4)))
    gdm = geometry.geometry_domain_create(n, n, n, 0.0, 0.0, 0.0, 1.0, 1.0, lz)
    sdm = geometry.geometry_subdomain_create(gdm)
    dev = torch.device("cuda:0")
    d_b = torch.zeros((n + 2,) * 3, dtype=torch.float64, device=dev)
    _lib.check(_lib.lib().pepsmg_fill_rhs_device(n, n, n, C.c_void_p(d_b.data_ptr()), 20210501, -1.0,
                                                 -3 * math.pi ** 2, _lib.dptr(sdm.xg), _lib.dptr(sdm.yg),
                                                 _lib.dptr(sdm.zg), None))
    d_x = torch.zeros_like(d_b)
    stream = torch.cuda.current_stream().cuda_stream

    def make(precision):
        kw = dict(precision=precision)
        if precision == 1 and a.inner:
            kw["inner_cycles"] = a.inner
        return Multigrid(sdm, levels, 200, **kw)

    def run(mg, steps, warmup, label):
        res = None
        for _ in range(warmup):
            d_x.zero_()
            res = mg.solve_vcycle_device(d_x.data_ptr(), 0, d_b.data_ptr(), 2, 1e-10, 1.0, stream)
        torch.cuda.synchronize()
        times = []
        t_all = time.perf_counter()
        for q in range(steps):
            t0 = time.perf_counter()
            d_x.zero_()
            res = mg.solve_vcycle_device(d_x.data_ptr(), 0, d_b.data_ptr(), 2, 1e-10, 1.0, stream)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            log(f"[bench] {label} step {q}: {times[-1]*1e3:.2f} ms, cycles={res.cycles}, rho={res.rho:.3e}")
        torch.cuda.synchronize()
        total = time.perf_counter() - t_all
        return total / steps, res

    log(f"[bench] {n}^3 lz={lz} levels={levels}")
    mg64 = make(0)
    t64, r64 = run(mg64, a.steps, a.warmup, "fp64")
    x64 = d_x.clone()
    mgmx = make(1)
    tmx, rmx = run(mgmx, a.steps, a.warmup, "mixed")
    default_times = {"fp64_time_s": round(t64, 6), "mixed_time_s": round(tmx, 6)}
    # The autotuner (peps_mg_amd.autotune.tune_solver, measured afresh on this
    # GPU, outside the timed region) picks the mixed solver's tile, coarse
    # switch and inner V-cycle schedule; the tuned solves are timed like the
    # default ones and give value.  The FP64 solve keeps its defaults: its
    # options are launch shapes only (same x bit for bit), and the per-level
    # default tiles are its tuned ones.
    tuned, extra = None, []
    if not a.no_tune:
        from peps_mg_amd.autotune import tune_solver
        names = {f for f, _ in _lib.Options._fields_}
        best = tune_solver(sdm, levels, d_x, d_b, 2, 1e-10, 1.0, precision=1, use_cache=False)
        tuned = {"mixed": {k: v for k, v in best.items() if k in names}}
        log(f"[bench] autotuned mixed: {tuned['mixed']}")
        mgmxt = Multigrid(sdm, levels, 200, **tuned["mixed"])
        tmx, rmx = run(mgmxt, a.steps, a.warmup, "mixed tuned")
        extra = [mgmxt]
    err_max = float((d_x - x64).abs().max() / x64.abs().max())
    err_l2 = float(torch.linalg.vector_norm(d_x - x64) / torch.linalg.vector_norm(x64))

    # fine-level smoother kernel, timed with events on the solve stream
    avg = C.c_double(0)
    nbytes = C.c_double(0)
    kname = C.c_char_p()
    _lib.check(_lib.lib().pepsmg_bench_smoother(mg64.h, 0, a.smoother_reps, 1.0, C.byref(avg), C.byref(nbytes),
                                                C.byref(kname)))
    kname64 = kname.value.decode()   # the library reuses its name buffer on the next call
    achieved = nbytes.value / (avg.value * 1e-3) / 1e9
    avg32 = C.c_double(0)
    nbytes32 = C.c_double(0)
    kname32 = C.c_char_p()
    _lib.check(_lib.lib().pepsmg_bench_smoother(mgmx.h, 1, a.smoother_reps, 1.0, C.byref(avg32),
                                                C.byref(nbytes32), C.byref(kname32)))
    kname32s = kname32.value.decode()
    achieved32 = nbytes32.value / (avg32.value * 1e-3) / 1e9

    # the bandwidth a kernel with the smoother's streams (x, b in; x out) and no
    # stencil reaches on this GPU: the practical ceiling for the smoother
    s_ms, s_bytes = C.c_double(0), C.c_double(0)
    _lib.check(_lib.lib().pepsmg_bench_stream(n ** 3, 10, C.byref(s_ms), C.byref(s_bytes)))
    stream_gbs = s_bytes.value / (s_ms.value * 1e-3) / 1e9
    traffic, traffic_note, tr = None, "not measured (--no-traffic)", None
    if not a.no_traffic:
        tr, err = measure_traffic(a, kname64)
        if tr:
            traffic = tr["FETCH_SIZE"] + tr["WRITE_SIZE"]
            traffic_note = (f"rocprofv3 FETCH_SIZE x {tr['FETCH_SIZE_scale']:.3f} + WRITE_SIZE x "
                            f"{tr['WRITE_SIZE_scale']:.3f} (scales from a 1 GiB device copy)")
        else:
            traffic_note = err

    cpu = None
    if not a.no_cpu:
        log("[bench] cpu_baseline: oracle, 2 V-cycles")
        del mg64, mgmx, extra
        torch.cuda.empty_cache()
        cpu = cpu_baseline(n, lz, r64.cycles)

    unknowns = n ** 3
    out = {
        "metric": "time to 1e-10 rel. residual at 512^3 (mixed vs FP64 GPU); smoother HBM GB/s",
        "value": round(tmx, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(tmx * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"Poisson 7-point, {n}^3 cells, unit cube x (1,1,{lz:g}), homogeneous Dirichlet, "
                               f"V(2,2) RBGS omega=1, {levels} coarse levels to 4^3, tol 1e-10, x0=0",
                   "grid": [n, n, n], "levels": levels, "precision": "fp32 V-cycles in fp64 refinement",
                   "options": "mixed: autotuned (tuned_options); fp64: defaults" if tuned else "defaults"},
        "tuned_options": tuned,
        "default_options_times": default_times,
        "fp64_time_s": round(t64, 6),
        "fp64_cycles": r64.cycles,
        "mixed_time_s": round(tmx, 6),
        "mixed_cycles": rmx.cycles,
        "mixed_refine_steps": rmx.refine_steps,
        "speedup_mixed_vs_fp64": round(t64 / tmx, 4),
        "fp64_vcycles_per_s": round(r64.cycles / t64, 3),
        "fp64_unknowns_per_s": round(unknowns / t64, 1),
        "mixed_unknowns_per_s": round(unknowns / tmx, 1),
        "fp64_cell_vcycles_per_s": round(unknowns * r64.cycles / t64, 1),
        "rho_fp64": r64.rho,
        "rho_mixed": rmx.rho,
        "mixed_vs_fp64_error": {"max_rel": err_max, "l2_rel": err_l2},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": round(traffic) if traffic else None, "traffic_note": traffic_note,
                     "traffic_read": round(tr["FETCH_SIZE"]) if tr else None,
                     "traffic_write": round(tr["WRITE_SIZE"]) if tr else None,
                     "kernel": kname64, "avg_ms": avg.value, "bytes_per_launch": nbytes.value,
                     "stream_ceiling_gbs": round(stream_gbs, 1),
                     "frac_of_stream_ceiling": round(achieved / stream_gbs, 4)},
        "roofline_fp32_smoother": {"achieved": round(achieved32, 1), "frac": round(achieved32 / HBM_PEAK_GBS, 4),
                                   "kernel": kname32s, "avg_ms": avg32.value,
                                   "bytes_per_launch": nbytes32.value},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
