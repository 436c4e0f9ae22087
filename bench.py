"""Benchmark: time to 1e-10 relative residual at 512^3, mixed precision vs FP64.

Workload: -lap(u) = f on the unit cube, 512^3 cells, homogeneous Dirichlet,
7-point stencil, b = -f with f = U[-1,1] (SplitMix64, seed 20210501) +
3 pi^2 sin(pi x) sin(pi y) sin(pi z); V(2,2) red/black Gauss-Seidel
cycles (omega = 1), 8-child volume-weighted restriction, trilinear
prolongation, 7 levels down to 4^3, coarsest solved to the tolerance.

A step is one complete solve from x = 0 to rho < 1e-10 with the inputs
already resident in HBM.  value = mean wall time of a mixed-precision step
(FP32 V-cycles inside FP64 iterative refinement, inner schedule picked by
the autotuner before the timed region); the FP64 solve is timed the same
way and reported beside it, together with two runs that separate the gain of
precision from that of the inner schedule (FP64 with the inner cycles' SOR
factor; mixed with plain V(2,2) omega = 1 inner cycles).  The roofline
object is for the fine-level smoother kernel: every launch of it on the
512^3 level inside FP64 solves run after the timed ones is bracketed by HIP
events on the solve stream (the library's measurement hook, which is off
while the solves are timed: each event adds a few microseconds to the
stream); beside that, the same kernel
launched back to back, a rocprofv3 kernel trace of those launches (child
process) and its HBM traffic from two counter passes.

The default run also carries the other workloads (each in one key):
aniso_512 (the 512^3 grid with hx:hy:hz = 1:1:2, FP64 and mixed),
batch_cfg3 (48 cubes of edges 32..192, 8 with 4 right-hand sides, batched
mixed precision against looping the single-grid FP64 solver) and
adjoint_128 (the inverse-source problem: central differences, a CPU
reference gradient, 50 L-BFGS iterations).

cpu_baseline times the oracle (the C restatement of the reference loops,
OpenMP) on a complete FP64 solve of the same 512^3 workload; the same run
gives the oracle's solution, against which the GPU FP64 solve (bit for bit)
and the mixed-precision solve (max / l2 relative error) are checked, and
likewise on the 1:1:2 grid.
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
SEED = 20210501


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, default=512)
    p.add_argument("--lz", type=float, default=1.0, help="domain length in z (2.0: hx:hy:hz = 1:1:2)")
    p.add_argument("--inner", type=int, default=0, help="FP32 V-cycles per refinement step (0: default)")
    p.add_argument("--no-cpu", action="store_true", help="skip the oracle legs (cpu_baseline and parity)")
    p.add_argument("--no-traffic", action="store_true",
                   help="skip the rocprofv3 child passes (HBM traffic counters and the kernel trace)")
    p.add_argument("--no-extras", action="store_true",
                   help="default workload only: skip aniso_512, batch_cfg3 and adjoint_128")
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
    p.add_argument("--cg-profile", action="store_true",
                   help="run only two FP64 CG(4)-smoothed solves of the 512^3 workload (persistent coarse kernel) "
                        "and a 1 GiB device copy, for rocprofv3 passes over the CG kernels")
    p.add_argument("--kernels-only", action="store_true",
                   help="run only the fine-level smoother (FP64 and FP32) and a 1 GiB device copy, for "
                        "rocprofv3 passes; prints one JSON line with the event timings")
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def levels_for(n):
    return int(round(math.log2(n // (3 if n % 3 == 0 else 4))))


def kernels_only(a):
    """Smoother launches on the 512^3 workload plus a plain copy of known
    size, the calibration for rocprofv3 FETCH_SIZE / WRITE_SIZE.  Each
    context is loaded with the workload first (FP64: a solve capped at 0
    cycles stores b and x = 0 in the solver's layout; mixed: one FP32
    V-cycle also fills the FP32 correction), so the smoother runs on the
    data it sees in a solve.  Of each smoother's 2 x smoother_reps launches
    the second half is timed with HIP events; profiler passes take the same
    last smoother_reps dispatches."""
    import torch
    from peps_mg_amd import _lib, geometry
    from peps_mg_amd.multigrid import Multigrid
    n = a.n
    gdm = geometry.geometry_domain_create(n, n, n, 0.0, 0.0, 0.0, 1.0, 1.0, a.lz)
    sdm = geometry.geometry_subdomain_create(gdm)
    d_b = workload_rhs_device(n, a.lz, sdm, torch)
    d_x = torch.zeros_like(d_b)
    out = {}
    for prec in (0, 1):
        mg = Multigrid(sdm, levels_for(n), prec, precision=prec)
        mg.solve_vcycle_device(d_x.data_ptr(), 0, d_b.data_ptr(), 2, 1e-10, 1.0)
        torch.cuda.synchronize()
        avg = C.c_double(0)
        nb = C.c_double(0)
        kn = C.c_char_p()
        _lib.check(_lib.lib().pepsmg_bench_smoother(mg.h, prec, a.smoother_reps, 1.0, C.byref(avg), C.byref(nb),
                                                    C.byref(kn)))
        out[kn.value.decode()] = {"avg_ms": avg.value, "bytes": nb.value, "launches": 2 * a.smoother_reps,
                                  "timed_launches": a.smoother_reps}
        mg.destroy()
    src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    out["copy_1GiB"] = {"bytes_read": 1 << 30, "bytes_written": 1 << 30}
    print(json.dumps(out), flush=True)


def cg_profile(a):
    """Two FP64 solves of the workload with the CG smoother (4 iterations per
    smoothing step, src/cg_poisson_matrix.f90:157-241), the second one timed
    with events around the solve, then a 1 GiB device copy (the counter
    calibration).  Prints one JSON line."""
    import torch
    from peps_mg_amd import geometry
    from peps_mg_amd.multigrid import Multigrid
    n = a.n
    sdm = geometry.geometry_subdomain_create(geometry.geometry_domain_create(n, n, n, 0.0, 0.0, 0.0, 1.0, 1.0, a.lz))
    d_b = workload_rhs_device(n, a.lz, sdm, torch)
    d_x = torch.zeros_like(d_b)
    mg = Multigrid(sdm, levels_for(n), 200, smoother=1)
    t, r = timed_solves(mg, d_x, d_b, 1, 1, "fp64 CG(4)", torch, nu=4)
    mg.destroy()
    src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"cg4_fp64_solve_s": t, "cycles": r.cycles, "rho": r.rho, "device_ms": r.device_ms,
                      "grid": [n, n, n]}), flush=True)


def profiler_child(a, args, tag, timeout=150):
    """Runs rocprofv3 <args> -- python bench.py --kernels-only in a child
    process with its own time limit; returns (output dir, stdout) or
    (None, reason)."""
    import shutil
    import signal
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if not prof:
        return None, "rocprofv3 not found"
    d = tempfile.mkdtemp(prefix="pepsmg_prof_")
    cmd = [prof] + args + ["--output-format", "csv", "-d", d, "-o", tag, "--", sys.executable,
                           os.path.abspath(__file__), "--kernels-only", "--smoother-reps", str(a.smoother_reps),
                           "--n", str(a.n), "--lz", str(a.lz)]
    log(f"[bench] rocprofv3 child pass: {' '.join(args)}")
    proc = subprocess.Popen(cmd, cwd=d, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                            start_new_session=True)
    try:
        out, _ = proc.communicate(timeout=timeout)
    except subprocess.TimeoutExpired:
        os.killpg(proc.pid, signal.SIGKILL)
        proc.communicate()
        return None, f"{' '.join(args)} pass timed out"
    if proc.returncode != 0:
        return None, f"{' '.join(args)} pass exited with {proc.returncode}"
    return d, out


def is_smoother(kernel, name):
    """Whether a traced kernel name is the plain smoother `kernel` or its
    downward-marching twin (the solver alternates the two)."""
    name = name.replace(" ", "")
    k = kernel.replace(" ", "")
    return k in name or k.replace("k_rb_fused<", "k_rb_fused_down<") in name


def measure_traffic(a, kernel):
    """HBM bytes per launch of `kernel` from two rocprofv3 counter passes
    (FETCH_SIZE, then WRITE_SIZE, in KiB per dispatch) over this script's
    --kernels-only mode, each scaled by what the same counter reads for the
    1 GiB device copy that mode also runs (copy bytes / counter)."""
    import csv
    import shutil
    out = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d, msg = profiler_child(a, ["--pmc", counter], "pmc")
        if d is None:
            return None, msg
        path = os.path.join(d, "pmc_counter_collection.csv")
        if not os.path.exists(path):
            return None, f"{counter} pass wrote no counters"
        vals, copy = [], []
        for r in csv.DictReader(open(path)):
            name = r["Kernel_Name"]
            if is_smoother(kernel, name):
                vals.append((int(r.get("Dispatch_Id", len(vals))), float(r["Counter_Value"])))
            elif "copyBuffer" in name and float(r["Counter_Value"]) > 1e5:
                copy.append(float(r["Counter_Value"]))
        shutil.rmtree(d, ignore_errors=True)
        if len(vals) < a.smoother_reps or not copy:
            return None, f"{counter}: kernel or calibration copy not found"
        vals = [v for _, v in sorted(vals)[-a.smoother_reps:]]   # the event-timed launches
        scale = (1 << 30) / 1024.0 / (sum(copy) / len(copy))   # KiB the copy moves / KiB counted
        out[counter] = sum(vals) / len(vals) * 1024.0 * scale
        out[counter + "_scale"] = scale
    return out, None


def trace_smoothers(a, kernels):
    """rocprofv3 --kernel-trace --stats of --kernels-only: per kernel, the mean
    dispatch duration of the event-timed launches (its last smoother_reps
    dispatches), the mean over all its dispatches (what --stats reports:
    also the untimed warm-up launches and the loading solve's), the span
    from the first timed start to the last timed end divided by the launch
    count, and the HIP-event mean the child measured over the same
    launches."""
    import csv
    import shutil
    d, out = profiler_child(a, ["--kernel-trace", "--stats"], "kt")
    if d is None:
        return None, out
    path = os.path.join(d, "kt_kernel_trace.csv")
    if not os.path.exists(path):
        return None, "kernel trace pass wrote no trace"
    events = {}
    for line in out.splitlines():
        line = line.strip()
        if line.startswith("{"):
            try:
                events = json.loads(line)
            except ValueError:
                pass
    rows = list(csv.DictReader(open(path)))
    res = {}
    for k in kernels:
        ds = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in rows
                    if is_smoother(k, r["Kernel_Name"]))
        if len(ds) < a.smoother_reps:
            return None, f"kernel trace: {len(ds)} dispatches of {k}"
        timed = ds[-a.smoother_reps:]
        durs = [(e - s) * 1e-6 for s, e in timed]
        ev = events.get(k, {}).get("avg_ms")
        res[k] = {"trace_avg_ms": sum(durs) / len(durs),
                  "trace_avg_all_ms": sum((e - s) * 1e-6 for s, e in ds) / len(ds),
                  "trace_span_per_launch_ms": (timed[-1][1] - timed[0][0]) * 1e-6 / len(timed),
                  "event_avg_ms_same_run": ev, "dispatches": len(ds), "timed": len(timed)}
    shutil.rmtree(d, ignore_errors=True)
    return res, None


def workload_rhs_device(n, lz, sdm, torch):
    from peps_mg_amd import _lib
    d_b = torch.zeros((n + 2,) * 3, dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().pepsmg_fill_rhs_device(n, n, n, C.c_void_p(d_b.data_ptr()), SEED, -1.0,
                                                 -3 * math.pi ** 2, _lib.dptr(sdm.xg), _lib.dptr(sdm.yg),
                                                 _lib.dptr(sdm.zg), None))
    torch.cuda.synchronize()
    return d_b


def vcycle_roofline(n, levels, t, cycles):
    """Whole FP64 V(2,2) cycle against the HBM peak: algorithmic elements per
    cycle E = n (16.25 sum_{l<L} 8^-l + 2) (per level four red/black
    iterations at 3 n, residual + restriction and prolongation + correction
    at 2.125 n each; the convergence residual at 2 n on the finest level),
    8 bytes each, over the measured time per cycle."""
    e = n ** 3 * (16.25 * sum(8.0 ** -l for l in range(levels)) + 2.0)
    ms = t / cycles * 1e3
    gbs = e * 8 / (ms * 1e-3) / 1e9
    return {"bytes_per_cycle": round(e * 8), "ms_per_cycle": round(ms, 4), "achieved_gbs": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4),
            "model": "4 red/black iterations (3 n each), residual+restriction and prolongation+correction "
                     "(2.125 n each) per level, the convergence residual (2 n) once; the solver folds that "
                     "residual into the next cycle's first smoothing pass and reads no x in the first "
                     "iteration on coarse levels, so it moves fewer bytes than counted"}


def kernel_timing(mg, enable):
    """The library's measurement hook: HIP events on the solve stream around
    every plain fused smoother launch on the finest level of the solves that
    follow; returns (mean ms, launches, kernel) when switched off."""
    from peps_mg_amd import _lib
    L = _lib.lib()
    if enable:
        _lib.check(L.pepsmg_kernel_timing(mg.h, 1))
        return None
    _lib.check(L.pepsmg_kernel_timing(mg.h, 0))
    tot, cnt, name = C.c_double(0), C.c_int32(0), C.c_char_p()
    _lib.check(L.pepsmg_kernel_timing_read(mg.h, C.byref(tot), C.byref(cnt), C.byref(name)))
    return (tot.value / cnt.value if cnt.value else None), cnt.value, name.value.decode()


def timed_solves(mg, d_x, d_b, steps, warmup, label, torch, nu=2, omega=1.0, time_smoother=False):
    """Mean wall time of `steps` solves from x = 0 (after `warmup` untimed
    ones), each bracketed by device synchronisation.  time_smoother: the
    fine-level smoother launches of the timed solves are also timed with HIP
    events (kernel_timing), read back after the last one."""
    stream = torch.cuda.current_stream().cuda_stream
    res = None
    for _ in range(warmup):
        d_x.zero_()
        res = mg.solve_vcycle_device(d_x.data_ptr(), 0, d_b.data_ptr(), nu, 1e-10, omega, stream)
    torch.cuda.synchronize()
    if time_smoother:
        kernel_timing(mg, True)
    t_all = time.perf_counter()
    for q in range(steps):
        t0 = time.perf_counter()
        d_x.zero_()
        res = mg.solve_vcycle_device(d_x.data_ptr(), 0, d_b.data_ptr(), nu, 1e-10, omega, stream)
        torch.cuda.synchronize()
        log(f"[bench] {label} step {q}: {(time.perf_counter() - t0) * 1e3:.2f} ms, cycles={res.cycles}, "
            f"rho={res.rho:.3e}")
    torch.cuda.synchronize()
    t = (time.perf_counter() - t_all) / steps
    if time_smoother:
        res.smoother = kernel_timing(mg, False)
        log(f"[bench] {label}: fine-level smoother {res.smoother}")
    return t, res


def oracle_solve(n, lz, b, cycles=200):
    """The oracle's FP64 solve of the workload (C restatement of the
    reference loops, OpenMP) for the right-hand side b the GPU solved (the
    padded host copy of the device field, so both see the same bits):
    (x, cycles, rho history, res0, seconds of the solve alone, threads)."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    g = O.Grid(n, n, n, l=(1.0, 1.0, lz))
    A = g.matrix()
    mo = O.Multigrid(g, levels_for(n), cycles)
    x = O.padded(n, n, n)
    r = O.padded(n, n, n)
    t0 = time.perf_counter()
    nc, hist, cit, res0 = mo.solve_vcycle(x, r, A, b, 2, 1e-10, 1.0)
    dt = time.perf_counter() - t0
    del A, r, mo
    return x, nc, hist, res0, dt, threads


def compare_to_oracle(xo, nc, hist, fp64, mixed):
    """fp64 / mixed: (host x, SolveResult) of the GPU solves."""
    import numpy as np
    x64, r64 = fp64
    d = {"fp64_bitwise_equal": bool(np.array_equal(x64, xo)), "fp64_cycles_equal": r64.cycles == nc,
         "max_abs_rho_diff": float(np.max(np.abs(r64.history - hist))) if r64.cycles == nc else None}
    if mixed is not None:
        xm, _ = mixed
        diff = xm - xo
        d["mixed_vs_oracle_error"] = {"max_rel": float(np.abs(diff).max() / np.abs(xo).max()),
                                      "l2_rel": float(np.linalg.norm(diff) / np.linalg.norm(xo))}
    return d


# ---------------------------------------------------------------------------
def adjoint_run(n, rng_seed=SEED, log_prefix="adjoint"):
    """Inverse source problem on an n^3 grid: recover f from u_obs, where
    -lap(u) = f (A u = -f), u_obs = u(f*) + N(0, 1e-3) and f* is a sum of six
    Gaussians (centres U[0.2,0.8]^3, widths U[0.03,0.1], amplitudes U[-1,1]).
    Loss 0.5 |u(f) - u_obs|^2 + 1e-4 |grad f|^2 (forward differences / h);
    gradients through the adjoint multigrid solve.  For FP64 and mixed
    precision: the worst central-difference deviation (eps 1e-6, 32 random
    unit directions), the relative l2 error of the gradient against a CPU
    FP64 reference gradient (the oracle's multigrid solves of the forward
    and adjoint problems to 1e-13, regularisation gradient in numpy), the
    same for the data term alone (at a random f0 the regularisation term
    dominates the total gradient; the data term's is the adjoint solve's
    output), and the wall time of 50 L-BFGS iterations
    (torch.optim.LBFGS)."""
    import numpy as np
    import torch
    from oracle import oracle as O
    from peps_mg_amd.adjoint import PoissonSolver, poisson_solve
    levels = levels_for(n)
    rng = np.random.default_rng(rng_seed)
    h = 1.0 / n
    c = (np.arange(n) + 0.5) * h
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    fstar = np.zeros((n, n, n))
    for _ in range(6):
        cx, cy, cz = rng.uniform(0.2, 0.8, 3)
        w = rng.uniform(0.03, 0.1)
        amp = rng.uniform(-1, 1)
        fstar += amp * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2) / (2 * w * w))
    del X, Y, Z
    noise = rng.normal(scale=1e-3, size=(n, n, n))
    f0 = rng.uniform(-1, 1, (n, n, n))
    dirs_seed = int(rng.integers(1 << 31))
    dev = torch.device("cuda:0")

    # CPU FP64 reference: u_obs and the gradient at f0 from oracle solves
    g = O.Grid(n, n, n)
    A = g.matrix()

    def cpu_solve(rhs_interior):
        b = O.padded(n, n, n)
        b[1:-1, 1:-1, 1:-1] = rhs_interior
        x = O.padded(n, n, n)
        O.Multigrid(g, levels, 200).solve_vcycle(x, O.padded(n, n, n), A, b, 2, 1e-13, 1.0)
        return x[1:-1, 1:-1, 1:-1].copy()

    uobs_cpu = cpu_solve(-fstar) + noise
    u0 = cpu_solve(-f0)
    lam = cpu_solve(u0 - uobs_cpu)          # A symmetric (uniform axes): A^-T = A^-1
    grad_data_cpu = -lam                    # d/df of 0.5 |u(f) - u_obs|^2 with A u = -f
    grad_cpu = grad_data_cpu.copy()
    for ax in range(3):
        dd = np.diff(f0, axis=ax) / h
        sl_lo = [slice(None)] * 3
        sl_hi = [slice(None)] * 3
        sl_lo[ax] = slice(None, -1)
        sl_hi[ax] = slice(1, None)
        grad_cpu[tuple(sl_lo)] -= 2e-4 * dd / h
        grad_cpu[tuple(sl_hi)] += 2e-4 * dd / h
    del A

    out = {}
    for prec in (0, 1):
        S = PoissonSolver.uniform(n, levels, tolerance=1e-10, precision=prec)
        ft = torch.tensor(fstar, device=dev)
        with torch.no_grad():
            uobs = S.solve(-ft) + torch.tensor(noise, device=dev)

        def loss_fn(f):
            u = poisson_solve(-f, S)
            reg = sum(((torch.diff(f, dim=d) / h) ** 2).sum() for d in range(3))
            return 0.5 * ((u - uobs) ** 2).sum() + 1e-4 * reg

        f0t = torch.tensor(f0, device=dev, requires_grad=True)
        # the data term alone: its gradient is the adjoint solve's output
        (0.5 * ((poisson_solve(-f0t, S) - uobs) ** 2).sum()).backward()
        gd = f0t.grad.detach().cpu().numpy()
        grad_data_err = float(np.linalg.norm(gd - grad_data_cpu) / np.linalg.norm(grad_data_cpu))
        f0t.grad = None
        loss_fn(f0t).backward()
        gr = f0t.grad.detach()
        grad_err = float(np.linalg.norm(gr.cpu().numpy() - grad_cpu) / np.linalg.norm(grad_cpu))
        eps, worst = 1e-6, 0.0
        drng = np.random.default_rng(dirs_seed)
        with torch.no_grad():
            for _ in range(32):
                d = torch.tensor(drng.normal(size=(n, n, n)), device=dev)
                d /= d.norm()
                fd = (loss_fn(f0t + eps * d) - loss_fn(f0t - eps * d)).item() / (2 * eps)
                worst = max(worst, abs(fd - (gr * d).sum().item()) / gr.norm().item())
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
        log(f"[bench] {log_prefix} {tag}: 50 L-BFGS iterations in {dt:.3f} s ({evals[0]} evaluations), "
            f"loss {l0:.4e} -> {l1:.4e}, FD worst {worst:.2e}, gradient vs CPU reference {grad_err:.2e} "
            f"(data term alone {grad_data_err:.2e})")
        out[tag] = {"lbfgs50_s": round(dt, 4), "evaluations": evals[0], "loss0": l0, "loss50": l1,
                    "rel_error_vs_fstar": rel, "fd_worst_rel": worst, "grad_rel_l2_vs_cpu_reference": grad_err,
                    "data_term_grad_rel_l2_vs_cpu_reference": grad_data_err}
        S.destroy()
        del ft, uobs, f0t, f
        torch.cuda.empty_cache()
    return out


def adjoint_workload(a):
    out = adjoint_run(a.adjoint_n)
    n = a.adjoint_n
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
    from peps_mg_amd import geometry
    from peps_mg_amd.multigrid import Multigrid
    cases = [("rbgs_fp64_persist16", dict(), 2), ("rbgs_fp64_full_coarsening", dict(coarse_switch=0), 2),
             ("cg4_fp64_persist16", dict(smoother=1), 4), ("cg4_fp64_full_coarsening", dict(smoother=1, coarse_switch=0), 4),
             ("rbgs_mixed", dict(precision=1), 2), ("cg4_mixed", dict(smoother=1, precision=1), 4)]
    out = {}
    for n in (256, 512):
        sdm = geometry.geometry_subdomain_create(geometry.geometry_domain_create(n, n, n, 0.0, 0.0, 0.0, 1.0, 1.0,
                                                                                 a.lz))
        d_b = workload_rhs_device(n, a.lz, sdm, torch)
        d_x = torch.zeros_like(d_b)
        stream = torch.cuda.current_stream().cuda_stream
        for label, kw, nu in cases:
            mg = Multigrid(sdm, levels_for(n), 200, **kw)
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


def batch_run(steps, warmup, tune=True):
    """48 independent Dirichlet Poisson problems on unit cubes with edges
    drawn from {32, 48, 64, 96, 128, 192} (seeded), each with its own
    U[-1,1] right-hand side; 8 of them carry 4 right-hand sides (72 solves).
    Coarsest levels 4^3 (edges 32, 64, 128) or 3^3 (48, 96, 192).  The
    batched solver (every operation one launch over all solves, steps
    replayed from a hipGraph; tiles and schedule from the autotuner) in mixed
    precision and in FP64, against looping the single-grid FP64 solver over
    the same 72 solves."""
    import numpy as np
    import torch
    from peps_mg_amd import geometry
    from peps_mg_amd.batch import BatchSolver
    from peps_mg_amd.multigrid import Multigrid
    rng = np.random.default_rng(SEED)
    edges = [int(e) for e in rng.choice([32, 48, 64, 96, 128, 192], size=48)]
    nrhs = [4 if p < 8 else 1 for p in range(48)]
    levels = [levels_for(e) for e in edges]
    sdms = [geometry.geometry_subdomain_create(geometry.geometry_domain_create(e, e, e, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0))
            for e in edges]
    gen = torch.Generator(device="cuda").manual_seed(SEED)
    rhss, sols = [], []
    for e, r in zip(edges, nrhs):
        for _ in range(r):
            b = torch.zeros((e + 2,) * 3, dtype=torch.float64, device="cuda")
            b[1:-1, 1:-1, 1:-1] = torch.rand((e,) * 3, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
            rhss.append(b)
            sols.append(torch.zeros_like(b))
    unknowns = sum(e ** 3 * r for e, r in zip(edges, nrhs))
    log(f"[bench] batch: {len(rhss)} solves, {unknowns / 1e6:.1f} M unknowns, edges {sorted(set(edges))}")

    def timed(fn):
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
    xs = {}
    for prec, tag in ((1, "mixed"), (0, "fp64")):
        B = BatchSolver(sdms, levels, nrhs, ncycle=100, precision=prec)
        if tune:
            from peps_mg_amd.autotune import tune_batch
            tuned[tag] = tune_batch(B, sols, rhss, 2, 1e-10, 1.0)
            log(f"[bench] autotuned {tag}: {tuned[tag]}")
        res = []

        def run():
            for s in sols:
                s.zero_()
            res[:] = B.solve_device(sols, rhss, 2, 1e-10, 1.0, use_graph=True)

        t = timed(run)
        ok = all(r.converged for r in res)
        out["batch_" + tag] = {"s": round(t, 5), "unknowns_per_s": round(unknowns / t, 1), "all_converged": ok,
                               "max_rho": max(r.rho for r in res)}
        xs[tag] = [s.clone() for s in sols]
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

    t_loop = timed(loop)
    log(f"[bench] looped single-grid fp64: {t_loop * 1e3:.2f} ms")
    # batched FP64 items against the looped single-grid solves (same kernels: equal bit for bit)
    fp64_equal = all(torch.equal(a, b) for a, b in zip(xs["fp64"], sols))
    mixed_err = max(((a - b).abs().max() / b.abs().max()).item() for a, b in zip(xs["mixed"], sols))
    for m in mgs.values():
        m.destroy()
    return {"workload": "48 unit cubes, edges from {32,48,64,96,128,192} (seed 20210501), 8 with 4 RHS, "
                        "U[-1,1] RHS, V(2,2) RBGS, tol 1e-10", "solves": len(rhss), "unknowns": unknowns,
            "batch_mixed": out["batch_mixed"], "batch_fp64": out["batch_fp64"], "autotuned": tuned,
            "looped_single_fp64_s": round(t_loop, 5), "looped_single_fp64_unknowns_per_s": round(unknowns / t_loop, 1),
            "speedup_batch_mixed_vs_looped_fp64": round(t_loop / out["batch_mixed"]["s"], 3),
            "batch_fp64_equals_looped_fp64_bitwise": bool(fp64_equal),
            "batch_mixed_max_rel_diff_vs_looped_fp64": mixed_err,
            "note": "right-hand sides are purely random: a solution with relative residual 1e-10 is only ~1e-9 "
                    "(relative) from the exact one, so two such solutions differ by about that much; the tests "
                    "compare mixed items with oracle solves at 1e-12, within 1e-10"}


def batch_workload(a):
    r = batch_run(a.steps, a.warmup, tune=not a.no_tune)
    value = r["unknowns"] / r["batch_mixed"]["s"]
    print(json.dumps(dict({
        "metric": "batched solves: aggregate unknowns/s to 1e-10 rel. residual", "value": round(value, 1),
        "unit": "unknowns/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(r["batch_mixed"]["s"] * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": r["workload"], "solves": r["solves"], "unknowns": r["unknowns"]}}, **r)))


def aniso_run(n, steps, warmup, tune, torch):
    """The 512^3 grid with hx:hy:hz = 1:1:2 (domain 1 x 1 x 2): FP64 with the
    defaults and mixed precision with the defaults and autotuned; returns the
    times and the host copies of the FP64 and (tuned) mixed solutions."""
    from peps_mg_amd import _lib, geometry
    from peps_mg_amd.multigrid import Multigrid
    lz = 2.0
    sdm = geometry.geometry_subdomain_create(geometry.geometry_domain_create(n, n, n, 0.0, 0.0, 0.0, 1.0, 1.0, lz))
    d_b = workload_rhs_device(n, lz, sdm, torch)
    b_host = d_b.cpu().numpy()
    d_x = torch.zeros_like(d_b)
    out = {}
    mg = Multigrid(sdm, levels_for(n), 200)
    t64, r64 = timed_solves(mg, d_x, d_b, steps, warmup, "1:1:2 fp64", torch)
    x64 = d_x.cpu().numpy()
    mg.destroy()
    mg = Multigrid(sdm, levels_for(n), 200, precision=1)
    tmx, rmx = timed_solves(mg, d_x, d_b, steps, warmup, "1:1:2 mixed", torch)
    mg.destroy()
    out["fp64"] = {"s": round(t64, 6), "cycles": r64.cycles, "rho": r64.rho}
    out["mixed_default"] = {"s": round(tmx, 6), "cycles": rmx.cycles, "refine_steps": rmx.refine_steps,
                            "rho": rmx.rho}
    if tune:
        from peps_mg_amd.autotune import tune_solver
        names = {f for f, _ in _lib.Options._fields_}
        opts = tune_solver(sdm, levels_for(n), d_x, d_b, 2, 1e-10, 1.0, precision=1, use_cache=False)
        best = {k: v for k, v in opts.items() if k in names}
        mg = Multigrid(sdm, levels_for(n), 200, **best)
        tmx, rmx = timed_solves(mg, d_x, d_b, steps, warmup, "1:1:2 mixed tuned", torch)
        mg.destroy()
        out["mixed_tuned"] = {"s": round(tmx, 6), "cycles": rmx.cycles, "refine_steps": rmx.refine_steps,
                              "rho": rmx.rho, "options": best}
    xm = d_x.cpu().numpy()   # the solution of the mixed run reported in mixed_s
    out["mixed_s"] = round(tmx, 6)
    out["speedup_mixed_vs_fp64"] = round(t64 / tmx, 4)
    del d_x, d_b
    torch.cuda.empty_cache()
    return out, (x64, r64), (xm, rmx), b_host


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "the solver runs on one GPU; sharding a grid over several GPUs is not "
                                   "implemented", "n_gpus": a.gpus}))
        sys.exit(2)
    if a.kernels_only:
        kernels_only(a)
        return
    if a.cg_profile:
        cg_profile(a)
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
    levels = levels_for(n)
    sdm = geometry.geometry_subdomain_create(geometry.geometry_domain_create(n, n, n, 0.0, 0.0, 0.0, 1.0, 1.0, lz))
    d_b = workload_rhs_device(n, lz, sdm, torch)
    b_host = None if a.no_cpu else d_b.cpu().numpy()
    d_x = torch.zeros_like(d_b)

    log(f"[bench] {n}^3 lz={lz} levels={levels}")
    mg64 = Multigrid(sdm, levels, 200)
    t64, r64 = timed_solves(mg64, d_x, d_b, a.steps, a.warmup, "fp64", torch)
    x64 = d_x.cpu().numpy()
    kw = dict(precision=1)
    if a.inner:
        kw["inner_cycles"] = a.inner
    mgmx = Multigrid(sdm, levels, 200, **kw)
    tmx_def, rmx_def = timed_solves(mgmx, d_x, d_b, a.steps, a.warmup, "mixed", torch)
    # the fine-level smoother inside solves, timed by the library's event
    # hook in solves of their own (each event adds a few microseconds to the
    # stream, so the solve times above are taken without it)
    ev_steps = max(2, a.steps // 2)
    _, r64s = timed_solves(mg64, d_x, d_b, ev_steps, 0, "fp64 (smoother events)", torch, time_smoother=True)
    _, rmxs = timed_solves(mgmx, d_x, d_b, ev_steps, 0, "mixed (smoother events)", torch, time_smoother=True)
    # precision alone: mixed with the FP64 solve's own V(2,2) omega = 1 cycles,
    # and FP64 with the inner cycles' SOR factor 1.35
    mgp = Multigrid(sdm, levels, 200, precision=1, inner_omega=0.0, inner_nu=2, inner_nu_post=2)
    t_prec, r_prec = timed_solves(mgp, d_x, d_b, a.steps, a.warmup, "mixed V(2,2) omega 1", torch)
    mgp.destroy()
    t64w, r64w = timed_solves(mg64, d_x, d_b, a.steps, a.warmup, "fp64 omega 1.35", torch, omega=1.35)
    tuned = None
    tmx, rmx = tmx_def, rmx_def
    mg_best = mgmx
    if not a.no_tune:
        # the autotuner (measured afresh on this GPU, outside the timed region)
        # picks the mixed solver's tile, coarse switch and inner V-cycle
        # schedule; the FP64 solve's options are launch shapes only (same x bit
        # for bit) and its per-level default tiles are the tuned ones
        from peps_mg_amd.autotune import tune_solver
        names = {f for f, _ in _lib.Options._fields_}
        best = tune_solver(sdm, levels, d_x, d_b, 2, 1e-10, 1.0, precision=1, use_cache=False)
        tuned = {k: v for k, v in best.items() if k in names}
        log(f"[bench] autotuned mixed: {tuned}")
        mg_best = Multigrid(sdm, levels, 200, **tuned)
        tmx, rmx = timed_solves(mg_best, d_x, d_b, a.steps, a.warmup, "mixed tuned", torch)
    xmx = d_x.cpu().numpy()

    # fine-level smoother kernels, timed with events on the solve stream
    def smoother(mg, prec):
        avg, nbytes, kn = C.c_double(0), C.c_double(0), C.c_char_p()
        _lib.check(_lib.lib().pepsmg_bench_smoother(mg.h, prec, a.smoother_reps, 1.0, C.byref(avg), C.byref(nbytes),
                                                    C.byref(kn)))
        return avg.value, nbytes.value, kn.value.decode()   # (the library reuses its name buffer)

    bb64, nbytes64, kname64 = smoother(mg64, 0)
    bb32, nbytes32, kname32 = smoother(mgmx, 1)
    # the fine-level smoother inside the timed solves (the library's event hook)
    avg64, n64, kin64 = r64s.smoother
    avg32, n32, kin32 = rmxs.smoother
    assert kin64 == kname64 and kin32 == kname32, (kin64, kname64, kin32, kname32)
    achieved = nbytes64 / (avg64 * 1e-3) / 1e9
    achieved32 = nbytes32 / (avg32 * 1e-3) / 1e9
    # the bandwidth a kernel with the smoother's streams (x, b in; x out) and no
    # stencil reaches on this GPU: the practical ceiling for the smoother
    s_ms, s_bytes = C.c_double(0), C.c_double(0)
    _lib.check(_lib.lib().pepsmg_bench_stream(n ** 3, 10, C.byref(s_ms), C.byref(s_bytes)))
    stream_gbs = s_bytes.value / (s_ms.value * 1e-3) / 1e9
    mg64.destroy()
    mgmx.destroy()
    if mg_best is not mgmx:
        mg_best.destroy()
    del d_x, d_b
    torch.cuda.empty_cache()

    traffic, traffic_note, tr, trace, trace_note = None, "not measured (--no-traffic)", None, None, "not measured"
    if not a.no_traffic:
        tr, err = measure_traffic(a, kname64)
        if tr:
            traffic = tr["FETCH_SIZE"] + tr["WRITE_SIZE"]
            traffic_note = (f"rocprofv3 FETCH_SIZE x {tr['FETCH_SIZE_scale']:.3f} + WRITE_SIZE x "
                            f"{tr['WRITE_SIZE_scale']:.3f} (scales from a 1 GiB device copy)")
        else:
            traffic_note = err
        trace, trace_note = trace_smoothers(a, [kname64, kname32])

    extras = {}
    aniso_x = None
    if not a.no_extras:
        extras["aniso_512"], a64, amx, ab = aniso_run(n, max(1, a.steps // 2), 1, not a.no_tune, torch)
        aniso_x = (a64, amx, ab)
        extras["batch_cfg3"] = batch_run(max(1, a.steps // 2), 1, tune=not a.no_tune)
        torch.cuda.empty_cache()
        extras["adjoint_128"] = adjoint_run(128)
        torch.cuda.empty_cache()

    cpu, parity = None, None
    if not a.no_cpu:
        log("[bench] cpu_baseline: oracle FP64 solve of the same workload")
        xo, nco, histo, res0o, dto, threads = oracle_solve(n, lz, b_host)
        cpu = {"value": round(dto, 3), "unit": "s", "cores": threads, "kind": "port",
               "sample": f"one complete FP64 solve of the {n}^3 workload to 1e-10 by the oracle (C restatement of "
                         f"the reference loops, AoS coefficients, OpenMP {threads} threads): {nco} V(2,2) cycles "
                         f"in {dto:.2f} s; set-up (matrix, right-hand side) not timed"}
        parity = compare_to_oracle(xo, nco, histo, (x64, r64), (xmx, rmx))
        log(f"[bench] oracle: {nco} cycles in {dto:.2f} s; parity {parity}")
        del xo, b_host
        if aniso_x is not None:
            log("[bench] oracle FP64 solve of the 1:1:2 grid")
            xo, nco, histo, _, dta, _ = oracle_solve(n, 2.0, aniso_x[2])
            extras["aniso_512"]["oracle"] = dict(compare_to_oracle(xo, nco, histo, aniso_x[0], aniso_x[1]),
                                                 oracle_cycles=nco, oracle_solve_s=round(dta, 3))
            del xo
    if parity is None:
        # no oracle leg: the mixed solution against the GPU FP64 one
        diff = xmx - x64
        mixed_err = {"reference": "GPU FP64 solve", "max_rel": float(np.abs(diff).max() / np.abs(x64).max()),
                     "l2_rel": float(np.linalg.norm(diff) / np.linalg.norm(x64))}
    else:
        mixed_err = dict(parity["mixed_vs_oracle_error"], reference="oracle FP64 solve")

    unknowns = n ** 3
    sched = (f"mixed: FP32 V({tuned.get('inner_nu', 2)},{tuned.get('inner_nu_post', 0) or tuned.get('inner_nu', 2)}) "
             f"inner cycles, SOR {tuned.get('inner_omega', 1.35):g}, {tuned.get('inner_cycles', 2)} per FP64 "
             f"refinement step (autotuned)" if tuned else
             "mixed: FP32 V(2,2) inner cycles, SOR 1.35, 2 per FP64 refinement step (defaults)")
    def frac(nbytes, ms):
        return round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)

    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": round(traffic) if traffic else None, "traffic_note": traffic_note,
            "traffic_read": round(tr["FETCH_SIZE"]) if tr else None,
            "traffic_write": round(tr["WRITE_SIZE"]) if tr else None,
            "kernel": kname64, "kernel_down": kname64.replace("k_rb_fused<", "k_rb_fused_down<"),
            "avg_ms": avg64, "launches": n64, "bytes_per_launch": nbytes64,
            "timing": f"HIP events on the solve stream around every plain iteration on the {n}^3 level inside "
                      f"{ev_steps} FP64 solves run after the timed ones (both kernels: a plain iteration that "
                      "follows an upward launch on the level marches down)",
            "stream_ceiling_gbs": round(stream_gbs, 1), "frac_of_stream_ceiling": round(achieved / stream_gbs, 4),
            "back_to_back": {"avg_ms": bb64, "frac": frac(nbytes64, bb64),
                             "note": f"{a.smoother_reps} launches in a row after {a.smoother_reps} untimed ones "
                                     "(pepsmg_bench_smoother, alternating upward and downward), events around the whole run"}}
    roof32 = {"achieved": round(achieved32, 1), "frac": round(achieved32 / HBM_PEAK_GBS, 4), "kernel": kname32,
              "kernel_down": kname32.replace("k_rb_fused<", "k_rb_fused_down<"),
              "avg_ms": avg32, "launches": n32, "bytes_per_launch": nbytes32,
              "timing": f"HIP events around every launch on the {n}^3 level inside {ev_steps} mixed solves "
                        "(default options) run after the timed ones",
              "back_to_back": {"avg_ms": bb32, "frac": frac(nbytes32, bb32)}}
    if trace:
        t64r, t32r = trace[kname64], trace[kname32]
        roof["back_to_back"].update({
            "trace_avg_ms": t64r["trace_avg_ms"], "trace_event_avg_ms": t64r["event_avg_ms_same_run"],
            "trace_span_per_launch_ms": t64r["trace_span_per_launch_ms"],
            "trace_avg_all_dispatches_ms": t64r["trace_avg_all_ms"],
            "frac_trace": frac(nbytes64, t64r["trace_avg_ms"]),
            "trace_note": f"rocprofv3 --kernel-trace of the {a.smoother_reps} event-timed launches in a child run of "
                          "--kernels-only (trace_event_avg_ms: the HIP events of that run)"})
        roof32["back_to_back"].update({"trace_avg_ms": t32r["trace_avg_ms"],
                                       "trace_event_avg_ms": t32r["event_avg_ms_same_run"],
                                       "frac_trace": frac(nbytes32, t32r["trace_avg_ms"])})
    elif not a.no_traffic:
        roof["trace_note"] = trace_note
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
                               f"tol 1e-10, x0=0, {levels} coarse levels to 4^3; fp64: V(2,2) RBGS omega=1 (the "
                               f"reference's cycle); {sched}",
                   "grid": [n, n, n], "levels": levels, "precision": "fp32 V-cycles in fp64 refinement"},
        "tuned_options": {"mixed": tuned} if tuned else None,
        "default_options_times": {"fp64_time_s": round(t64, 6), "mixed_time_s": round(tmx_def, 6),
                                  "mixed_cycles": rmx_def.cycles},
        "fp64_time_s": round(t64, 6),
        "fp64_cycles": r64.cycles,
        "mixed_time_s": round(tmx, 6),
        "mixed_cycles": rmx.cycles,
        "mixed_refine_steps": rmx.refine_steps,
        "speedup_mixed_vs_fp64": round(t64 / tmx, 4),
        "precision_only": {
            "mixed_inner_v22_omega1_s": round(t_prec, 6), "mixed_inner_v22_omega1_cycles": r_prec.cycles,
            "speedup_vs_fp64": round(t64 / t_prec, 4),
            "fp64_omega135_s": round(t64w, 6), "fp64_omega135_cycles": r64w.cycles,
            "note": "mixed with the FP64 solve's own V(2,2) omega=1 cycles isolates the gain of precision; "
                    "FP64 with omega=1.35 shows what the SOR factor alone buys"},
        "fp64_vcycles_per_s": round(r64.cycles / t64, 3),
        "fp64_unknowns_per_s": round(unknowns / t64, 1),
        "mixed_unknowns_per_s": round(unknowns / tmx, 1),
        "fp64_cell_vcycles_per_s": round(unknowns * r64.cycles / t64, 1),
        "vcycle_roofline": vcycle_roofline(n, levels, t64, r64.cycles),
        "rho_fp64": r64.rho,
        "rho_mixed": rmx.rho,
        "mixed_vs_fp64_error": mixed_err,
        "oracle_parity": parity,
        "roofline": roof,
        "roofline_fp32_smoother": roof32,
        "cpu_baseline": cpu,
    }
    out.update(extras)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
