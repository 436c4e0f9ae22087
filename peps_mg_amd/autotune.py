"""Autotuner for the solver's launch and schedule options.

Everything is measured on the GPU through the C-ABI; nothing is modelled:

* ``tune_tiles`` times the fine-level fused smoother of one grid
  (``pepsmg_bench_smoother``, HIP events) for each tile shape and prefetch
  depth and returns the fastest;
* ``tune_solver`` times complete solves over a small space of schedule
  options (FP32 V-cycles per refinement step, their pre- and post-smoothing
  steps and SOR factor, size of the levels handed to the persistent coarse
  kernel) with the tuned tile;
* ``tune_batch`` gives every problem of a batch the tile tuned for its grid
  size (one measurement per distinct size) and then picks the schedule
  options that make the whole batch fastest.

Results are cached per grid size and precision in a JSON file
(``$PEPSMG_TUNE_CACHE`` or ``~/.cache/pepsmg_amd/autotune.json``), so a size
is measured once per machine.  Nothing here changes numerical results of
the FP64 path: tiles and prefetch depth only change the launch shape.
"""
import ctypes as C
import json
import os
import time

from . import _lib

TILES_FP64 = [(64, 32, 1), (64, 16, 1), (64, 8, 1), (64, 16, 2)]
TILES_FP32 = [(128, 32, 1), (128, 16, 1), (128, 8, 1), (64, 8, 1), (64, 16, 1), (128, 16, 2), (64, 32, 1)]


def _cache_path():
    return os.environ.get("PEPSMG_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "pepsmg_amd", "autotune.json"))


def _load():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(db):
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def _key(kind, sdm, precision, extra=""):
    return f"{kind}:{sdm.nx}x{sdm.ny}x{sdm.nz}:p{precision}{extra}"


def tune_tiles(sdm, nlevel, precision, reps=10, use_cache=True):
    """Fastest (tile_x, tile_y, fused_smoother) for the fine-level smoother
    of this grid at this precision; returns a dict of option values."""
    from .multigrid import Multigrid
    db = _load() if use_cache else {}
    key = _key("tile", sdm, precision)
    if key in db:
        return db[key]
    mg = Multigrid(sdm, nlevel, 1, precision=precision)
    best, best_ms, timings = None, float("inf"), {}
    for tx, ty, pf in (TILES_FP32 if precision == 1 else TILES_FP64):
        mg.set_options(tile_x=tx, tile_y=ty, fused_smoother=pf, precision=precision)
        avg = C.c_double(0)
        nb = C.c_double(0)
        kn = C.c_char_p()
        _lib.check(_lib.lib().pepsmg_bench_smoother(mg.h, precision, reps, 1.0, C.byref(avg), C.byref(nb),
                                                    C.byref(kn)))
        timings[f"{tx}x{ty}/pf{pf}"] = avg.value
        if avg.value < best_ms:
            best_ms, best = avg.value, {"tile_x": tx, "tile_y": ty, "fused_smoother": pf}
    mg.destroy()
    best = dict(best, smoother_ms=best_ms, timings=timings)
    if use_cache:
        db = _load()
        db[key] = best
        _store(db)
    return best


def _opts(d):
    return {k: d[k] for k in ("tile_x", "tile_y", "fused_smoother") if k in d}


def tune_solver(sdm, nlevel, d_sol, d_rhs, maxiteration=2, tolerance=1e-10, omega=1.0, precision=1,
                inner_cycles=(2, 3), inner_nu=(1, 2), inner_nu_post=(1, 2), inner_omega=(1.2, 1.35),
                coarse_switch=(512, 4096, 32768), repeats=2, use_cache=True):
    """Fastest schedule options for complete solves of this grid; d_sol /
    d_rhs are padded device arrays (torch tensors; d_sol is zeroed before
    each timed solve).  Returns the option dict (tile options included)."""
    import torch
    from .multigrid import Multigrid
    db = _load() if use_cache else {}
    key = _key("solve", sdm, precision, f":nl{nlevel}:nu{maxiteration}:v2")
    if key in db:
        return db[key]
    tiles = _opts(tune_tiles(sdm, nlevel, precision, use_cache=use_cache))
    mg = Multigrid(sdm, nlevel, 200, precision=precision, **tiles)
    best, best_t, timings = None, float("inf"), {}
    dflt = _lib.default_options()
    sched = ([(ic, nu, nup, om) for ic in inner_cycles for nu in inner_nu for nup in inner_nu_post
              for om in inner_omega] if precision == 1
             else [(dflt.inner_cycles, dflt.inner_nu, dflt.inner_nu_post, dflt.inner_omega)])
    for cs in coarse_switch:
        for ic, nui, nup, om in sched:
            try:
                mg.set_options(coarse_switch=cs, inner_cycles=ic, inner_nu=nui, inner_nu_post=nup, inner_omega=om)
            except _lib.PepsmgError:
                continue
            t = float("inf")
            for _ in range(repeats + 1):
                d_sol.zero_()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                r = mg.solve_vcycle_device(d_sol.data_ptr(), 0, d_rhs.data_ptr(), maxiteration, tolerance, omega)
                torch.cuda.synchronize()
                if r.converged:
                    t = min(t, time.perf_counter() - t0)
            timings[f"cs{cs}/ic{ic}/nu{nui}-{nup}/om{om}"] = t
            if t < best_t:
                best_t, best = t, dict(tiles, coarse_switch=cs, inner_cycles=ic, inner_nu=nui, inner_nu_post=nup,
                                       inner_omega=om, precision=precision)
    mg.destroy()
    best = dict(best, solve_s=best_t, timings=timings)
    if use_cache:
        db = _load()
        db[key] = best
        _store(db)
    return best


def tune_batch(batch, sols, rhss, maxiteration=2, tolerance=1e-10, omega=1.0, inner_cycles=(2, 3),
               inner_nu=(1, 2), inner_nu_post=(1, 2), inner_omega=(1.2, 1.35), repeats=1, use_cache=True):
    """Per-problem tiles by grid size, then the inner-cycle schedule that
    makes the whole batch fastest; applies both to `batch` (a BatchSolver)
    and returns {'tiles': {size: tile}, 'inner_cycles': n, 'inner_nu': n,
    'inner_nu_post': n, 'inner_omega': w, 'timings': {...}}."""
    import torch
    precision = batch.options.precision
    tiles_by_size = {}
    tx, ty = [], []
    for s, nl in zip(batch.sdms, batch.nlevels):
        key = (s.nx, s.ny, s.nz)
        if key not in tiles_by_size:
            tiles_by_size[key] = tune_tiles(s, nl, precision, use_cache=use_cache)
        t = tiles_by_size[key]
        tx.append(t["tile_x"])
        ty.append(t["tile_y"])
    batch.set_tiles(tx, ty)
    timings, best, best_t = {}, None, float("inf")
    o = batch.options
    sched = ([(ic, nu, nup, om) for ic in inner_cycles for nu in inner_nu for nup in inner_nu_post
              for om in inner_omega] if precision == 1
             else [(o.inner_cycles, o.inner_nu, o.inner_nu_post, o.inner_omega)])
    for ic, nui, nup, om in sched:
        batch.set_options(inner_cycles=ic, inner_nu=nui, inner_nu_post=nup, inner_omega=om)
        t = float("inf")
        for _ in range(repeats + 1):
            for x in sols:
                x.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            res = batch.solve_device(sols, rhss, maxiteration, tolerance, omega)
            torch.cuda.synchronize()
            if all(r.converged for r in res):
                t = min(t, time.perf_counter() - t0)
        timings[f"ic{ic}/nu{nui}-{nup}/om{om}"] = t
        if t < best_t:
            best_t, best = t, (ic, nui, nup, om)
    batch.set_options(inner_cycles=best[0], inner_nu=best[1], inner_nu_post=best[2], inner_omega=best[3])
    return {"tiles": {f"{k[0]}x{k[1]}x{k[2]}": {"tile_x": v["tile_x"], "tile_y": v["tile_y"]}
                      for k, v in tiles_by_size.items()},
            "inner_cycles": best[0], "inner_nu": best[1], "inner_nu_post": best[2], "inner_omega": best[3],
            "timings": timings}
