"""Autotuner for the fused stage kernel's tiling.

The stage kernel walks x rows in segments; the tile shape that matters is
the number of rows per wavefront (the y strip is fixed at 60 cells by the
64-lane wavefront).  Too few rows repeat the 4-row halo prologue and the
extra x face too often; too many leave wavefront slots idle or unbalanced.
The tuner times candidate row counts on the caller's own state (device
loop, graph replay), keeps the fastest, and caches it per
(device, nx, ny, precision, Powell) -- or per batch signature -- in a JSON
file (``$AMHD_TUNE_CACHE``, default ``~/.cache/aeromhd/tune.json``).
Row count 0 is the built-in rule (one segment per resident wavefront slot).
"""
from __future__ import annotations

import hashlib
import json
import os
import time

DEFAULT_CANDIDATES = (0, 2, 4, 6, 8, 12, 16, 24, 32)


def cache_path() -> str:
    return os.environ.get("AMHD_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "aeromhd",
                                                          "tune.json"))


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key: str, value: dict) -> None:
    data = _load()
    data[key] = value
    path = cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _device() -> str:
    import torch

    return torch.cuda.get_device_name(0)


def single_key(cfg) -> str:
    powell = int(bool(cfg.powell_source and cfg.B_transverse > 0))
    return f"single|{_device()}|{cfg.nx}x{cfg.ny}|{getattr(cfg, 'precision', 'fp64')}|powell{powell}"


def batch_key(cfgs, precision: str) -> str:
    sig = ";".join(f"{c.nx}x{c.ny}:{int(bool(c.powell_source and c.B_transverse > 0))}" for c in cfgs)
    return f"batch|{_device()}|{precision}|{hashlib.sha1(sig.encode()).hexdigest()[:16]}"


def _time(fn, reps: int) -> float:
    import torch

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fn(reps)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def tune_single(solver, candidates=DEFAULT_CANDIDATES, steps: int = 12, use_cache: bool = True) -> int:
    """Pick the row count for `solver` (an MHDSolver) and apply it.  The
    timed steps advance a copy of the state that is restored afterwards."""
    key = single_key(solver.cfg)
    if use_cache:
        hit = _load().get(key)
        if hit is not None:
            solver.ctx.set_rows(int(hit["rows"]))
            return int(hit["rows"])
    U0 = solver.U.clone()
    t0, s0 = solver.t, solver.step
    cfg = solver.cfg
    saved = (cfg.t_end, cfg.max_steps)
    cfg.t_end, cfg.max_steps = 1e300, 2 ** 31 - 1
    results = {}
    try:
        for rows in candidates:
            if rows > cfg.nx:
                continue
            solver.ctx.set_rows(rows)
            solver.U = U0
            solver.t, solver.step = t0, s0
            solver.advance(2)   # capture the graph, warm caches
            results[rows] = _time(lambda n: solver.advance(n), steps)
    finally:
        cfg.t_end, cfg.max_steps = saved
        solver.U = U0
        solver.t, solver.step = t0, s0
    best = min(results, key=results.get)
    solver.ctx.set_rows(best)
    _store(key, {"rows": best, "ms_per_step": {str(k): v * 1e3 for k, v in results.items()}})
    return best


def tune_batch(batch, candidates=DEFAULT_CANDIDATES, steps: int = 6, use_cache: bool = True) -> int:
    """Pick the task row count for a BatchSolver; the timed steps run on
    copies of the instance states, restored afterwards."""
    key = batch_key(batch.cfgs, batch.precision)
    if use_cache:
        hit = _load().get(key)
        if hit is not None:
            batch.set_rows(int(hit["rows"]))
            return int(hit["rows"])
    saved = [batch.download(k) for k in range(batch.n)]
    t0, s0 = batch.t.copy(), batch.steps.copy()
    limits = [(c.t_end, c.max_steps) for c in batch.cfgs]
    for c in batch.cfgs:
        c.t_end, c.max_steps = 1e300, 2 ** 31 - 1
    results = {}
    try:
        for rows in candidates:
            batch.set_rows(rows)
            batch.run(1)
            results[rows] = _time(lambda n: batch.run(n), steps)
    finally:
        for c, (te, ms) in zip(batch.cfgs, limits):
            c.t_end, c.max_steps = te, ms
        for k, U in enumerate(saved):
            batch.upload(k, U)
        batch.t, batch.steps = t0, s0
    best = min(results, key=results.get)
    batch.set_rows(best)
    _store(key, {"rows": best, "ms_per_step": {str(k): v * 1e3 for k, v in results.items()}})
    return best
