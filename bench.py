#!/usr/bin/env python3
"""Throughput of the gfx950 MHD update on one GPU.

Workload (default): the magnetic-sieve RMI on a 2048 x 1024 grid in FP32 --
MUSCL/van Leer + HLLD + GLM psi transport + Powell source + SSP-RK3 + CFL
reduction, dt computed on the device every step.  A "step" is one full
SSP-RK3 step of the whole grid (three fused stage launches; the first and
the third carry the step control).  Inputs are resident in HBM before the
timed region.

Prints ONE JSON line: Mcell-updates/s, the roofline of the fused stage
kernel (algorithmic bytes 8 * 10 * sizeof(real) per cell-update divided by
the stage kernels' device time, HIP events filled by the stage dispatches on
the context stream over the timed steps, replayed from the state they start
from), and the
CPU oracle timed on this host on a bounded sample.
Secondary records: the 1024x512 RMI in FP64 and FP32, the batched form (64
instances), the x-slab decomposition (4 slabs with halo exchange, one
GPU), and the learned closure's fine-equivalent rate over the FP64 fine
solve with its mixing-width and density errors.

  python bench.py [--steps K] [--warmup W] [--workload sieve2048|sieve2048-fp64|rmi1024-fp64|rmi1024-fp32]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "aeromhd-cuda_amd")
sys.path.insert(0, PKG)

METRIC = "Mcell-updates/s per RK3 step + % HBM roofline; closure speedup at equal L1 err"
HBM_PEAK_GBS = 8000.0
NVAR = 10


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name: str, seed: int):
    from amhd_ics import magnetic_sieve, rmi_uniform_bx

    if name in ("sieve2048", "sieve2048-fp64"):
        prec = "fp64" if name.endswith("fp64") else "fp32"
        cfg, W = magnetic_sieve(2048, 1024, precision=prec, seed=seed)
        desc = (f"magnetic-sieve RMI 2048x1024 {prec.upper()} (seed {seed}): MUSCL+HLLD+GLM+Powell, SSP-RK3, "
                "device dt")
    elif name in ("rmi1024-fp64", "rmi1024-fp32"):
        prec = name.split("-")[1]
        cfg, W = rmi_uniform_bx(1024, 512, precision=prec)
        desc = f"single-mode RMI 1024x512 {prec.upper()} (Mach 10, A=0.5, beta=10 Bx): SSP-RK3, device dt"
    else:
        raise SystemExit(f"unknown workload {name}")
    cfg.t_end = 1e30
    cfg.max_steps = 2 ** 30
    return cfg, W, desc


def gpu_measure(name: str, steps: int, warmup: int, seed: int, rows: int | None, packing: int | None = None,
                tune: bool = False):
    import torch
    from physics import prim_to_cons
    from solver import MHDSolver

    cfg, W, desc = make_workload(name, seed)
    s = MHDSolver(cfg)
    if rows:
        s.ctx.set_rows(rows)
    if packing is not None:
        s.ctx.set_packing(packing)
    s.U = prim_to_cons(W, cfg.gamma)
    if tune:
        import amhd_tune

        amhd_tune.tune_single(s, use_cache=False)   # times the candidates on a copy of this state
    s.advance(warmup)
    # the state the timed steps start from (FP32 values round-trip exactly
    # through the float64 copy), so that the roofline pass times the same steps
    snap, t_snap, step_snap = s.ctx.download(), s.t, s.step
    # timed region 1 (value): the product loop -- device dt, hipGraph replay
    torch.cuda.synchronize()
    g0 = time.perf_counter()
    n = s.advance(steps, use_graph=True)
    torch.cuda.synchronize()
    gwall = time.perf_counter() - g0
    assert n == steps, "device loop stopped early"
    t_graph = s.t
    # timed region 2 (roofline): the same steps again from the same state,
    # launched without the graph, each launch on the context stream carrying
    # a HIP event pair that its dispatch fills with the kernel's start and end
    s.ctx.upload(snap)
    s.t, s.step = t_snap, step_snap
    del snap
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t, step, last, stage_ms, _ = s.ctx.run_profiled(steps, cfg.cfl, cfg.t_end, cfg.max_steps, s.t, s.step)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    assert step - s.step == steps, "device loop stopped early"
    s.t, s.step = t, step
    cells = cfg.nx * cfg.ny
    esize = 4 if cfg.precision == "fp32" else 8
    bytes_step = 8 * NVAR * esize * cells
    kern_s = sum(stage_ms) * 1e-3 / steps
    sums = s.global_sums()
    return {
        "desc": desc, "cfg": cfg, "cells": cells, "esize": esize, "rows": s.ctx.rows,
        "mcups": cells * n / gwall / 1e6, "ms_per_step": gwall * 1e3 / n,
        "profiled_mcups": cells * steps / wall / 1e6,
        "stage_avg_us": [1e3 * m / steps for m in stage_ms],
        "achieved_gbs": bytes_step / kern_s / 1e9, "bytes_per_step": bytes_step,
        "bytes_per_launch": [2 * NVAR * esize * cells, 3 * NVAR * esize * cells, 3 * NVAR * esize * cells],
        "finite": all(x == x for x in sums), "t": t, "dt": last,
        "same_steps": t == t_graph,   # the roofline pass ended where the timed graph run did
    }


def batch_measure(n: int, steps: int, warmup: int, seed: int, seq_steps: int = 20):
    """Batched sweep: n seeded single-mode RMI instances (grids 256x128 ..
    1024x512, Mach U[2,12], Atwood U[0.2,0.8], beta log-U[1,100]) in FP32,
    one persistent launch per RK stage, graph-replayed; compared with the
    same instances advanced one after another by their own device loops."""
    import torch
    from amhd_batch import BatchSolver
    from amhd_ics import batched_instances, cons_from_prim_host
    from solver import MHDSolver
    import amhd_tune

    insts = batched_instances(n=n, seed=seed, precision="fp32", max_steps=2 ** 30)
    cases = []
    for cfg, W in insts:
        cfg.t_end = 1e30
        cases.append((cfg, cons_from_prim_host(W, cfg.gamma)))
    del insts
    b = BatchSolver(cases)
    rows = amhd_tune.tune_batch(b, candidates=(0, 8, 12, 16, 20, 26, 32), steps=3, use_cache=False)
    b.run(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    took = b.run(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    assert int(took.min()) == steps
    cells = b.cell_count
    info = b.info()
    b.close()
    # the same instances one at a time (device loop + graph each)
    seq_wall = 0.0
    for cfg, U in cases:
        s = MHDSolver(cfg)
        s.U = U
        s.advance(2)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        s.advance(seq_steps)
        torch.cuda.synchronize()
        seq_wall += time.perf_counter() - t1
        s.ctx.close()
    return {"workload": f"batched sweep: {n} single-mode RMI instances, grids 256x128..1024x512 (seed {seed}), "
                        f"FP32, one persistent launch per RK stage, hipGraph replay",
            "value": round(cells * steps / wall / 1e6, 2), "unit": "Mcell-updates/s", "dtype": "f32",
            "steps": steps, "ms_per_step": round(wall * 1e3 / steps, 4), "cells": cells,
            "tasks": info["tasks"], "rows_per_task": rows or info["rows"],
            "sequential_value": round(cells * seq_steps / seq_wall / 1e6, 2)}


def slab_measure(n_slabs: int = 4, steps: int = 40, warmup: int = 3, seed: int = 0):
    """The x-slab decomposition of the 2048x1024 FP32 sieve into n_slabs
    slabs (512x1024 each for 4: one GPU's share on 4 GPUs), all run on this
    GPU with a halo exchange before every RK stage and the dt taken over all
    slabs; host-driven (no graph), so it also carries the per-stage host
    synchronisations a multi-GPU step would have."""
    import torch
    from amhd_ics import cons_from_prim_host, magnetic_sieve
    from amhd_slab import SlabSolver

    cfg, W = magnetic_sieve(2048, 1024, precision="fp32", seed=seed, max_steps=10 ** 6)
    cfg.t_end = 1e30
    sl = SlabSolver(cfg, n_slabs)
    sl.upload(cons_from_prim_host(W, cfg.gamma))
    sl.advance(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n = sl.advance(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    i0, i1 = sl.bounds[0]
    return {"workload": f"x-slab decomposition of the 2048x1024 FP32 sieve: {n_slabs} slabs of {i1 - i0}x1024 "
                        "on one GPU, halo exchange before each RK stage, global dt (host-driven)",
            "value": round(cfg.nx * cfg.ny * n / wall / 1e6, 2), "unit": "Mcell-updates/s",
            "ms_per_step": round(wall / n * 1e3, 4), "steps": n}


def closure_measure(seeds=tuple(range(200, 208)), fine_steps: int = 2000):
    """The learned closure in its batched form: 8 held-out magnetic-sieve
    seeds as one BatchClosure (coarse 256x128 FP32 instances, one batched
    step, network pass and limited correction per step), each compared with
    its fine 2048x1024 run restricted to the coarse grid.

    Error-matched comparison: the plain solver (no network) runs the same
    seeds as batches at 256x128, 512x256 and 1024x512, scored on the
    256x128 grid the same way.  The speed-up at equal error is the wall
    time of the cheapest plain resolution whose error is at most the
    closure's, over the closure's wall time (a lower bound when no plain
    resolution reaches it); it is given for the integral mixing width (the
    closure's target), for the L1 density error (which neither the network's
    mixing flux nor the sharpening changes) and for the mean |C - C_fine|
    (the field both correct).  The fine-equivalent
    rate over the FP64 fine solve is kept as a secondary figure.  The
    limited anti-diffusion of C without the network (the closure's
    sharpening pass at its own calibrated coefficient) is measured the same
    way beside it."""
    import amhd_closure as cl

    model = cl.FluxCNN.load(os.path.join(PKG, "closure_weights.npz"))
    bt = cl.evaluate_batched(list(seeds), model, steps=fine_steps)
    res = cl.evaluate_resolutions(list(seeds), factors=(4, 2), steps=fine_steps)
    fp64 = cl.fp64_fine_throughput(seeds[0], fine_steps)
    fine_cells = 2048 * 1024 * fine_steps * len(seeds)
    plain = [("256x128", bt["wall_plain_s"], bt["mw_err_plain_mean"], bt["l1rho_plain_mean"], bt["l1c_plain_mean"])]
    for tag in ("512x256", "1024x512"):
        plain.append((tag, res[tag]["wall_s"], res[tag]["mw_err_mean"], res[tag]["l1rho_mean"], res[tag]["l1c_mean"]))
    wc, mwc, l1r = bt["wall_closure_s"], bt["mw_err_closure_mean"], bt["l1rho_closure_mean"]

    def equal_error(idx, err_c, wall=None):
        """The plain solve of the same error, from the grids tested (cost
        rises and error falls with resolution): an exact match gives its
        time; an error between two grids' brackets the speed-up by theirs;
        an error below every grid's makes the finest one's time a lower
        bound, and above every grid's the coarsest one's an upper bound."""
        wall = wc if wall is None else wall
        for k, p in enumerate(plain):
            if p[idx] <= err_c * (1 + 1e-6):
                if p[idx] >= err_c * (1 - 1e-6):
                    return {"plain_grid": p[0], "speedup": round(p[1] / wall, 3), "bound": "exact match"}
                if k == 0:
                    return {"plain_grid": p[0], "speedup_at_most": round(p[1] / wall, 3),
                            "bound": "the coarsest plain grid tested is already below this error"}
                q = plain[k - 1]
                return {"plain_grid": f"between {q[0]} and {p[0]}",
                        "speedup_between": [round(q[1] / wall, 3), round(p[1] / wall, 3)],
                        "bound": "the plain grid of equal error lies between these two, its time between theirs"}
        p = plain[-1]
        return {"plain_grid": p[0], "speedup_at_least": round(p[1] / wall, 3),
                "bound": "no plain grid tested reaches this error; the finest one's time is a lower bound"}

    return {"workload": f"learned closure, batched: {len(seeds)} held-out magnetic-sieve seeds "
                        f"({seeds[0]}-{seeds[-1]}) in one BatchClosure, coarse 256x128 FP32 + BF16-MFMA CNN, "
                        "against plain batches at 256x128, 512x256 and 1024x512 and the FP64 fine 2048x1024 solve",
            "equal_mixing_width_error": equal_error(2, mwc),
            "equal_l1rho_error": equal_error(3, l1r),
            "equal_l1c_error": equal_error(4, bt["l1c_closure_mean"]),
            "mw_err_closure": round(mwc, 4), "l1rho_closure": round(l1r, 5),
            "mw_target_5pct_met": bool(mwc <= 0.05),
            "plain": {t: {"wall_s": round(w, 4), "mw_err": round(m, 4), "l1rho": round(l, 5), "l1c": round(lc, 6)}
                      for t, w, m, l, lc in plain},
            "fine_equiv_over_fp64_fine": round(fine_cells / wc / 1e6 / fp64["mcups"], 2),
            "fp64_fine_mcups": round(fp64["mcups"], 1),
            "fallbacks": bt["fallbacks_closure"], "wall_closure_s": round(wc, 4),
            "wall_closure_runs_s": bt["wall_closure_runs_s"],
            **({"sharpen_alone": {"kappa": bt["kappa_alone"], "mw_err": round(bt["mw_err_sharpen_mean"], 4),
                                  "l1rho": round(bt["l1rho_sharpen_mean"], 5),
                                  "l1c": round(bt["l1c_sharpen_mean"], 5), "wall_s": round(bt["wall_sharpen_s"], 4),
                                  "equal_mixing_width_error": equal_error(2, bt["mw_err_sharpen_mean"],
                                                                          bt["wall_sharpen_s"]),
                                  "equal_l1c_error": equal_error(4, bt["l1c_sharpen_mean"], bt["wall_sharpen_s"]),
                                  "fine_equiv_over_fp64_fine":
                                      round(fine_cells / bt["wall_sharpen_s"] / 1e6 / fp64["mcups"], 2)}}
               if "wall_sharpen_s" in bt else {}),
            "l1c_closure": round(bt["l1c_closure_mean"], 5),
            "timing": "median of 3 timed runs after an untimed 2-step run; FP64 fine timed after 5 warm-up steps"}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(seed: int, steps: int, threads: int, steps_1core: int = 2):
    """The CPU oracle (C restatement of the reference math, FP64, OpenMP
    over x rows) on the 2048 x 1024 sieve: `steps` SSP-RK3 steps with the
    reference's per-step dt on `threads` threads, and `steps_1core` on one."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    from amhd_ics import cons_from_prim_host, magnetic_sieve

    oracle.build()
    cfg, W = magnetic_sieve(2048, 1024, precision="fp64", seed=seed)
    cfg.t_end = 1e30
    U0 = cons_from_prim_host(W, cfg.gamma)
    oracle.set_threads(threads)
    U, *_ = oracle.run_loop(U0, cfg, 1)          # untimed first step (page-in)
    t0 = time.perf_counter()
    _, _, n, _, _ = oracle.run_loop(U, cfg, steps)
    wall = time.perf_counter() - t0
    oracle.set_threads(1)
    t1 = time.perf_counter()
    _, _, n1, _, _ = oracle.run_loop(U, cfg, steps_1core)
    wall1 = time.perf_counter() - t1
    cells = cfg.nx * cfg.ny
    return {"value": round(cells * n / wall / 1e6, 3), "unit": "Mcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"{n} SSP-RK3 steps (with per-step CFL dt) of the 2048x1024 magnetic-sieve RMI in FP64 "
                      f"(oracle/oracle.c, gcc -O3 -fopenmp, {threads} threads), {wall:.1f} s; "
                      f"{n1} steps on 1 thread, {wall1:.1f} s",
            "single_core_value": round(cells * n1 / wall1 / 1e6, 3),
            "nproc": os.cpu_count(), "cpu_model": cpu_model()}


def copy_bandwidth(mib: int = 1024, reps: int = 10) -> float:
    """GB/s of a plain 1 GiB device-to-device copy (read + write bytes over
    the event-timed duration), the practical ceiling the roofline is also
    quoted against."""
    import torch

    x = torch.empty(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    y.copy_(x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        y.copy_(x)
    e1.record()
    torch.cuda.synchronize()
    return 2.0 * mib * (1 << 20) * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9


def copy_calibration(mib: int = 1024, reps: int = 5):
    """A plain device copy of known size (read mib MiB + write mib MiB per
    launch), run before the bench so that a rocprofv3 --pmc pass over this
    command sees FETCH_SIZE / WRITE_SIZE for a known byte count."""
    import torch

    x = torch.empty(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    log(f"[bench] copy calibration: {reps} x {mib} MiB device copy")


def traffic_from_profile(workload: str):
    """HBM bytes per stage launch measured by rocprofv3 FETCH_SIZE /
    WRITE_SIZE (profiles/stage_kernel_traffic_<workload>.json), or None."""
    path = os.path.join(REPO, "profiles", f"stage_kernel_traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="sieve2048")
    ap.add_argument("--rows", type=int, default=0)
    ap.add_argument("--packing", type=int, default=None,
                    help="most segments of a narrow last y strip per wavefront: 0 (off), 2, 4 or 8")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=16)
    ap.add_argument("--cpu-steps-1core", type=int, default=2)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--copy-calib", action="store_true", help="run a 1 GiB device copy first (PMC scaling)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--no-closure", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the single-instance solver does not shard across GPUs).", file=sys.stderr)
        sys.exit(2)

    if args.copy_calib:
        copy_calibration()
    log(f"[bench] workload {args.workload}: warmup {args.warmup}, steps {args.steps}")
    m = gpu_measure(args.workload, args.steps, args.warmup, args.seed, args.rows or None, args.packing)
    if not m["finite"]:
        print(f"bench.py: the {args.workload} state is not finite after the timed steps", file=sys.stderr)
        sys.exit(3)
    log(f"[bench] {m['mcups']:.1f} Mcell/s, stage us {m['stage_avg_us']}, achieved {m['achieved_gbs']:.0f} GB/s")

    secondary = []
    if not args.no_secondary and args.workload == "sieve2048":
        # the same workload late in the run: the mixing layer has developed
        # and fewer rows repeat the row beside them
        log("[bench] secondary sieve2048 late (after 1500 untimed steps)")
        x = gpu_measure("sieve2048", args.steps, 1500, args.seed, args.rows or None, args.packing)
        secondary.append({"workload": x["desc"] + ", timed after 1500 untimed steps", "value": round(x["mcups"], 2),
                          "unit": "Mcell-updates/s", "dtype": "f32", "ms_per_step": round(x["ms_per_step"], 4),
                          "achieved_GBps": round(x["achieved_gbs"], 1),
                          "frac": round(x["achieved_gbs"] / HBM_PEAK_GBS, 4),
                          "stage_avg_us": [round(u, 2) for u in x["stage_avg_us"]], "t": x["t"]})
    if not args.no_secondary:
        for wl in ("rmi1024-fp64", "rmi1024-fp32"):
            if wl == args.workload:
                continue
            log(f"[bench] secondary {wl}")
            d = gpu_measure(wl, 200, 20, args.seed, None)
            x = gpu_measure(wl, 200, 20, args.seed, None, tune=True)
            secondary.append({"workload": x["desc"], "value": round(x["mcups"], 2), "unit": "Mcell-updates/s",
                              "dtype": "f32" if x["esize"] == 4 else "f64",
                              "rows_per_segment_tuned": x["rows"], "value_default_rows": round(d["mcups"], 2),
                              "ms_per_step": round(x["ms_per_step"], 4),
                              "achieved_GBps": round(x["achieved_gbs"], 1),
                              "frac": round(x["achieved_gbs"] / HBM_PEAK_GBS, 4),
                              "stage_avg_us": [round(u, 2) for u in x["stage_avg_us"]]})

    if not args.no_secondary and not args.no_batch:
        log("[bench] secondary batch64")
        secondary.append(batch_measure(64, args.batch_steps, 5, args.seed))

    if not args.no_secondary:
        log("[bench] secondary slabs")
        secondary.append(slab_measure())

    if not args.no_secondary and not args.no_closure:
        log("[bench] secondary closure")
        secondary.append(closure_measure())

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        log(f"[bench] cpu baseline: oracle, {threads} threads, {args.cpu_steps} steps")
        cpu = cpu_baseline(args.seed, args.cpu_steps, threads, args.cpu_steps_1core)

    copy_gbs = copy_bandwidth()
    copy_ws_gbs = copy_bandwidth(mib=80, reps=50)
    log(f"[bench] device copy {copy_gbs:.0f} GB/s (1 GiB), {copy_ws_gbs:.0f} GB/s (80 MiB, one state's size)")
    traffic = traffic_from_profile(args.workload)
    avg_launch_bytes = sum(m["bytes_per_launch"]) / 3
    out = {
        "metric": METRIC,
        "value": round(m["mcups"], 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(m["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if m["esize"] == 4 else "f64",
        "data": "synthetic (seeded initial condition generated on the host)",
        "config": {"workload": m["desc"], "nx": m["cfg"].nx, "ny": m["cfg"].ny, "rows_per_segment": m["rows"],
                   "cfl": m["cfg"].cfl, "powell": bool(m["cfg"].powell_source and m["cfg"].B_transverse > 0)},
        "roofline": {"bound": "hbm", "achieved": round(m["achieved_gbs"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(m["achieved_gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": "stage_kernel (3 launches per step, the step control carried by stages 1 and 3)",
                     "algorithmic_bytes_per_launch": [int(b) for b in m["bytes_per_launch"]],
                     "avg_launch_bytes": int(avg_launch_bytes),
                     "stage_avg_us": [round(u, 2) for u in m["stage_avg_us"]],
                     "same_steps_as_value": m["same_steps"],
                     "copy_GBps": round(copy_gbs, 1),
                     "frac_of_copy": round(m["achieved_gbs"] / copy_gbs, 4),
                     "copy_GBps_80MiB": round(copy_ws_gbs, 1)},
        "cpu_baseline": cpu,
        "profiled_pass_value": round(m["profiled_mcups"], 2),
        "secondary": secondary,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
