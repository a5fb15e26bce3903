"""Per-wavefront timing of the fused stage kernels (diagnostic).

Needs the diagnostic build of the library, which records the start and end
time of every stage-kernel wavefront (100 MHz s_memrealtime ticks) and
exports amhd_debug_wave_times:

    python aeromhd-cuda_amd/build_lib.py --wave-times aeromhd-cuda_amd/_wave_times
    AMHD_LIB=aeromhd-cuda_amd/_wave_times/libaeromhd_hip.so python aeromhd-cuda_amd/wave_times.py

It advances the 2048x1024 FP32 magnetic-sieve workload on the device loop
and prints, for the last launch of each RK stage, the spread of wavefront
durations and how it splits by dispatch order: the hardware deals
workgroups round-robin over the XCDs and then over their CUs, so with two
workgroups per CU the first CU-count workgroups are the first on their CU
and share every SIMD with a later one.
"""
from __future__ import annotations

import argparse
import ctypes as C
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

WT_WAVES = 8192   # the diagnostic array's wave capacity (amhd_kernels.hpp)
WPB = 4           # wavefronts per workgroup of the stage kernel


def wave_times(lib) -> np.ndarray:
    """(4, WT_WAVES, 2) int64 start/end ticks of the latest launch per stage."""
    if not hasattr(lib, "amhd_debug_wave_times"):
        raise RuntimeError("this libaeromhd_hip.so is not the diagnostic build (build_lib.py --wave-times)")
    buf = (C.c_ulonglong * (4 * WT_WAVES * 2))()
    if lib.amhd_debug_wave_times(buf) != 0:
        raise RuntimeError("amhd_debug_wave_times failed")
    return np.frombuffer(buf, dtype=np.uint64).reshape(4, WT_WAVES, 2).astype(np.int64)


def summarize(t: np.ndarray, n_waves: int, cus_per_xcd: int) -> dict:
    """Durations (us) of the n_waves wavefronts of one launch, overall and
    split into each CU's first and second workgroup."""
    t = t[:n_waves]
    dur = (t[:, 1] - t[:, 0]) / 100.0
    span = (t[:, 1].max() - t[:, 0].min()) / 100.0
    n_blocks = (n_waves + WPB - 1) // WPB
    per_xcd = (n_blocks & ~7) // 8
    logical = np.arange(n_waves) // WPB
    in_xcd = np.where(logical < 8 * per_xcd, logical % max(per_xcd, 1), per_xcd)
    first = in_xcd < cus_per_xcd
    pct = lambda x: np.percentile(x, [0, 10, 50, 90, 100]).round(1).tolist()
    return {"span_us": round(float(span), 1), "duration_us_min_p10_med_p90_max": pct(dur),
            "first_on_cu_median_us": round(float(np.median(dur[first])), 1),
            "second_on_cu_median_us": round(float(np.median(dur[~first])), 1) if (~first).any() else None,
            "slot_occupancy": round(float(dur.sum() / (span * 2 * 4 * cus_per_xcd * 8)), 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--detail", type=int, default=0,
                    help="also list the N slowest wavefronts and the last N (packed ones come last)")
    args = ap.parse_args()
    import amhd_lib as L
    torch = L.require_gpu()
    from amhd_ics import cons_from_prim_host, magnetic_sieve
    from solver import MHDSolver

    cfg, W = magnetic_sieve(2048, 1024, precision="fp32", seed=0, max_steps=10 ** 6)
    s = MHDSolver(cfg)
    s.U = cons_from_prim_host(W, cfg.gamma)
    s.advance(args.steps)
    torch.cuda.synchronize()
    t = wave_times(L.load())
    cus_per_xcd = torch.cuda.get_device_properties(0).multi_processor_count // 8
    for stage in (1, 2, 3):
        n = int((t[stage, :, 1] > 0).sum())
        print(f"stage {stage}: {n} wavefronts", summarize(t[stage], n, cus_per_xcd), flush=True)
        if args.detail:
            tt = t[stage, :n]
            dur = (tt[:, 1] - tt[:, 0]) / 100.0
            t0 = tt[:, 0].min()
            slow = np.argsort(dur)[::-1][:args.detail]
            fmt = lambda ids: [(int(w), round(float(dur[w]), 1), round(float(tt[w, 0] - t0) / 100.0, 1)) for w in ids]
            print("  slowest (wave, us, start us):", fmt(slow), flush=True)
            print("  last (wave, us, start us):", fmt(range(max(0, n - args.detail), n)), flush=True)


if __name__ == "__main__":
    main()
