"""Build recipe for libaeromhd_hip.so (gfx950).

Three translation units, compiled by hipcc in-tree:

* ``csrc/amhd_f64.hip``  FP64 fused update, ``-ffp-contract=off`` (IEEE
  per-operation arithmetic, as the reference's ufuncs);
* ``csrc/amhd_f32.hip``  FP32 fused update, ``-ffp-contract=on`` (multiply-adds fused per
  source expression, so a body inlined into different kernels rounds the same);
* ``csrc/amhd_api.hip``  the C ABI of ``include/aeromhd.h`` and the FP64
  standalone operators, ``-ffp-contract=off``;
* ``csrc/amhd_cnn.hip``  BF16-MFMA 3x3 convolutions of the learned closure.

Every TU is built with packed-FP32 instructions disabled
(``-target-feature -packed-fp32-ops``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(REPO, "include")
LIB_NAME = "libaeromhd_hip.so"
ARCH = os.environ.get("AMHD_OFFLOAD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", f"-I{INCLUDE}", f"-I{CSRC}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
UNITS = [
    ("amhd_f64.hip", ["-ffp-contract=off"]),
    ("amhd_f32.hip", ["-ffp-contract=on"]),
    ("amhd_api.hip", ["-ffp-contract=off"]),
    ("amhd_cnn.hip", ["-ffp-contract=on"]),
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found: the gfx950 library cannot be built")
    return path


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build(out_dir: str | None = None, verbose: bool = False, extra_flags=()) -> str:
    """Compile the translation units and link ``libaeromhd_hip.so`` into
    out_dir (default: this package); extra_flags (e.g. -D tuning macros) go
    to every compile and get their own object directory."""
    out_dir = out_dir or HERE
    tag = "_".join(f.strip("-").replace("=", "") for f in extra_flags)
    obj_dir = os.path.join(REPO, "_build", "amhd" + ("_" + tag if tag else ""))
    os.makedirs(obj_dir, exist_ok=True)
    cc = hipcc()
    jobs = []
    objs = []
    for src, extra in UNITS:
        obj = os.path.join(obj_dir, src.replace(".hip", ".o"))
        objs.append(obj)
        jobs.append([cc, *COMMON, *extra, *extra_flags, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        for cmd, fut in zip(jobs, [ex.submit(_run, c) for c in jobs]):
            if verbose:
                print(" ".join(cmd), flush=True)
            fut.result()
    os.makedirs(out_dir, exist_ok=True)
    lib = os.path.join(out_dir, LIB_NAME)
    tmp = lib + ".tmp"
    _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    # python build_lib.py [-v] [--wave-times OUT_DIR]: the regular library, or
    # a diagnostic one that records every stage-kernel wavefront's start and
    # end time (exported as amhd_debug_wave_times: [stage][wave][start, end] in
    # ticks of the 100 MHz s_memrealtime clock)
    if "--wave-times" in sys.argv:
        out = sys.argv[sys.argv.index("--wave-times") + 1]
        print(build(out, verbose="-v" in sys.argv, extra_flags=("-DAMHD_WAVE_TIMES",)))
    else:
        print(build(verbose="-v" in sys.argv))
