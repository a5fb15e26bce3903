"""Recipe for oracle/_ref/libhlld_ref.so -- TEST INFRASTRUCTURE ONLY.

The reference's only native code is the HLLD kernel held as a string in
physics.py (physics.py:43-276, ``_hlld_kernel_code``).  Its text is plain
CUDA C that HIP compiles unchanged.  This recipe reads that file as text
(``ast.parse``; nothing is imported or executed), writes the kernel string
to oracle/_ref/hlld_ref_kernel.inc, and compiles oracle/ref_launcher.hip
(a host harness that includes it) with hipcc for gfx950 with
-ffp-contract=off, so the reference kernel runs on the GPU with IEEE
per-operation arithmetic.  Outputs go to oracle/_ref/ only (git-ignored).

It needs /root/reference; on a machine without it the prebuilt
oracle/_ref/libhlld_ref.so (if any) is used as is.
"""
from __future__ import annotations

import ast
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF_PHYSICS = os.environ.get("AMHD_REFERENCE_PHYSICS", "/root/reference/physics.py")
OUT = os.path.join(HERE, "_ref")
LIB = os.path.join(OUT, "libhlld_ref.so")


def extract_kernel(path: str = REF_PHYSICS) -> str:
    tree = ast.parse(open(path, encoding="utf-8").read(), filename=path)
    for node in ast.walk(tree):
        if isinstance(node, ast.Assign) and any(getattr(t, "id", None) == "_hlld_kernel_code" for t in node.targets):
            if isinstance(node.value, ast.Constant) and isinstance(node.value.value, str):
                return node.value.value
    raise RuntimeError(f"_hlld_kernel_code not found in {path}")


def build(arch: str = "gfx950") -> str | None:
    if not os.path.exists(REF_PHYSICS):
        return LIB if os.path.exists(LIB) else None
    os.makedirs(OUT, exist_ok=True)
    inc = os.path.join(OUT, "hlld_ref_kernel.inc")
    text = extract_kernel()
    if not os.path.exists(inc) or open(inc).read() != text:
        with open(inc, "w") as f:
            f.write(text)
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    tmp = LIB + ".tmp"
    subprocess.run([hipcc, f"--offload-arch={arch}", "-O3", "-fPIC", "-shared", "-ffp-contract=off",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    f"-I{OUT}", "-o", tmp, os.path.join(HERE, "ref_launcher.hip")], check=True,
                   capture_output=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build() or "reference not present; nothing built", file=sys.stderr)
