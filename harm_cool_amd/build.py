"""Build libharmcool.so (gfx950) in-tree with hipcc, and its fast-mode twin.

FP64 parity with the CPU path needs IEEE semantics: -ffp-contract=off (no FMA
contraction), no fast-math.  Packed FP32 ops are disabled for all device code.

libharmcool_fast.so is the same sources with FMA contraction in the device code
only (-Xarch_device -ffp-contract=fast): every a*b+c of the kernels becomes one
FMA, which cuts the FP64 instruction count of the instruction-bound stencil
kernels; its results differ from the parity build within rounding.  The host
code (geometry tables, initial data) stays uncontracted, so both builds start
from the same bits.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libharmcool.so")
LIB_FAST = os.path.join(HERE, "libharmcool_fast.so")
SOURCES = ["harm_cool.hip", "harm_geometry.cpp", "harm_mlp.hip", "harm_eos.cpp", "harm_restart.cpp", "harm_fermi.hip", "harm_cache.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HARM_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-variable",
]


VARIANTS = {
    "parity": (LIB, []),
    "fast": (LIB_FAST, ["-Xarch_device", "-ffp-contract=fast"]),
}


def build(verbose: bool = False, variants=("parity", "fast")) -> str:
    """Compile the requested variants (in parallel); returns the parity
    library's path."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    procs = []
    for v in variants:
        out, extra = VARIANTS[v]
        tmp = out + ".tmp"
        cmd = [HIPCC, *FLAGS, *extra, "-I", os.path.join(HERE, "..", "include"), "-o", tmp, *srcs]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((v, out, tmp, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for v, out, tmp, p in procs:
        log, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(log)
            failed.append(v)
        else:
            os.replace(tmp, out)
    if failed:
        raise RuntimeError(f"hipcc failed building {', '.join(failed)}")
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
