"""Builds libalamo_elastic.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m alamo_amd.build   (or alamo_amd.build.build()).
The library is written next to this file so that it travels with the
source tree; nothing is installed.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libalamo_elastic.so")
ARCH = os.environ.get("ALAMO_AMD_ARCH", "gfx950")

SOURCES = ["elastic.hip", "unet.hip"]
# headers the two sources include (all of csrc/ is compiled on every build)
HEADERS = ["common.h", "elastic_kernels.h", "stencil.h", "stencil_lds.h", "bottom_kernels.h", "amr_kernels.h",
           "learned_kernels.h", "unet_kernels.h", "psi_kernels.h", "cluster.h", "inverse_kernels.h"]

# Packed FP32 instructions are disabled for all device code: hipcc -O3 forms
# them from ordinary float code and some of their forms have given wrong
# results on this hardware pool.
HIPFLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"),
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    """Compile every HIP source into libalamo_elastic.so (always recompiles)."""
    cmd = [hipcc(), *HIPFLAGS, "-o", LIB + ".tmp", *[os.path.join(CSRC, s) for s in SOURCES]]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libalamo_elastic.so")
    warn = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if verbose and warn:
        sys.stderr.write(warn + "\n")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
