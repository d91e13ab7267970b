"""Build libmg_sph.so (hipcc, gfx950) in-tree: morphogenesis_amd/lib/libmg_sph.so.

Flags:
  -ffp-contract=off                      no FMA contraction: the FP32 results are the IEEE
                                         results of the expression trees the oracle uses;
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE division and sqrt;
  -Xclang -target-feature -Xclang -packed-fp32-ops  no packed FP32 instructions.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
ARCH = os.environ.get("MG_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]

SOURCES = ["mg_capi.hip", "fluid_system.cpp"]


def hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.sep not in c or os.path.exists(c):
            return c
    return "hipcc"


def run(cmd, cwd=None):
    p = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass reports the packed-fp32 target feature as unknown; that line is noise
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build failed ({p.returncode}): {' '.join(cmd)}")
    return out


def build_library(verbose: bool = False) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "libmg_sph.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    cmd = [hipcc()] + HIP_FLAGS + ["-shared", "-o", tmp] + srcs
    log = run(cmd)
    if verbose and log.strip():
        print(log)
    os.replace(tmp, out)
    return out


def build_rounding_check() -> str:
    """The GPU test library tests/hip/rounding_check.hip (the short forms of the correctly
    rounded sqrt and 1/x against the compiler's expansions), with the library's flags."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "hip", "rounding_check.hip")
    out_dir = os.path.join(root, "tests", "hip", "lib")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "librounding_check.so")
    tmp = out + ".tmp"
    run([hipcc()] + HIP_FLAGS + ["-shared", "-o", tmp, src])
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build_library(verbose=True))
