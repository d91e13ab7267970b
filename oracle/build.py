"""Build the CPU oracle (test infrastructure): oracle/lib/libmg_oracle.so.

g++ with -ffp-contract=off (no FMA contraction, matching the HIP build) and OpenMP. The
reference itself is not built here: its step lives in CUDA/OpenCL sources whose host
needs jsoncpp, VTK and Boost, none of which this image has.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")


def build_oracle(verbose: bool = False) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "libmg_oracle.so")
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared",
           "-Wall", "-Wextra", "-o", tmp, os.path.join(HERE, "mg_oracle.cpp")]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout)
        raise RuntimeError("oracle build failed")
    if verbose and p.stdout.strip():
        print(p.stdout)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build_oracle(verbose=True))
