"""FP32 vs FP64 drift of the step: the same scene stepped by the FP32 and the FP64 oracle.

TEST INFRASTRUCTURE (the GPU path equals the FP32 oracle bit for bit, so this is also the
drift of the GPU path against FP64 arithmetic).

  python -m oracle.drift --workload tissue_262k --steps 100 [--threads 8]

For each field, prints max |a32 - a64| / max |a64| and the same ratio of RMS values.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from morphogenesis_amd import _ffi as F  # noqa: E402
from morphogenesis_amd import scenes  # noqa: E402
from oracle.mg_oracle import Oracle  # noqa: E402

WORKLOADS = {
    "tissue_262k": lambda: scenes.tissue(64, 32, seed=1),
    "tissue_4k": lambda: scenes.tissue(16, 32, seed=1),
    "fluid_block_32k": lambda: scenes.fluid_block(32, seed=1),
}

FIELDS = [(F.FPOS, "pos", 3), (F.FVEL, "vel", 3), (F.FVEVAL, "veval", 3), (F.FPRESS, "press", 0),
          (F.FDENSITY, "density", 0), (F.FCONC, "conc", 0)]


def run(scene, precision: int, steps: int, threads: int) -> dict:
    o = Oracle(scene.n, scene.b, precision=precision, threads=threads)
    o.load_scene(scene)
    o.Run2PhysicalSort()
    o.step(steps, scene.dt, scene.flags)
    out = {}
    for buf, name, k in FIELDS:
        a = o.download(buf)
        out[name] = np.asarray(a[:, :k] if k else a, dtype=np.float64)
    nump, na = o.counts()
    o.close()
    out["_active"] = na
    return out


def drift(scene, steps: int, threads: int = 0) -> dict:
    r32 = run(scene, 32, steps, threads)
    r64 = run(scene, 64, steps, threads)
    na = min(r32["_active"], r64["_active"])
    res = {"steps": steps, "particles": scene.n, "compared": na}
    for _, name, _ in FIELDS:
        a, b = r32[name][:na], r64[name][:na]
        scale = np.abs(b).max()
        rms = np.sqrt(np.mean(b * b))
        d = np.abs(a - b)
        res[name] = {"max_rel": float(d.max() / scale) if scale > 0 else 0.0,
                     "rms_rel": float(np.sqrt(np.mean(d * d)) / rms) if rms > 0 else 0.0}
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="tissue_262k", choices=sorted(WORKLOADS))
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args(argv)
    t = time.time()
    res = drift(WORKLOADS[a.workload](), a.steps, a.threads)
    res["workload"] = a.workload
    res["seconds"] = round(time.time() - t, 1)
    print(json.dumps(res, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
