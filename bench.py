"""Benchmark: the per-timestep SPH step on the 262,144-particle bonded tissue (B = 32) on one
MI355X, reported as particle-steps/s with the roofline of the dominant kernel and the CPU
oracle timed beside it on the host cores.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload tissue_262k] [--no-cpu]

One step = Run2InnerPhysicalLoop (pressure, force, advance) + Run2GeneAction (diffusion,
genes) + Run2PhysicalSort (insert, scan, counting sort), with all inputs resident in HBM.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = ("particle-steps/sec + % HBM peak, 262k bonded tissue; learned 4×dt vs plain at 1e-3")
HBM_PEAK_GBS = 8000.0
PHASES = ["pressure", "force", "advance", "diffusion", "genes", "insert", "scan", "place",
          "sort", "lists", "fill"]
KERNEL_OF_PHASE = {"pressure": "k_pressure", "force": "k_force", "advance": "k_advance",
                   "diffusion": "k_diffusion", "genes": "k_genes", "insert": "k_insert",
                   "scan": "k_scan", "place": "k_place+k_cellsort", "sort": "k_sort_gather",
                   "lists": "k_dense_lists", "fill": "k_fill_u32"}


def algorithmic_bytes(phase: str, scene, stats: dict) -> float:
    """Compulsory HBM bytes per launch of a phase's kernel(s) in this build's layout, with
    every array read or written once (neighbour data counted once, as the particle's own
    read). n = particles, C = cells, B = bond slots, bo / bi = mean outgoing / incoming
    bonds per particle (the scene's, from its bond table)."""
    n, c, b = stats["n_active"], scene.c, scene.b
    bo, bi = stats["bonds_out"], stats["bonds_in"]
    if phase == "pressure":   # gcell 4, P 16 read, press/dens 8 written; CNT/OFF once
        return n * (4 + 16 + 8) + 8 * c
    if phase == "force":      # gcell, home, P, V: 40; BPARTNER all B slots; per bond BPAR 16,
        # SLOT 4, BINT 8 r + 8 w; INC_OFF 8, INC 4 per incoming bond; F 16 written
        return n * (40 + 4 * b + bo * (16 + 4 + 16) + 8 + 4 * bi + 16) + 8 * c
    if phase == "advance":    # gcell 4, F 16, P 16, V 16, VEL 16 read; P, V, VEL 48 written
        return n * (4 + 64 + 48)
    if phase == "diffusion":  # gcell 4, P 16, CONC 64 read; CONC 64 written
        return n * (4 + 16 + 64 + 64) + 8 * c
    if phase == "genes":      # GMASK 4, HOME 4, EPIGEN 64 r/w, CONC 64 r/w
        return n * (8 + 128 + 128)
    if phase == "insert":     # P 16 read; GCELL, GNDX written; CNT atomics
        return n * (16 + 8) + 4 * c
    if phase == "scan":
        return 8 * c * (1 + 16) + 16 * 4
    if phase == "place":      # GCELL, GNDX read, PERM scatter; cellsort PERM r/w, GNDX w
        return n * (8 + 4 + 8 + 4) + 8 * c
    if phase == "sort":       # PERM 4; P V VEL 48, CONC 64, HOME GCELL GNDX 12 read and written
        return n * (4 + 2 * (48 + 64 + 12) + 64 + 4 + 4)  # + EPIGEN 64 read, GMASK 4, SLOT 4
    if phase == "lists":
        return n * (4 + 16 * 4 * 0.3) + 8 * c * 17
    return 0.0


def scene_stats(scene) -> dict:
    import numpy as np

    from morphogenesis_amd import _ffi as F

    rec = scene.buffers.get(F.FELASTIDX)
    if rec is None or scene.b == 0:
        bo = bi = 0.0
    else:
        j = rec[:, :, 0]
        valid = j != F.NONE
        bo = float(valid.sum()) / scene.n
        bi = bo
    return {"bonds_out": bo, "bonds_in": bi}


def cpu_baseline(scene, budget_s: float = 15.0) -> dict:
    """The FP32 oracle (a plain C++ restatement of the reference step, OpenMP over all host
    cores) on a bounded sample of the same workload: the same scene, a few steps."""
    import multiprocessing

    from oracle.mg_oracle import Oracle

    threads = int(os.environ.get("MG_CPU_THREADS", "0")) or multiprocessing.cpu_count()
    o = Oracle(scene.n, scene.b, precision=32, threads=threads)
    o.load_scene(scene)
    o.Run2PhysicalSort()
    t = time.perf_counter()
    o.step(1, scene.dt, scene.flags)
    one = time.perf_counter() - t
    k = max(2, min(40, int(budget_s / max(one, 1e-3))))
    t = time.perf_counter()
    o.step(k, scene.dt, scene.flags)
    el = time.perf_counter() - t
    o.close()
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": scene.n * k / el, "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{k} steps of the same {scene.n}-particle scene after one sort, "
                      f"{el:.1f} s, FP32 oracle (CPU restatement), {threads} threads on {model}"}


def load_traffic(kernel: str, workload: str):
    p = os.path.join(REPO, "profiles", f"traffic_{kernel}_{workload}.json")
    if not os.path.exists(p):
        return None
    try:
        return json.load(open(p)).get("bytes_per_launch")
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="tissue_262k")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--timed-steps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the step would shard by spatial slabs; see DESIGN.md). Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from morphogenesis_amd import _ffi as F
    from morphogenesis_amd import scenes
    from morphogenesis_amd.fluid_system import FluidSystem

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    builders = {"tissue_262k": lambda: scenes.tissue(64, 32, seed=1),
                "fluid_block_32k": lambda: scenes.fluid_block(32, seed=1)}
    scene = builders[args.workload]()
    stats = scene_stats(scene)

    sim = FluidSystem(scene.n, scene.b)
    sim.load_scene(scene)
    sim.Run2PhysicalSort()
    sim.step(args.warmup, scene.dt, scene.flags)
    sim.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.step(args.steps, scene.dt, scene.flags)
    sim.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    nump, na = sim.counts()
    stats["n_active"] = na

    # per-kernel durations with HIP events on the library's own stream
    ms = (C.c_double * 11)()
    calls = (C.c_int * 11)()
    lib = sim.lib
    lib.mg_step_timed.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_uint,
                                  C.POINTER(C.c_double), C.POINTER(C.c_int)]
    r = lib.mg_step_timed(sim.h, args.timed_steps, scene.dt, scene.flags, ms, calls)
    if r != 0:
        raise F.MgError(sim.last_error())
    per_kernel = {}
    for i, ph in enumerate(PHASES):
        if calls[i]:
            per_kernel[ph] = {"avg_ms": ms[i] / calls[i] * (1 if ph not in ("place", "sort") else 1),
                              "calls": calls[i]}
    # launches per step of each phase: place/sort scopes bracket two kernels
    dom = max((p for p in per_kernel if p != "fill"), key=lambda p: per_kernel[p]["avg_ms"])
    dom_ms = per_kernel[dom]["avg_ms"] / (calls[PHASES.index(dom)] / args.timed_steps)
    alg = algorithmic_bytes(dom, scene, stats)
    achieved = alg / (dom_ms * 1e-3) / 1e9
    breakdown = {p: round(v["avg_ms"] * v["calls"] / args.timed_steps, 5) for p, v in per_kernel.items()}
    ms_per_step = elapsed / args.steps * 1e3
    value = scene.n * args.steps / elapsed
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": args.workload, "particles": scene.n, "active": na,
                   "bonds_per_particle": scene.b, "bonds_per_particle_mean": stats["bonds_out"] * 2,
                   "cells": scene.c, "dt": scene.dt, "genes": bool(scene.flags & F.STEP_GENES),
                   "step": "pressure+force+advance+diffusion+genes+insert+scan+sort"},
        "roofline": {"bound": "hbm", "kernel": KERNEL_OF_PHASE[dom], "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": load_traffic(KERNEL_OF_PHASE[dom], args.workload),
                     "algorithmic_bytes_per_launch": alg, "avg_launch_ms": dom_ms},
        "step_breakdown_ms": breakdown,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
