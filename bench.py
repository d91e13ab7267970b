"""Benchmark: the per-timestep SPH step on the 262,144-particle bonded tissue (B = 32) on one
MI355X, reported as particle-steps/s with the roofline of the dominant kernel and the CPU
oracle timed beside it on the host cores.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload tissue_262k] [--no-cpu]
                  [--learned]

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
          "sort", "lists", "fill", "learn"]
# kernels of the fused step (mg_step) behind each timed phase
KERNEL_OF_PHASE = {"pressure": "k_pressure", "force": "k_force_advance",
                   "diffusion": "k_diffuse_genes_insert", "scan": "k_scan_lookback",
                   "place": "k_place", "sort": "k_sort_gather",
                   "lists": "k_dense_lists", "fill": "k_fill_u32"}


def bond_records_in_use(scene) -> bool:
    """The fused step hands bond records from the pressure to the force pass together with
    the hit masks (rows of up to 4 cells, cut-off 1) when the particles have bond slots."""
    return float(scene.params.r2) == 1.0 and int(scene.grid.srch) <= 4 and scene.b > 0


def algorithmic_bytes(phase: str, scene, stats: dict) -> float:
    """Compulsory HBM bytes per launch of the fused step's kernels in this build's layout:
    every array element read or written once, neighbour data counted once (as that
    particle's own read). n = active particles, C = cells, B = bond slots, bo / bi = mean
    outgoing / incoming bonds per particle."""
    n, c, b = stats["n_active"], scene.c, scene.b
    bo, bi = stats["bonds_out"], stats["bonds_in"]
    genes = bool(scene.flags & 1)
    srch = int(scene.grid.srch)
    # hit masks handed from the pressure to the force sweep: 8 bytes per row of cells; the
    # bond records go with them (a 16-byte record per owned bond)
    hm = 8 * srch * srch if (float(scene.params.r2) == 1.0 and srch <= 4) else 0
    rec = bond_records_in_use(scene)
    if phase == "pressure":
        # GCELL 4, P 16 read; press, 1/density 8 written; hit masks written; OFF read once,
        # next grid's CNT cleared; with bond records HOME 4, LIVE 8, V 16 and per owned bond
        # BPARTNER 4, SLOT 4, BPAR 16 read, the record 16 written
        return n * (4 + 16 + 8 + hm + ((28 + 40 * bo) if rec else 0)) + 8 * c
    if phase == "force":
        # force: GCELL 4, HOME 4, P 16, V 16 (+VEL 16 for the advance); hit masks read;
        # without records BPARTNER 4 per slot and per bond BPAR 16 + SLOT 4, with them LIVE 8
        # and per owned bond its record 16 (an incoming bond's record is its owner's, counted
        # there) plus SLOT 4 per incoming bond; BINT 8 read + 8 written per bond; INC_OFF 8,
        # INC 4 per incoming bond; advance writes P, V, VEL: 48
        if rec:
            return n * (56 + hm + 8 + bo * 32 + 8 + 8 * bi + 48)
        return n * (56 + hm + 4 * b + bo * 36 + 8 + 4 * bi + 48)
    if phase == "diffusion":
        # P 16, CONC 64 read, CONC 64 written, GCELL 4 read/written, GNDX 4, GMASK 4, HOME 4,
        # EPIGEN 64 read (+64 written with genes), GMN 4; the hit masks read (within the skin)
        return n * (16 + 128 + 8 + 4 + 4 + 4 + 64 + (64 if genes else 0) + 4 + hm) + 4 * c
    if phase == "scan":
        return 8 * c
    if phase == "place":  # k_place: GCELL, GNDX read, PERM written; OFF read per cell
        return n * (8 + 4) + 4 * c
    if phase == "sort":
        # PERM 4 (the cell runs the ranks count re-read it); P, V, VEL 48 and CONC 64 read and
        # written; HOME/GCELL 8 read, HOME/GCELL/GNDX 12 written; GMN 4 read, GMASK 4 and
        # SLOT 4 written; OFF/CNT 8 per cell
        return n * (4 + 2 * (48 + 64) + 8 + 12 + 4 + 4 + 4) + 8 * c
    return 0.0


def reference_layout_bytes(phase: str, scene, stats: dict):
    """Compulsory HBM bytes per launch in the reference's own buffer layout (src/fluid.h:
    138-206; bond records of 9 words per slot, FEPIGEN / FCONC of 16 words), per particle
    times the active particles, for the reference kernels that each fused kernel of this
    build performs. Returns (bytes, {term: bytes per particle}). Terms per particle:
      insert     16 position + 64 FEPIGEN read + 8 FGCELL/FGNDX written            = 88
      pressure   16 position + 4 cell read, 8 pressure and density written        = 28
      force      16 position + 16 veval + 4 press + 4 density + 4 cell
                 + 36 per bond slot (the record) + 8 per used bond (integrators written)
                 + 32 force read-modify-write
      advance    64 position, veval, velocity, force + 4 cell read, 48 written      = 116
      diffusion  64 FCONC read + 64 written                                          = 128
      genes      FEPIGEN and FCONC read and written with the ftemp accumulators      = 256
    Used bonds are the particle's own live bond slots (bonds_out per particle). When the
    pressure pass evaluates the bond terms for the force (the bond records), the force's reads
    of the bond records (36 per slot) and of veval (16) count as the pressure kernel's."""
    b, bo = scene.b, stats["bonds_out"]
    genes = bool(scene.flags & 1)
    rec = bond_records_in_use(scene)
    terms = {}
    if phase == "pressure":
        terms = {"pressure": 28.0}
        if rec:
            terms["bond terms"] = 16 + 36.0 * b
    elif phase == "force":
        bonds = 0.0 if rec else 16 + 36.0 * b
        terms = {"force": 16 + 4 + 4 + 4 + bonds + 8.0 * bo + 32, "advance": 116.0}
    elif phase == "diffusion":
        terms = {"diffusion": 128.0, "insert": 88.0}
        if genes:
            terms["genes"] = 256.0
    else:
        return 0.0, {}
    return stats["n_active"] * sum(terms.values()), terms


def scene_stats(scene, neighbours: bool = True) -> dict:
    import numpy as np
    from scipy.spatial import cKDTree

    from morphogenesis_amd import _ffi as F

    rec = scene.buffers.get(F.FELASTIDX)
    if rec is None or scene.b == 0:
        bo = bi = 0.0
    else:
        valid = rec[:, :, 0] != F.NONE
        bo = float(valid.sum()) / scene.n
        bi = bo
    if not neighbours:
        return {"bonds_out": bo, "bonds_in": bi, "neighbours": None}
    pos = scene.buffers[F.FPOS][:, :3].astype(np.float64)
    rng = np.random.default_rng(0)
    idx = rng.choice(pos.shape[0], size=min(4096, pos.shape[0]), replace=False)
    tree = cKDTree(pos)
    # neighbours within the force radius (dsq < 1), the particle itself excluded
    nbrs = float(np.mean([len(x) for x in tree.query_ball_point(pos[idx], 1.0)])) - 1.0
    return {"bonds_out": bo, "bonds_in": bi, "neighbours": nbrs}


# the step count each workload's configuration names (the 262k tissue: 1000 steps of 1e-4)
CONFIG_STEPS = {"tissue_262k": 1000, "fluid_block_32k": 100, "mixed_4m": 200}


def config_run(scene, steps: int, flags: int) -> dict:
    """The configuration's whole run from its initial state: `steps` steps on a fresh system,
    timed in chunks. The state changes along the run (the tissue settles and compresses under
    gravity, so cells fill up and the neighbour sweeps slow down); `value` above is the K-step
    window after W warmup steps that the benchmark contract times."""
    from morphogenesis_amd.fluid_system import FluidSystem

    sim = FluidSystem(scene.n, scene.b)
    sim.load_scene(scene)
    sim.Run2PhysicalSort()
    chunk = max(steps // 10, 1)
    done, total_s, per_chunk = 0, 0.0, []
    while done < steps:
        k = min(chunk, steps - done)
        sim.sync()
        t = time.perf_counter()
        sim.step(k, scene.dt, flags)
        sim.sync()
        el = time.perf_counter() - t
        total_s += el
        done += k
        per_chunk.append(round(el / k * 1e3, 4))
    sim.close()
    return {"steps": steps, "value": scene.n * steps / total_s, "unit": "particle-steps/s",
            "ms_per_step": total_s / steps * 1e3, "ms_per_step_by_chunk": per_chunk,
            "chunk_steps": chunk, "note": "from the initial state; includes graph capture and re-homing"}


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


def run_batch(args, count: int = 256):
    """The batched workload: `count` bonded tissue blocks of log-uniform size in
    [2048, 65536] particles, each with its own domain, grid and genome, stepped together."""
    from morphogenesis_amd import _ffi as F
    from morphogenesis_amd import scenes
    from morphogenesis_amd.batch import Batch
    from morphogenesis_amd.fluid_system import FluidSystem

    def progress(msg):
        print(f"[bench batch] {msg}", file=sys.stderr, flush=True)

    scs = scenes.tissue_batch(count, 32, seed=3)
    total = sum(sc.n for sc in scs)
    progress(f"{count} scenes built, {total} particles")
    systems = []
    for sc in scs:
        sim = FluidSystem(sc.n, sc.b)
        sim.load_scene(sc)
        sim.Run2PhysicalSort()
        systems.append(sim)
    progress("uploaded and sorted")
    dt, flags = scs[0].dt, scs[0].flags
    bat = Batch(systems)
    chosen = bat.autotune(2, dt, flags)
    if args.batch_form != "auto":
        bat.set_persistent(args.batch_form == "persistent")
    persistent, workgroups = bat.launch_config()
    progress(f"autotuned: block sizes {chosen}, persistent launch {persistent} ({workgroups} workgroups)")
    bat.step(args.warmup, dt, flags | F.STEP_GRAPH)
    bat.sync()
    t0 = time.perf_counter()
    bat.step(args.steps, dt, flags | F.STEP_GRAPH)
    bat.sync()
    elapsed = time.perf_counter() - t0
    # the other launch form over the same number of steps, for comparison (not the value)
    forms = {("persistent" if persistent else "dispatch"): elapsed / args.steps * 1e3}
    if persistent or bat.launch_config()[0] == 0:
        bat.set_persistent(not persistent)
        other = bat.launch_config()[0]
        if other != persistent:
            bat.step(2, dt, flags | F.STEP_GRAPH)
            bat.sync()
            t1 = time.perf_counter()
            bat.step(args.steps, dt, flags | F.STEP_GRAPH)
            bat.sync()
            forms["persistent" if other else "dispatch"] = (time.perf_counter() - t1) / args.steps * 1e3
        bat.set_persistent(bool(persistent))
    ms = (C.c_double * len(PHASES))()
    calls = (C.c_int * len(PHASES))()
    bat._call("batch_step_timed", bat.h, args.timed_steps, dt, flags, ms, calls)
    per_step = {ph: ms[i] / args.timed_steps for i, ph in enumerate(PHASES) if calls[i]}
    kern = {"pressure": "kb_pressure", "force": "kb_force_advance",
            "diffusion": "kb_diffuse_genes_insert", "scan": "kb_scan", "place": "kb_place",
            "sort": "kb_sort_gather"}
    # algorithmic bytes of a launch group = sum over the instances
    stats = []
    for sc, sim in zip(scs, systems):
        st = scene_stats(sc, neighbours=False)
        st["n_active"] = sim.counts()[1]
        stats.append(st)
    kernels = {}
    for p, v in per_step.items():
        bts = sum(algorithmic_bytes(p, sc, st) for sc, st in zip(scs, stats))
        if p in kern and bts > 0:
            kernels[kern[p]] = {"ms": round(v, 5), "alg_bytes": int(bts), "GB/s": round(bts / (v * 1e-3) / 1e9, 1)}
    dom = max((p for p in per_step if p in kern and p != "place"), key=lambda p: per_step[p])
    ref = sum(reference_layout_bytes(dom, sc, st)[0] for sc, st in zip(scs, stats))
    terms = reference_layout_bytes(dom, scs[0], stats[0])[1]
    lay = sum(algorithmic_bytes(dom, sc, st) for sc, st in zip(scs, stats))
    achieved = ref / (per_step[dom] * 1e-3) / 1e9
    lay_ach = lay / (per_step[dom] * 1e-3) / 1e9
    classes = bat.classes()
    out = {
        "metric": METRIC, "value": total * args.steps / elapsed, "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": args.workload, "instances": count, "particles": total,
                   "bonds_per_particle": 32, "dt": dt, "genes": bool(flags & F.STEP_GENES),
                   "hipgraph": True,
                   "launch": {"persistent": bool(persistent), "workgroups": workgroups,
                              "ms_per_step_by_form": {k: round(v, 4) for k, v in forms.items()}},
                   "size_classes": bat.class_config(),
                   "autotuned_threads": chosen,
                   "step": ("pressure+force+advance+diffusion+genes+insert+scan+sort, "
                            + ("one persistent launch per step for all instances" if persistent
                               else "one launch per phase and size class"))},
        "roofline": {"bound": "hbm", "kernel": kern[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": load_traffic(kern[dom], args.workload),
                     "algorithmic_bytes_per_launch": ref,
                     "algorithmic_basis": "reference layout, per particle: " +
                                          ", ".join(f"{k} {v:.1f} B" for k, v in terms.items()),
                     "avg_launch_ms": per_step[dom],
                     "timing": "per-phase times of the dispatch form (HIP events around each launch group)",
                     "layout": {"bytes_per_launch": lay, "achieved": lay_ach, "frac": lay_ach / HBM_PEAK_GBS,
                                "basis": "this build's HBM layout, every element once"}},
        "step_breakdown_ms": {p: round(v, 5) for p, v in per_step.items()},
        "kernels": kernels,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_batch(scs)
    bat.close()
    print(json.dumps(out))


def cpu_baseline_batch(scs, budget_s: float = 15.0) -> dict:
    """The FP32 oracle stepping the instances one after another (as a CPU loop over the
    batch would) until the time budget is spent; rate = particle-steps / time."""
    import multiprocessing

    from oracle.mg_oracle import Oracle

    threads = int(os.environ.get("MG_CPU_THREADS", "0")) or multiprocessing.cpu_count()
    done, spent, used = 0, 0.0, 0
    for sc in scs:
        o = Oracle(sc.n, sc.b, precision=32, threads=threads)
        o.load_scene(sc)
        o.Run2PhysicalSort()
        t = time.perf_counter()
        o.step(2, sc.dt, sc.flags)
        spent += time.perf_counter() - t
        done += 2 * sc.n
        used += 1
        o.close()
        if spent > budget_s:
            break
    return {"value": done / spent, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"2 steps of each of the first {used} of {len(scs)} instances, {spent:.1f} s, "
                      f"FP32 oracle (CPU restatement), {threads} threads"}


def learned_measurement(coarse_steps: int = 1000, progress=None) -> dict:
    """The learned 4 x dt step on the 262,144-particle tissue. Training data: 200 fine
    trajectory segments (4 steps of dt each) along the fine trajectories of four other 262k
    tissues (seeds 21-24, 50 start states each, 80 fine steps apart, 65,536 random rows per
    segment); Adam on the GPU. Then from the sorted start state of the seed-1 tissue:
    4 K fine steps of dt, K learned steps of 4 dt and K plain steps of 4 dt (the ablation),
    positions compared by particle ID."""
    from morphogenesis_amd import scenes
    from morphogenesis_amd.learned import Net, make_training_data, rollout_errors

    say = progress or (lambda msg: None)
    t0 = time.perf_counter()
    train = [scenes.tissue(64, 32, seed=s) for s in (21, 22, 23, 24)]
    data = make_training_data(train, samples=50, ratio=4, gap=80, rows_per_sample=65536)
    del train
    gen_s = time.perf_counter() - t0
    say(f"training data: {data.segments} segments, {data.X.shape[0]} rows in {gen_s:.1f} s")
    net = Net(seed=5)
    t1 = time.perf_counter()
    loss = net.fit(data, steps=3000, batch=16384, lr=1e-3)
    train_s = time.perf_counter() - t1
    say(f"trained in {train_s:.1f} s, loss {loss[:5].mean():.3f} -> {loss[-100:].mean():.3f}")
    rows = int(data.X.shape[0])
    del data
    sc = scenes.tissue(64, 32, seed=1)
    r = rollout_errors(sc, net, coarse_steps, ratio=4)
    net.close()
    return {"speedup_vs_fine": r["speedup_vs_fine"],
            "speedup_plain_coarse_vs_fine": r["fine_s"] / r["coarse_s"],
            "rel_error_vs_fine_disp": r["learned_rel_disp"],
            "rel_error_vs_fine_pos": r["learned_rel_pos"],
            "ablation_plain_coarse_rel_error_disp": r["coarse_rel_disp"],
            "ablation_plain_coarse_rel_error_pos": r["coarse_rel_pos"],
            "error_ratio_learned_over_plain": r["learned_rel_disp"] / max(r["coarse_rel_disp"], 1e-30),
            "fallback_particle_steps": r["fallbacks"], "coarse_steps": coarse_steps,
            "fine_s": r["fine_s"], "learned_s": r["learned_s"], "coarse_s": r["coarse_s"],
            "train_segments": 200, "train_rows": rows, "train_data_s": gen_s, "train_s": train_s,
            "fallback_rule": "normalised input beyond 6 sigma of the training data, or a drift correction moving the particle more than 1e-3 in the step",
            "train_loss_first": float(loss[:5].mean()), "train_loss_last": float(loss[-100:].mean()),
            "error_norm": "|x - x_fine| / |x_fine - x_0| over all particles (2-norms), by particle ID",
            "particles": sc.n, "dt_fine": sc.dt, "dt_coarse": 4 * sc.dt}


def run_learned(args):
    """The learned workload as its own line: `value` counts a learned step as the 4 fine
    steps it replaces (fine-equivalent particle-steps/s)."""
    def progress(msg):
        print(f"[bench learned] {msg}", file=sys.stderr, flush=True)

    lm = learned_measurement(args.steps, progress)
    n = lm["particles"]
    out = {
        "metric": METRIC, "value": n * 4 * args.steps / lm["learned_s"], "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 0, "ms_per_step": lm["learned_s"] / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16+fp32",
        "data": "synthetic",
        "config": {"workload": args.workload, "particles": n, "bonds_per_particle": 32,
                   "dt_fine": lm["dt_fine"], "dt_coarse": lm["dt_coarse"], "coarse_steps": args.steps,
                   "hipgraph": True,
                   "mlp": "53(64)-128-128-6(16) SiLU, bf16 MFMA, fp32 accumulate; outputs: drift and veval corrections",
                   "value_counts": "each learned step as 4 fine-equivalent steps"},
        "learned": lm,
    }
    print(json.dumps(out))


def roofline(dom: str, scene, stats: dict, dom_ms: float, workload: str) -> dict:
    """The dominant kernel against HBM bandwidth. `achieved` counts the compulsory bytes of
    the reference's layout for the work the kernel does (reference_layout_bytes, terms
    listed); `layout` counts the compulsory bytes of this build's own layout
    (algorithmic_bytes), which move far fewer bytes per particle; `traffic` is the measured
    HBM traffic per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)."""
    ref_bytes, terms = reference_layout_bytes(dom, scene, stats)
    lay = algorithmic_bytes(dom, scene, stats)
    ach = ref_bytes / (dom_ms * 1e-3) / 1e9
    lay_ach = lay / (dom_ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": KERNEL_OF_PHASE[dom], "achieved": ach, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": load_traffic(KERNEL_OF_PHASE[dom], workload),
            "algorithmic_bytes_per_launch": ref_bytes,
            "algorithmic_basis": "reference layout, per particle: " +
                                 ", ".join(f"{k} {v:.1f} B" for k, v in terms.items()),
            "avg_launch_ms": dom_ms,
            "layout": {"bytes_per_launch": lay, "achieved": lay_ach, "frac": lay_ach / HBM_PEAK_GBS,
                       "basis": "this build's HBM layout, every element once"}}


def load_traffic(kernel: str, workload: str):
    p = os.path.join(REPO, "profiles", f"traffic_{kernel}_{workload}.json")
    if not os.path.exists(p):
        return None
    try:
        return json.load(open(p)).get("bytes_per_launch")
    except Exception:
        return None


def run_bonds(args):
    """Bond formation (mg_initialize_bonds) on the 262,144-particle tissue with gene 1 ("solid")
    on 30% of the particles: each call searches every bond slot of every listed particle.
    `value` = slot searches per second over whole calls (kernel + device rebuild of the
    incoming-bond lists); the CPU baseline is the oracle's initialize_bonds on the same state."""
    import numpy as np

    from morphogenesis_amd import _ffi as F
    from morphogenesis_amd import scenes
    from morphogenesis_amd.fluid_system import FluidSystem

    def scene(n_side):
        sc = scenes.tissue(n_side, 32, seed=1)
        sc.genome = scenes.example_genome()
        return sc

    sc = scene(64)
    sim = FluidSystem(sc.n, sc.b)
    sim.load_scene(sc)
    sim.Run2PhysicalSort()
    listed = int(sim.download(F.FDENSE_LIST_LENGTHS)[1])
    sim.InitializeBondsCL()
    sim.sync()
    reps = max(args.steps // 40, 3)
    t = time.perf_counter()
    for _ in range(reps):
        sim.InitializeBondsCL()
    sim.sync()
    dt = (time.perf_counter() - t) / reps
    out = {"metric": "bond-slot searches/s (initialize_bonds), 262k bonded tissue", "value": listed * sc.b / dt,
           "unit": "slot-searches/s", "n_gpus": 1, "steps": reps, "warmup": 1, "ms_per_step": dt * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32 (fp64 angle sum)",
           "data": "synthetic",
           "config": {"workload": args.workload, "particles": sc.n, "listed": listed, "bonds_per_particle": sc.b}}
    if not args.no_cpu:
        from oracle.mg_oracle import Oracle

        o = Oracle(sc.n, sc.b)
        o.load_scene(sc)
        o.Run2PhysicalSort()
        ls = int(o.download(F.FDENSE_LIST_LENGTHS)[1])
        cpu_reps = 2
        t = time.perf_counter()
        for _ in range(cpu_reps):
            o.InitializeBondsCL()
        ct = (time.perf_counter() - t) / cpu_reps
        out["cpu_baseline"] = {"value": ls * sc.b / ct, "unit": "slot-searches/s", "cores": 1,
                               "kind": "port", "sample": f"{cpu_reps} initialize_bonds calls over the same {ls} "
                               f"listed particles of the {sc.n}-particle tissue, {ct:.2f} s per call, FP32 oracle, "
                               "1 thread"}
    print(json.dumps(out), flush=True)


def remodel_scene(n_side):
    """Bonded tissue (B = 32) for the remodelling workload: gene 2 ("living") on 70% of the
    particles, the tissue genes 6 / 7 / 9 / 10 on 10-15% each, 5% of the bonds broken (rest
    length 0), strain integrators from the first steps, and gentle tanh remodelling rows."""
    import numpy as np

    from morphogenesis_amd import _ffi as F
    from morphogenesis_amd import scenes

    sc = scenes.tissue(n_side, 32, seed=1)
    g = scenes.example_genome()
    rows = [(1.0, 0.05, 10.0, 0.5, 1.0, 0.03, 40.0, 0.2), (1.01, 0.04, 6.0, 1.0, 0.99, 0.02, 20.0, 0.1),
            (0.995, 0.02, 3.0, 0.3, 1.02, 0.05, 8.0, 0.4)]
    for r, row in enumerate(rows):
        for k, v in enumerate(row):
            g.tanh_param[r][k] = v
    sc.genome = g
    rng = np.random.default_rng(11)
    epi = sc.buffers[F.FEPIGEN]
    for gi, frac in ((2, 0.7), (6, 0.1), (7, 0.1), (9, 0.15), (10, 0.1)):
        epi[gi] = np.where(rng.random(sc.n) < frac, 2147483647, 0).astype(np.uint32)
    rec = sc.buffers[F.FELASTIDX]
    broken = (rec[:, :, 0] != F.NONE) & (rng.random(rec.shape[:2]) < 0.05)
    rec[:, :, 2][broken] = 0
    return sc


def run_remodel(args):
    """Remodelling (Run2Remodelling = mg_compute_bond_changes, mg_prefix_sum_changes,
    mg_counting_sort_changes, mg_compute_particle_changes) on the 262,144-particle tissue of
    remodel_scene after 4 steps. `value` = bond slots of the living particles processed per
    second over whole calls (kernels + the device rebuild of the incoming-bond lists after
    heal); the CPU baseline is 25 of the oracle's calls on the same 262,144-particle state."""
    from morphogenesis_amd import _ffi as F
    from morphogenesis_amd.fluid_system import FluidSystem

    sc = remodel_scene(64)
    sim = FluidSystem(sc.n, sc.b)
    sim.load_scene(sc)
    sim.Run2PhysicalSort()
    sim.step(4, sc.dt, sc.flags | F.STEP_GRAPH)
    living = int(sim.download(F.FDENSE_LIST_LENGTHS)[2])
    sim.Run2Remodelling(4, seed=0)
    healed0 = int(sim.download(F.FDENSE_LIST_LENGTHS_CHANGES)[0])
    sim.sync()
    reps = max(args.steps // 40, 3)
    parts = {"bond_changes": 0.0, "prefix_sum": 0.0, "counting_sort": 0.0, "heal": 0.0}
    t0 = time.perf_counter()
    for r in range(reps):
        for name, fn in (("bond_changes", lambda: sim.ComputeBondChangesCL(4)),
                         ("prefix_sum", lambda: sim.PrefixSumChangesCL(1)),
                         ("counting_sort", sim.CountingSortChangesCL),
                         ("heal", lambda: sim.ComputeParticleChangesCL(r + 1))):
            t = time.perf_counter()
            fn()
            sim.sync()
            parts[name] += time.perf_counter() - t
    dt = (time.perf_counter() - t0) / reps
    out = {"metric": "bond slots remodelled/s (Run2Remodelling), 262k bonded tissue",
           "value": living * sc.b / dt, "unit": "slots/s", "n_gpus": 1, "steps": reps, "warmup": 1,
           "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "fp32 (fp64 angle sum in heal)", "data": "synthetic",
           "config": {"workload": args.workload, "particles": sc.n, "living": living, "bonds_per_particle": sc.b,
                      "heal_list_first_call": healed0,
                      "ms_per_call": {k: round(v / reps * 1e3, 3) for k, v in parts.items()}}}
    if not args.no_cpu:
        from oracle.mg_oracle import Oracle

        o = Oracle(sc.n, sc.b)
        o.load_scene(sc)
        o.Run2PhysicalSort()
        o.step(4, sc.dt, sc.flags)
        ls = int(o.download(F.FDENSE_LIST_LENGTHS)[2])
        cpu_reps = 25
        t = time.perf_counter()
        for r in range(cpu_reps):
            o.Run2Remodelling(4, r)
        ct = (time.perf_counter() - t) / cpu_reps
        out["cpu_baseline"] = {"value": ls * sc.b / ct, "unit": "slots/s", "cores": 1, "kind": "port",
                               "sample": f"{cpu_reps} Run2Remodelling calls over the same {ls} living particles "
                               f"of the {sc.n}-particle tissue, {ct:.3f} s per call, FP32 oracle, 1 thread"}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="tissue_262k")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--timed-steps", type=int, default=20)
    ap.add_argument("--no-graph", action="store_true", help="launch the step kernels eagerly")
    ap.add_argument("--batch-form", choices=("auto", "persistent", "dispatch"), default="auto",
                    help="tissue_batch_256: the launch form (auto: the autotuner's choice)")
    ap.add_argument("--no-config-run", action="store_true",
                    help="leave out the timing of the configuration's whole run")
    ap.add_argument("--learned", action="store_true",
                    help="add the learned 4 x dt measurement to the tissue_262k line (training "
                         "and three 1,000-coarse-step rollouts, about 1.5 min; the same as "
                         "--workload tissue_262k_learned)")
    ap.add_argument("--no-learned", action="store_true",
                    help="leave out the learned measurement (the default)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run a plain device copy of this size (traffic counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path (the step "
              "would shard the domain into spatial slabs, one per GPU, with a halo of one cell "
              "layer exchanged each step). Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    if args.calib_copy_mib > 0:
        # a runtime device copy of known size, for the traffic counters' scale; torch is
        # loaded before the library, so each keeps its own HIP runtime
        import torch

        n = args.calib_copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # the library is the only GPU user: creating a system fails loudly without a HIP device
    from morphogenesis_amd import _ffi as F
    from morphogenesis_amd import scenes
    from morphogenesis_amd.fluid_system import FluidSystem

    if args.workload == "tissue_batch_256":
        run_batch(args)
        return
    if args.workload == "tissue_262k_learned":
        run_learned(args)
        return
    if args.workload == "tissue_262k_bonds":
        run_bonds(args)
        return
    if args.workload == "tissue_262k_remodel":
        run_remodel(args)
        return
    builders = {"tissue_262k": lambda: scenes.tissue(64, 32, seed=1),
                "fluid_block_32k": lambda: scenes.fluid_block(32, seed=1),
                "mixed_4m": lambda: scenes.mixed(4194304, 32, seed=2)}
    scene = builders[args.workload]()
    stats = scene_stats(scene)


    sim = FluidSystem(scene.n, scene.b)
    sim.load_scene(scene)
    sim.Run2PhysicalSort()
    flags = scene.flags | (0 if args.no_graph else F.STEP_GRAPH)
    sim.step(args.warmup, scene.dt, flags)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(args.steps, scene.dt, flags)
    sim.sync()
    elapsed = time.perf_counter() - t0
    nump, na = sim.counts()
    stats["n_active"] = na

    # per-kernel durations with HIP events on the library's own stream
    ms = (C.c_double * len(PHASES))()
    calls = (C.c_int * len(PHASES))()
    lib = sim.lib
    lib.mg_step_timed.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_uint,
                                  C.POINTER(C.c_double), C.POINTER(C.c_int)]
    r = lib.mg_step_timed(sim.h, args.timed_steps, scene.dt, scene.flags, ms, calls)
    if r != 0:
        raise F.MgError(sim.last_error())
    per_step = {ph: ms[i] / args.timed_steps for i, ph in enumerate(PHASES) if calls[i]}
    dom = max((p for p in per_step if p in KERNEL_OF_PHASE and p not in ("fill", "place")),
              key=lambda p: per_step[p])
    dom_ms = per_step[dom]  # one launch per step
    breakdown = {p: round(v, 5) for p, v in per_step.items()}
    kernels = {}
    for p, v in per_step.items():
        if p in KERNEL_OF_PHASE and algorithmic_bytes(p, scene, stats) > 0:
            bts = algorithmic_bytes(p, scene, stats)
            rb, _ = reference_layout_bytes(p, scene, stats)
            kernels[KERNEL_OF_PHASE[p]] = {"ms": round(v, 5), "alg_bytes": int(bts),
                                           "GB/s": round(bts / (v * 1e-3) / 1e9, 1)}
            if rb:
                kernels[KERNEL_OF_PHASE[p]]["ref_layout_GB/s"] = round(rb / (v * 1e-3) / 1e9, 1)
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
                   "hipgraph": not args.no_graph,
                   "step": "pressure+force+advance+diffusion+genes+insert+scan+sort"},
        "roofline": roofline(dom, scene, stats, dom_ms, args.workload),
        "step_breakdown_ms": breakdown,
        "kernels": kernels,
        "mean_neighbours": round(stats["neighbours"], 2),
    }
    if args.workload in CONFIG_STEPS and not args.no_config_run:
        out["config_run"] = config_run(scene, CONFIG_STEPS[args.workload], flags)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene)
    if args.workload == "tissue_262k" and args.learned and not args.no_learned:
        sim.close()
        out["learned"] = learned_measurement(1000, lambda msg: print(f"[bench learned] {msg}",
                                                                      file=sys.stderr, flush=True))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
