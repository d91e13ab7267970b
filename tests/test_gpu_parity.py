"""GPU parity: the HIP library (through its C-ABI) against the FP32 CPU oracle on the same
seeded inputs. Integer/index results (cells, counts, offsets, order, bond indices, gene
states) must be bit-exact; floating-point state is compared bit-exactly too, which is
stronger than the 1e-5 relative bar (both sides evaluate the same expression trees in the
same order with IEEE division/sqrt and no FMA contraction)."""
import ctypes as C
import sys

import numpy as np
import pytest

from morphogenesis_amd import _ffi as F
from morphogenesis_amd import scenes
from tests.helpers import assert_bits_equal, mini_scene, per_cell_sets

pytestmark = pytest.mark.gpu

COMPARE = [F.FPOS, F.FVEL, F.FVEVAL, F.FFORCE, F.FPRESS, F.FDENSITY, F.FAGE, F.FCOLOR, F.FGCELL,
           F.FGNDX, F.FELASTIDX, F.FNERVEIDX, F.FCONC, F.FEPIGEN, F.FBIN, F.FPARTICLEIDX,
           F.FPARTICLE_ID, F.FMASS_RADIUS, F.FBIN_COUNT, F.FBIN_OFFSET,
           F.FBIN_COUNT_ACTIVE_GENES, F.FBIN_OFFSET_ACTIVE_GENES, F.FDENSE_LIST_LENGTHS,
           F.FDENSE_LISTS]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pair(scene):
    from morphogenesis_amd.fluid_system import FluidSystem
    from oracle.mg_oracle import Oracle

    g = FluidSystem(scene.n, scene.b)
    o = Oracle(scene.n, scene.b)
    g.load_scene(scene)
    o.load_scene(scene)
    return g, o


def compare_all(g, o, what, bufs=COMPARE):
    assert g.counts() == o.counts(), (what, g.counts(), o.counts())
    for b in bufs:
        assert_bits_equal(g.download(b), o.download(b), f"{what}: {F.BUFFER_NAMES[b]}")


def tricky_tissue(seed=5, n_side=10, b=16):
    """Tissue with reserve (null) particles, particles outside the scan window and genes."""
    sc = scenes.tissue(n_side=n_side, b=b, seed=seed, reserve=29)
    sc.buffers[F.FPOS][:7, 0] = 0.4        # outside the window: dropped at the first sort
    sc.buffers[F.FVEL][7:40, :3] = np.random.default_rng(seed).normal(0, 0.3, (33, 3))
    return sc


def test_phase_by_phase(built):
    sc = tricky_tissue()
    say("case: phase_by_phase", sc.name, sc.n)
    g, o = pair(sc)
    compare_all(g, o, "upload")
    gi = g.InsertParticlesCL(readback=True)
    oi = o.InsertParticlesCL(readback=True)
    np.testing.assert_array_equal(gi[0], oi[0])          # FGCELL
    np.testing.assert_array_equal(gi[2], oi[2])          # FBIN_COUNT
    acc = gi[0] != F.GRID_UNDEF
    # FGNDX after insert is an atomicAdd slot (any order); per-cell it is a permutation
    for c in np.unique(gi[0][acc])[:300]:
        m = gi[0] == c
        assert sorted(gi[1][m].tolist()) == list(range(int(m.sum())))
    for name, fg, fo in (("prefix", g.PrefixSumCellsCL, o.PrefixSumCellsCL),
                         ("sort", g.CountingSortFullCL, o.CountingSortFullCL),
                         ("pressure", g.ComputePressureCL, o.ComputePressureCL),
                         ("force", g.ComputeForceCL, o.ComputeForceCL)):
        say("phase:", name)
        fg(); fo()
        bufs = COMPARE if name != "prefix" else [F.FBIN_COUNT, F.FBIN_OFFSET,
                                                 F.FBIN_COUNT_ACTIVE_GENES,
                                                 F.FBIN_OFFSET_ACTIVE_GENES,
                                                 F.FDENSE_LIST_LENGTHS]
        compare_all(g, o, name, bufs)
    say("phase: advance")
    g.AdvanceCL(0.0, sc.dt, 1.0); o.AdvanceCL(0.0, sc.dt, 1.0)
    compare_all(g, o, "advance")
    say("phase: diffusion")
    g.ComputeDiffusionCL(); o.ComputeDiffusionCL()
    compare_all(g, o, "diffusion")
    say("phase: genes")
    g.ComputeGenesCL(); o.ComputeGenesCL()
    compare_all(g, o, "genes")
    say("phase: second sort")
    g.Run2PhysicalSort(); o.Run2PhysicalSort()
    compare_all(g, o, "sort2")
    # force accumulates across inner steps until the next sort (src/fluid_system_cuda.cu:225)
    for k in range(2):
        g.ComputePressureCL(); o.ComputePressureCL()
        g.ComputeForceCL(); o.ComputeForceCL()
    compare_all(g, o, "force accumulated twice")
    g.ZeroVelCL(); o.ZeroVelCL()
    compare_all(g, o, "zero vel")


@pytest.mark.parametrize("seed,n_side,b,steps", [(1, 12, 32, 100), (2, 9, 6, 60)])
def test_tissue_steps_bit_exact(built, seed, n_side, b, steps):
    sc = scenes.tissue(n_side=n_side, b=b, seed=seed)
    say("case: tissue steps", sc.name, "seed", seed, "steps", steps)
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
    for chunk in range(steps // 20):
        g.step(20, sc.dt, sc.flags)
        o.step(20, sc.dt, sc.flags)
        compare_all(g, o, f"after {20 * (chunk + 1)} steps")


def test_fluid_block_100_steps(built):
    sc = scenes.fluid_block(32, seed=1)
    say("case: fluid block 32^3, 100 steps")
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
        sim.step(100, sc.dt, 0)
    compare_all(g, o, "fluid block 100 steps")


def test_escaping_and_breaking_bonds(built):
    # stretched bonds beyond the elastic limit break; fast particles leave the window
    sc = scenes.tissue(n_side=8, b=8, seed=9, limit=0.62)
    rng = np.random.default_rng(3)
    sc.buffers[F.FVEL][:, :3] = rng.normal(0, 2.5, (sc.n, 3)).astype(np.float32)
    sc.buffers[F.FVEVAL][:, :3] = sc.buffers[F.FVEL][:, :3]
    say("case: breaking bonds / escaping particles")
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
    for k in range(4):
        g.step(10, 1e-3, sc.flags)
        o.step(10, 1e-3, sc.flags)
        compare_all(g, o, f"escape chunk {k}")
    rest = g.download(F.FELASTIDX)[:, :, 2].view(np.float32)
    assert (rest == 0).any(), "no bond broke"


def test_gpu_deterministic(built):
    from morphogenesis_amd.fluid_system import FluidSystem

    sc = scenes.tissue(n_side=12, b=32, seed=7)
    say("case: determinism")
    outs = []
    for _ in range(2):
        g = FluidSystem(sc.n, sc.b)
        g.load_scene(sc)
        g.Run2PhysicalSort()
        g.step(30, sc.dt, sc.flags)
        outs.append({b: g.download(b) for b in (F.FPOS, F.FVEL, F.FCONC, F.FELASTIDX, F.FGCELL)})
        g.close()
    for b in outs[0]:
        assert_bits_equal(outs[0][b], outs[1][b], F.BUFFER_NAMES[b])


def test_edge_cases(built):
    from morphogenesis_amd.fluid_system import FluidSystem
    from oracle.mg_oracle import Oracle

    cases = {
        "all_outside": mini_scene(np.full((5, 3), 0.2), b=2),
        "single": mini_scene([[5.0, 5.0, 5.0]], b=0),
        "one_cell_crowd": mini_scene(np.random.default_rng(1).uniform(5.0, 5.99, (70, 3)), b=3),
        "nan_particle": mini_scene([[5.0, 5.0, 5.0], [np.nan, 5.0, 5.0], [5.5, 5.2, 5.0]], b=1),
    }
    for name, sc in cases.items():
        say("case: edge", name)
        g, o = pair(sc)
        for sim in (g, o):
            sim.Run2PhysicalSort()
            sim.step(3, 1e-4, 0)
        compare_all(g, o, name)


def test_full_size_headline_workload(built):
    """The 262,144-particle bonded tissue (B = 32) at its real size: the first sort's counts,
    offsets and order are bit-exact, and 3 full steps match the oracle bit for bit."""
    sc = scenes.tissue(64, 32, seed=1)
    say("case: full-size tissue 262k")
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
    compare_all(g, o, "initial sort", [F.FBIN_COUNT, F.FBIN_OFFSET, F.FGCELL, F.FGNDX,
                                       F.FPARTICLE_ID, F.FELASTIDX, F.FPARTICLEIDX])
    g.step(3, sc.dt, sc.flags)
    o.step(3, sc.dt, sc.flags)
    compare_all(g, o, "3 steps", [F.FPOS, F.FVEL, F.FVEVAL, F.FPRESS, F.FDENSITY, F.FCONC,
                                  F.FEPIGEN, F.FELASTIDX, F.FGCELL, F.FBIN_OFFSET])


def test_cpp_facade_matches_capi(built):
    """The C++ FluidSystem facade (Run2PhysicalSort + n x [Run2InnerPhysicalLoop,
    Run2GeneAction, Run2PhysicalSort]) gives the same state as mg_step."""
    from morphogenesis_amd.fluid_system import FluidSystem

    lib = F.load_library()
    lib.mgfs_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int]
    lib.mgfs_set_host.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]
    lib.mgfs_get_host.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]
    lib.mgfs_setup.argtypes = [C.c_void_p, C.POINTER(F.MgGrid), C.POINTER(F.MgParams),
                               C.POINTER(F.MgGenome), C.c_double]
    lib.mgfs_transfer_to.argtypes = [C.c_void_p]
    lib.mgfs_transfer_from.argtypes = [C.c_void_p]
    lib.mgfs_run.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.mgfs_destroy.argtypes = [C.c_void_p]
    sc = scenes.tissue(n_side=10, b=16, seed=12)
    say("case: C++ facade")
    fh = C.c_void_p()
    assert lib.mgfs_create(C.byref(fh), sc.n, sc.b) == 0
    assert lib.mgfs_setup(fh, C.byref(sc.grid), C.byref(sc.params), C.byref(sc.genome), sc.dt) == 0
    for b, arr in sc.buffers.items():
        a = np.ascontiguousarray(arr)
        assert lib.mgfs_set_host(fh, b, a.ctypes.data_as(C.c_void_p), a.nbytes) == 0
    assert lib.mgfs_transfer_to(fh) == 0
    assert lib.mgfs_run(fh, 15, 1) == 0
    assert lib.mgfs_transfer_from(fh) == 0
    g = FluidSystem(sc.n, sc.b)
    g.load_scene(sc)
    g.Run2PhysicalSort()
    g.step(15, sc.dt, sc.flags)
    for b in (F.FPOS, F.FVEL, F.FCONC, F.FEPIGEN, F.FELASTIDX, F.FGCELL):
        shape, dt = F.buffer_shape(b, sc.n, sc.c, sc.b)
        a = np.empty(shape, dt)
        assert lib.mgfs_get_host(fh, b, a.ctypes.data_as(C.c_void_p), a.nbytes) == 0
        assert_bits_equal(a, g.download(b), f"facade {F.BUFFER_NAMES[b]}")
    lib.mgfs_destroy(fh)


def test_graph_replay_matches_eager(built):
    """mg_step with MG_STEP_GRAPH (two fused steps captured once, replayed) equals eager
    stepping bit for bit, also with eager work, downloads and syncs between replays."""
    from morphogenesis_amd.fluid_system import FluidSystem

    sc = scenes.tissue(n_side=12, b=32, seed=17)
    say("case: graph replay")
    sims = []
    for _ in range(2):
        s = FluidSystem(sc.n, sc.b)
        s.load_scene(sc)
        s.Run2PhysicalSort()
        sims.append(s)
    eager, graph = sims
    eager.step(40, sc.dt, sc.flags)
    for chunk in (10, 7, 1, 2, 20):
        graph.step(chunk, sc.dt, sc.flags | F.STEP_GRAPH)
        _ = graph.download(F.FPOS)  # eager work and a sync between replays
        graph.sync()
    for b in COMPARE:
        assert_bits_equal(graph.download(b), eager.download(b), f"graph {F.BUFFER_NAMES[b]}")


@pytest.mark.parametrize("mask", [0x0000, 0x0006, 0x8008, 0xF0F0, 0x1111, 0xFFFF])
def test_diffusion_group_plans(built, mask):
    """Diffusing-TF sets that select each FCONC group layout of the diffusion sweep (no group,
    one group, two apart, two adjacent, all four with one TF each, all sixteen), through
    the fused step and through the phase path."""
    sc = scenes.tissue(n_side=8, b=6, seed=31)
    for t in range(F.NUM_TF):
        sc.genome.tf_diffusability[t] = (t + 2) if (mask >> t) & 1 else 0
        sc.genome.tf_breakdown_rate[t] = 1
    say("case: diffusion plan mask", hex(mask))
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
        sim.step(12, sc.dt, sc.flags)
    compare_all(g, o, f"fused, mask {mask:#06x}", [F.FPOS, F.FCONC, F.FEPIGEN, F.FGCELL])
    for sim in (g, o):
        sim.Run2InnerPhysicalLoop(sc.dt)
        sim.Run2GeneAction()
        sim.Run2PhysicalSort()
    compare_all(g, o, f"phases, mask {mask:#06x}", [F.FPOS, F.FCONC, F.FEPIGEN, F.FGCELL])


def test_dense_cells_band_fallback(built):
    """Spacing 0.25 puts 64 particles in a cell: a wavefront's neighbour band outgrows the LDS
    stage and the sweeps read global memory; results still match bit for bit."""
    sc = scenes.tissue(n_side=16, b=6, seed=4, spacing=0.25, vmax=12.0)
    say("case: dense cells (band fallback)", sc.n)
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
        sim.step(10, sc.dt, sc.flags)
    compare_all(g, o, "dense cells", [F.FPOS, F.FVEL, F.FPRESS, F.FDENSITY, F.FCONC, F.FGCELL])


def test_phase_calls_after_advance(built):
    """Pressure and force called on positions that moved since the last sort (no culling may
    apply) and diffusion called right after a sort (culling applies)."""
    sc = scenes.tissue(n_side=9, b=8, seed=12)
    say("case: phases after advance")
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
        sim.ComputeDiffusionCL()
        sim.Run2InnerPhysicalLoop(sc.dt)
        sim.ComputePressureCL()
        sim.ComputeForceCL()
    compare_all(g, o, "after advance", [F.FPOS, F.FFORCE, F.FPRESS, F.FDENSITY, F.FCONC])


def test_batch_matches_single(built):
    """Instances of three size classes (different genomes, bond counts and grids) stepped as
    one batch match the same instances stepped alone, bit for bit, through graph replay, odd
    step counts, an autotuning pass (the solo instances then restart from the batched state)
    and a second batch call; one instance is also checked against the oracle."""
    from morphogenesis_amd.batch import Batch
    from morphogenesis_amd.fluid_system import FluidSystem
    from oracle.mg_oracle import Oracle

    specs = [(9, 6, 41), (12, 32, 42), (21, 8, 43), (30, 8, 44), (13, 4, 45)]
    scs = [scenes.tissue(n_side=n, b=b, seed=sd) for n, b, sd in specs]
    say("case: batch of", [sc.n for sc in scs])
    solo, inb = [], []
    for sc in scs:
        for lst in (solo, inb):
            s = FluidSystem(sc.n, sc.b)
            s.load_scene(sc)
            s.Run2PhysicalSort()
            lst.append(s)
    bat = Batch(inb)
    assert [m for m, _ in bat.classes()] == [3, 1, 1]
    dt, fl = scs[0].dt, scs[0].flags
    bufs = [F.FPOS, F.FVEL, F.FVEVAL, F.FPRESS, F.FDENSITY, F.FCONC, F.FEPIGEN, F.FELASTIDX,
            F.FGCELL, F.FGNDX, F.FBIN_COUNT, F.FBIN_OFFSET, F.FPARTICLE_ID]

    def compare(what):
        for k, (a, b) in enumerate(zip(solo, inb)):
            assert a.counts() == b.counts(), (what, k)
            for buf in bufs:
                assert_bits_equal(b.download(buf), a.download(buf), f"{what}, instance {k}: {F.BUFFER_NAMES[buf]}")

    bat.step(7, dt, fl | F.STEP_GRAPH)
    bat.sync()
    for s in solo:
        s.step(7, dt, fl)
    compare("7 steps")
    # autotuning takes real steps of the batch (both launch forms); the solo instances restart
    # from the batched state
    chosen = bat.autotune(2, dt, fl)
    assert all(t in (64, 128, 256) for t in chosen)
    bat.sync()
    for k, sc in enumerate(scs):
        solo[k].close()
        s = FluidSystem(sc.n, sc.b)
        s.set_grid(sc.grid)
        s.set_params(sc.params)
        s.set_genome(sc.genome)
        s.TransferToCL(inb[k].TransferFromCL())
        s.Run2PhysicalSort()
        solo[k] = s
    bat.step(5, dt, fl | F.STEP_GRAPH)
    bat.sync()
    for s in solo:
        s.step(5, dt, fl)
    compare("after autotuning, 5 steps")
    o = Oracle(scs[1].n, scs[1].b)
    o.set_grid(scs[1].grid)
    o.set_params(scs[1].params)
    o.set_genome(scs[1].genome)
    o.TransferToCL(solo[1].TransferFromCL())
    o.Run2PhysicalSort()
    solo[1].step(3, dt, fl)
    o.step(3, dt, fl)
    compare_all(solo[1], o, "restarted instance vs oracle", [F.FPOS, F.FVEL, F.FCONC, F.FGCELL])
    bat.close()


def test_mixed_slab_and_fluid_column(built):
    """The mixed scene (bonded slab + falling fluid column with dense droplets: 1 to ~60
    particles per cell) at 65,536 particles, 12 steps, bit for bit."""
    sc = scenes.mixed(65536, 32, seed=2)
    say("case: mixed slab + column", sc.n, sc.info["n_slab"], sc.info["n_droplets"])
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
    compare_all(g, o, "mixed initial sort", [F.FBIN_COUNT, F.FBIN_OFFSET, F.FGCELL, F.FPARTICLE_ID])
    for sim in (g, o):
        sim.step(12, sc.dt, sc.flags)
    compare_all(g, o, "mixed 12 steps", [F.FPOS, F.FVEL, F.FPRESS, F.FDENSITY, F.FCONC, F.FEPIGEN,
                                         F.FELASTIDX, F.FGCELL, F.FBIN_OFFSET])


def test_smoothing_radius_not_a_power_of_two(built):
    """psmoothradius = 0.75 takes the correctly rounded r / sr path of the pressure sweep (a
    power-of-two radius scales exactly); both must match the oracle bit for bit."""
    sc = scenes.tissue(n_side=8, b=8, seed=17)
    sc.params.psmoothradius = 0.75
    sc.params.r2 = 0.75 * 0.75
    say("case: smoothing radius 0.75")
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
        sim.step(10, sc.dt, sc.flags)
    compare_all(g, o, "sr 0.75", [F.FPOS, F.FVEL, F.FPRESS, F.FDENSITY, F.FCONC])


@pytest.mark.parametrize("speed", [0.0, 60.0])
def test_diffusion_skin_masks_and_fallback(built, speed):
    """The diffusion sweep of the fused step reads the pressure sweep's masks (recorded within a
    skin of 0.01) when no particle moves more than 0.004 in the step, and tests its candidates
    itself otherwise: a particle at 60 units/s (0.006 per step) takes the second path. Both
    match the oracle bit for bit, through hipGraph replay."""
    sc = scenes.tissue(n_side=12, b=16, seed=31)
    if speed:
        sc.buffers[F.FVEL][100, :3] = [speed, 0.0, 0.0]
        sc.buffers[F.FVEVAL][100, :3] = [speed, 0.0, 0.0]
    say("case: diffusion skin, fast particle speed", speed)
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
    g.step(10, sc.dt, sc.flags | F.STEP_GRAPH)
    o.step(10, sc.dt, sc.flags)
    compare_all(g, o, f"skin speed {speed}", [F.FPOS, F.FVEL, F.FCONC, F.FEPIGEN, F.FPRESS, F.FDENSITY,
                                               F.FGCELL, F.FBIN_OFFSET])


@pytest.mark.parametrize("persistent", [1, 0])
def test_batch_launch_forms_match_single(built, persistent):
    """The persistent launch (one kb_step per step: every phase of every instance, grid barriers
    between phases) and the dispatch form, each forced, match the same instances stepped alone
    bit for bit, through graph replay and eager steps of both parities."""
    from morphogenesis_amd.batch import Batch
    from morphogenesis_amd.fluid_system import FluidSystem

    specs = [(9, 6, 51), (12, 32, 52), (21, 8, 53), (30, 8, 54), (13, 4, 55), (10, 16, 56)]
    scs = [scenes.tissue(n_side=n, b=b, seed=sd) for n, b, sd in specs]
    say("case: batch launch form persistent =", persistent, [sc.n for sc in scs])
    solo, inb = [], []
    for sc in scs:
        for lst in (solo, inb):
            s = FluidSystem(sc.n, sc.b)
            s.load_scene(sc)
            s.Run2PhysicalSort()
            lst.append(s)
    bat = Batch(inb)
    bat.set_persistent(bool(persistent))
    dt, fl = scs[0].dt, scs[0].flags
    bat.step(6, dt, fl | F.STEP_GRAPH)
    bat.step(3, dt, fl)
    bat.step(4, dt, fl | F.STEP_GRAPH)
    bat.sync()
    on, wg = bat.launch_config()
    assert on == persistent and (wg > 0) == bool(persistent)
    for s in solo:
        s.step(13, dt, fl)
    bufs = [F.FPOS, F.FVEL, F.FVEVAL, F.FPRESS, F.FDENSITY, F.FCONC, F.FEPIGEN, F.FELASTIDX,
            F.FGCELL, F.FGNDX, F.FBIN_COUNT, F.FBIN_OFFSET, F.FPARTICLE_ID]
    for k, (a, b) in enumerate(zip(solo, inb)):
        assert a.counts() == b.counts(), k
        for buf in bufs:
            assert_bits_equal(b.download(buf), a.download(buf), f"instance {k}: {F.BUFFER_NAMES[buf]}")
    bat.close()


def test_batch_tile_shapes_match_single(built):
    """Every pressure sweep tile shape the batch autotuner chooses from (candidate batch 8 or 4,
    staged band of 128 or 64 slots; a 64-slot band sends most rows of an 8-per-cell tissue to
    the global-memory path) steps the instances of three size classes bit for bit as alone."""
    from morphogenesis_amd.batch import Batch
    from morphogenesis_amd.fluid_system import FluidSystem

    specs = [(9, 6, 61), (21, 8, 62), (30, 8, 63)]
    scs = [scenes.tissue(n_side=n, b=b, seed=sd) for n, b, sd in specs]
    say("case: batch tile shapes", [sc.n for sc in scs])
    solo, inb = [], []
    for sc in scs:
        for lst in (solo, inb):
            s = FluidSystem(sc.n, sc.b)
            s.load_scene(sc)
            s.Run2PhysicalSort()
            lst.append(s)
    bat = Batch(inb)
    bat.set_persistent(False)
    dt, fl = scs[0].dt, scs[0].flags
    ncls = len(bat.classes())
    bufs = [F.FPOS, F.FVEL, F.FVEVAL, F.FPRESS, F.FDENSITY, F.FCONC, F.FBIN_COUNT, F.FBIN_OFFSET]
    for kb, cap in [(8, 128), (4, 128), (8, 64), (4, 64)]:
        for c in range(ncls):
            bat.set_tile_shape(c, kb, cap)
        assert all(cc["cand_batch"] == kb and cc["band_cap"] == cap for cc in bat.class_config())
        bat.step(4, dt, fl | F.STEP_GRAPH)
        bat.sync()
        for s in solo:
            s.step(4, dt, fl)
        for k, (a, b) in enumerate(zip(solo, inb)):
            for buf in bufs:
                assert_bits_equal(b.download(buf), a.download(buf), f"shape {kb}/{cap}, instance {k}: {F.BUFFER_NAMES[buf]}")
    with pytest.raises(F.MgError):
        bat.set_tile_shape(0, 6, 128)
    bat.close()


def row_windows(sim, grid):
    """Slot counts of every particle's 9 neighbour rows (3 consecutive cells each) in the
    sorted state."""
    cnt = sim.download(F.FBIN_COUNT).astype(np.int64)
    off = np.concatenate([sim.download(F.FBIN_OFFSET).astype(np.int64), [0]])
    off[-1] = off[-2] + cnt[-1]
    rx, rz = int(grid.res[0]), int(grid.res[2])
    cell = np.repeat(np.arange(cnt.size), cnt)
    out = []
    for dy in (-1, 0, 1):
        for dz in (-1, 0, 1):
            c0 = np.clip(cell - 1 + (dy * rz + dz) * rx, 0, cnt.size - 3)
            out.append(off[c0 + 3] - off[c0])
    return np.stack(out, 1)


def test_near_coincident_particles(built):
    """Two pairs of particles 1e-20 apart near the origin of a domain around it: their squared
    distance (1e-40) lies below 2^-96, so the wavefronts holding them take the compiler's whole
    correctly rounded sqrt and 1/x rather than the short forms (one pair is bonded, so the bond
    term does too). Bit-equal to the oracle through the phase calls (the force evaluating its
    bonds) and fused steps (the bond records)."""
    rng = np.random.default_rng(11)
    pos = scenes.lattice(6, 6, 6, 0.5, (-1.25, -1.25, -1.25), 0.05, rng)
    extra = np.array([[1e-20, 0.1, 0.1], [2e-20, 0.1, 0.1],
                      [0.3, -3e-20, 0.2], [0.3, -1e-20, 0.2]], np.float32)
    pos = np.concatenate([pos.astype(np.float32), extra])
    n = len(pos)
    sc = mini_scene(pos, b=4, vmin=-6.0, vmax=6.0, bonds=[(n - 4, n - 3, 0.5, 1000.0, 1.0, 2.0)])
    d = pos[n - 4] - pos[n - 3]
    assert 0.0 < float((d * d).sum(dtype=np.float32)) < 2.0 ** -96
    say("case: near-coincident particles", n)
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
    assert g.counts()[1] == n
    for name, fg, fo in (("pressure", g.ComputePressureCL, o.ComputePressureCL),
                         ("force", g.ComputeForceCL, o.ComputeForceCL)):
        fg(); fo()
        compare_all(g, o, f"near-coincident: {name}")
    for sim in (g, o):
        sim.Run2PhysicalSort()
    g.step(6, sc.dt, sc.flags | F.STEP_GRAPH)
    o.step(6, sc.dt, sc.flags)
    compare_all(g, o, "near-coincident: 6 fused steps")


def test_crowded_cell_and_unbinned_low_indices(built):
    """The sort's ranks within a cell: one cell holding 700 particles (runs far longer than a
    wavefront) beside a jittered lattice, and the first 40 particles in the grid's outer layer
    of cells, which the insert leaves unbinned. The binned particles with the highest indices
    then lie past mNumPoints, so the reset tail keeps their ranks as FGNDX. Bit-equal to the
    oracle after the phase sort and after fused steps."""
    rng = np.random.default_rng(5)
    outer = np.column_stack([rng.uniform(0.05, 0.95, 40), rng.uniform(3.0, 8.0, 40),
                             rng.uniform(3.0, 8.0, 40)])
    crowd = rng.uniform(5.05, 5.95, (700, 3))
    lat = scenes.lattice(8, 8, 8, 0.5, (2.2, 2.2, 6.2), 0.05, rng)
    pos = np.concatenate([outer, crowd, lat]).astype(np.float32)
    n = len(pos)
    sc = mini_scene(pos, b=4)
    say("case: crowded cell and unbinned low indices", n)
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
    nump, _ = g.counts()
    assert g.counts() == o.counts() and nump == n - 40
    cnt = o.download(F.FBIN_COUNT)
    assert cnt.max() >= 700
    compare_all(g, o, "crowded cell: phase sort")
    g.step(3, sc.dt, sc.flags | F.STEP_GRAPH)
    o.step(3, sc.dt, sc.flags)
    compare_all(g, o, "crowded cell: 3 fused steps")


def recorded_per_row(sim, grid, r2, sample=3000, seed=0):
    """For a sample of active particles, the number of candidates of each of its 9 neighbour
    rows within the hit-mask skin (0 < dsq < 1.0201 r2): what the pressure sweep records."""
    rng = np.random.default_rng(seed)
    nump, na = sim.counts()
    pos = sim.download(F.FPOS)[:na, :3].astype(np.float32)
    cell = sim.download(F.FGCELL)[:na].astype(np.int64)
    cnt = sim.download(F.FBIN_COUNT).astype(np.int64)
    off = sim.download(F.FBIN_OFFSET).astype(np.int64)
    rx, rz = int(grid.res[0]), int(grid.res[2])
    out = []
    for i in rng.choice(na, size=min(sample, na), replace=False):
        row = []
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                c0 = cell[i] - 1 + (dy * rz + dz) * rx
                if c0 < 0 or c0 + 3 > cnt.size:
                    row.append(0)
                    continue
                lo, hi = off[c0], off[c0 + 2] + cnt[c0 + 2]
                d = pos[lo:hi] - pos[i]
                dsq = (d * d).sum(1)
                row.append(int(((dsq > 0) & (dsq < np.float32(1.0201) * np.float32(r2))).sum()))
        out.append(row)
    return np.array(out)


@pytest.mark.parametrize("speed", [0.0, 60.0])
def test_rows_of_many_recorded_candidates(built, speed):
    """About 16 particles per cell (lattice spacing 0.4): neighbour rows hold up to 64 slots, so
    one hit mask covers them, and central rows record more than 12 candidates, where the
    force and diffusion sweeps (and the pressure sweep's readers) fetch each recorded
    candidate one visit ahead. Bit-equal to the oracle through hipGraph replay."""
    sc = scenes.tissue(n_side=16, b=6, seed=7, spacing=0.4, vmax=12.0)
    if speed:
        sc.buffers[F.FVEL][300, :3] = [speed, 0.0, 0.0]
        sc.buffers[F.FVEVAL][300, :3] = [speed, 0.0, 0.0]
    say("case: rows of many recorded candidates, speed", speed, sc.n)
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
    w = row_windows(g, sc.grid)
    rec = recorded_per_row(o, sc.grid, float(sc.params.r2))
    say("  row windows max", int(w.max()), "| recorded per row max", int(rec.max()),
        "share of sampled particles with a row over 12:", round(float((rec.max(1) > 12).mean()), 3))
    assert (w <= 64).mean() > 0.9 and (rec.max(1) > 12).mean() > 0.3
    g.step(8, sc.dt, sc.flags | F.STEP_GRAPH)
    o.step(8, sc.dt, sc.flags)
    compare_all(g, o, f"many recorded candidates, speed {speed}",
                [F.FPOS, F.FVEL, F.FVEVAL, F.FCONC, F.FEPIGEN, F.FPRESS, F.FDENSITY, F.FGCELL,
                 F.FBIN_OFFSET])


@pytest.mark.parametrize("speed", [0.0, 60.0])
def test_rows_wider_than_one_hit_mask(built, speed):
    """27 particles per cell (lattice spacing 1/3): most neighbour rows hold 65-128 slots, more
    than one 64-bit hit mask covers, so the force and diffusion sweeps test those rows'
    candidates themselves while the other rows use the masks; the cells of more than 8
    particles take the flat row walk (with a fast particle, the diffusion sweep of its
    neighbourhood tests every row itself). Bit-equal to the oracle through hipGraph replay."""
    sc = scenes.tissue(n_side=15, b=6, seed=5, spacing=1.0 / 3.0, vmax=12.0)
    if speed:
        sc.buffers[F.FVEL][200, :3] = [speed, 0.0, 0.0]
        sc.buffers[F.FVEVAL][200, :3] = [speed, 0.0, 0.0]
    say("case: rows wider than one hit mask, speed", speed, sc.n)
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
    w = row_windows(g, sc.grid)
    frac = ((w > 64) & (w <= 128)).mean()
    say("  rows of 65-128 slots:", round(float(frac), 3), "max", int(w.max()))
    assert frac > 0.3
    g.step(8, sc.dt, sc.flags | F.STEP_GRAPH)
    o.step(8, sc.dt, sc.flags)
    compare_all(g, o, f"wide rows, speed {speed}", [F.FPOS, F.FVEL, F.FVEVAL, F.FCONC, F.FEPIGEN,
                                                         F.FPRESS, F.FDENSITY, F.FGCELL, F.FBIN_OFFSET])
