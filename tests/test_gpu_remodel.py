"""GPU remodelling (mg_compute_bond_changes, mg_prefix_sum_changes, mg_counting_sort_changes,
mg_compute_particle_changes; the reference's Run2Remodelling) against the oracle: bond records,
change counts, offsets, lengths, lists and healed bonds bit-equal, and the remodelled state
stepping on bit-equal."""
import sys

import numpy as np
import pytest

from morphogenesis_amd import _ffi as F
from tests.test_gpu_parity import COMPARE, compare_all, pair
from tests.test_remodel import remodel_scene

pytestmark = pytest.mark.gpu

ALL = COMPARE + F.CHANGE_BUFFERS


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("seed,b", [(21, 12), (22, 32)])
def test_remodelling_cycles_match_oracle(built, seed, b):
    """The reference's loop order: inner loops without a sort, gene action, remodelling, sort."""
    sc = remodel_scene(seed, n_side=8, b=b)
    say("case: remodelling cycles", sc.name, "seed", seed, "B", b)
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
    healed = 0
    for cycle in range(3):
        for sim in (g, o):
            for _ in range(4):
                sim.Run2InnerPhysicalLoop(sc.dt)
            sim.Run2GeneAction()
            sim.ComputeBondChangesCL(4)
        compare_all(g, o, f"cycle {cycle}: bond changes", ALL)
        for sim in (g, o):
            sim.PrefixSumChangesCL(1)
            sim.CountingSortChangesCL()
        compare_all(g, o, f"cycle {cycle}: change lists", ALL)
        healed += int(o.download(F.FDENSE_LIST_LENGTHS_CHANGES)[0])
        for sim in (g, o):
            sim.ComputeParticleChangesCL(1000 + cycle)
        compare_all(g, o, f"cycle {cycle}: heal", ALL)
        for sim in (g, o):
            sim.Run2PhysicalSort()
        compare_all(g, o, f"cycle {cycle}: sort", ALL)
    assert healed > 100
    say("phase: 20 fused steps")
    for sim in (g, o):
        sim.step(20, sc.dt, sc.flags | (F.STEP_GRAPH if sim is g else 0))
    compare_all(g, o, "20 fused steps after remodelling", ALL)


def test_remodelling_after_fused_steps_32k(built):
    """32,768 particles, B = 32: fused steps, then Run2Remodelling in the sorted state."""
    sc = remodel_scene(23, n_side=32, b=32)
    say("case: remodelling after fused steps", sc.name, sc.n)
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
        sim.step(6, sc.dt, sc.flags | (F.STEP_GRAPH if sim is g else 0))
        sim.Run2Remodelling(6, seed=77)
    compare_all(g, o, "remodelled", ALL)
    assert o.download(F.FDENSE_LIST_LENGTHS_CHANGES)[0] > 1000
    for sim in (g, o):
        sim.step(10, sc.dt, sc.flags | (F.STEP_GRAPH if sim is g else 0))
    compare_all(g, o, "10 fused steps after remodelling", ALL)


def test_remodelling_errors(built):
    sc = remodel_scene(24, n_side=5, b=4)
    say("case: remodelling argument and list checks")
    g, o = pair(sc)
    g.Run2PhysicalSort()
    with pytest.raises(F.MgError):
        g.ComputeBondChangesCL(0)
    g.ComputeBondChangesCL(1)
    cc = g.download(F.FBIN_COUNT_CHANGES)
    cc[3, 5] = 2
    g.upload(F.FBIN_COUNT_CHANGES, cc)
    g.PrefixSumChangesCL(1)
    assert g.download(F.FDENSE_LIST_LENGTHS_CHANGES)[3] == 2
    before = g.download(F.FELASTIDX)
    with pytest.raises(F.MgError):
        g.ComputeParticleChangesCL(1)
    assert np.array_equal(g.download(F.FELASTIDX), before)


def test_remodelling_demo_scene_from_reference_files(built):
    """The reference's demo inputs (20 living particles, B = 6, its genome: the elastin
    lengthen row is all zero, so every remodel breaks each elastin bond and heal reconnects
    it) through the reference's loop with remodelling, 12 cycles."""
    from morphogenesis_amd import fileio as fio
    from morphogenesis_amd.fluid_system import FluidSystem
    from oracle.mg_oracle import Oracle
    from tests.test_fileio import GENOME, PARAMS, POINTS

    say("case: demo scene with remodelling (SimParams.txt, genome.csv, point file)")
    g = FluidSystem.from_reference_files(PARAMS, POINTS, GENOME)
    sc = fio.load_scene(PARAMS, POINTS, GENOME)
    o = Oracle(sc.n, sc.b)
    o.load_scene(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
    healed = 0
    for cycle in range(12):
        for sim in (g, o):
            for _ in range(3):
                sim.Run2InnerPhysicalLoop(sc.dt)
            sim.Run2GeneAction()
            sim.Run2Remodelling(3, seed=cycle)
        compare_all(g, o, f"demo cycle {cycle}: remodelled", ALL)
        healed += int(o.download(F.FDENSE_LIST_LENGTHS_CHANGES)[0])
        for sim in (g, o):
            sim.Run2PhysicalSort()
        compare_all(g, o, f"demo cycle {cycle}: sort", ALL)
    assert healed >= 12 * 18


def test_heal_dense_and_sparse_neighbourhoods(built):
    """Heal where a dense cluster gives ~100 candidates per search (the staged fallback) and a
    sparse lattice gives few (one candidate per lane): bonds formed, one slot in three of the
    living particles broken, Run2Remodelling; bit-equal to the oracle."""
    from morphogenesis_amd.fluid_system import FluidSystem
    from oracle.mg_oracle import Oracle
    from tests.test_gpu_bonds import dense_sparse_scene

    sc = dense_sparse_scene(b=12, living=0.8)
    say("case: heal, dense cluster and sparse lattice,", sc.n, "particles, B 12")
    g = FluidSystem(sc.n, sc.b)
    g.load_scene(sc)
    o = Oracle(sc.n, sc.b)
    o.load_scene(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
        sim.InitializeBondsCL()
    compare_all(g, o, "dense and sparse: bonds formed", ALL)
    rec = o.download(F.FELASTIDX)
    rng = np.random.default_rng(3)
    broken = rng.random(sc.n) < 0.33
    rec[broken, rng.integers(0, sc.b, sc.n)[broken], 2] = 0
    for sim in (g, o):
        sim.upload(F.FELASTIDX, rec)
        sim.Run2Remodelling(2, seed=5)
    compare_all(g, o, "dense and sparse: remodelled", ALL)
    assert o.download(F.FDENSE_LIST_LENGTHS_CHANGES)[0] > 300


@pytest.mark.parametrize("spokes", [200, 300, 1700])
def test_hub_particle_incoming_lists(built, spokes):
    """One particle that every other one bonds to: 200 incoming bonds take the device build's
    wavefront rank sort (several lane passes), 300 its workgroup sort for long runs (a group of
    four partners over 1,024 entries), 1,700 the host build (runs whose group may exceed 6,144
    entries). Steps bit-equal to the oracle either way, particle IDs shuffled so the incoming
    order is not the home order."""
    from tests.helpers import mini_scene

    rng = np.random.default_rng(spokes)
    n = spokes + 1
    hub = np.array([6.0, 6.0, 6.0])
    pos = [hub] + [hub + rng.normal(0, 1, 3) / 1.0 for _ in range(spokes)]
    pos = np.clip(np.array(pos), 3.5, 8.5)
    sc = mini_scene(pos, b=2)
    ids = rng.permutation(n).astype(np.uint32) + 1
    sc.buffers[F.FPARTICLE_ID][:] = ids
    rec = sc.buffers[F.FELASTIDX]          # FPARTICLEIDX cannot hold more than B incoming: left empty
    f = lambda x: np.float32(x).view(np.uint32)
    for k in range(1, n):
        rest = float(np.linalg.norm(pos[k] - hub)) + 0.01
        rec[k, 0] = [0, f(4.0), f(rest), f(5.0), f(0.1), ids[0], f(0), f(0), 0]
    say("case: hub particle with", spokes, "incoming bonds")
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
        sim.step(12, sc.dt, sc.flags | (F.STEP_GRAPH if sim is g else 0))
    compare_all(g, o, f"hub {spokes}: 12 steps")


def test_cpp_facade_remodelling(built):
    """The C++ FluidSystem's Run2Remodelling (seed = its call count) inside the reference's loop
    equals the same sequence through the C-ABI."""
    import ctypes as C

    from morphogenesis_amd.fluid_system import FluidSystem
    from tests.helpers import assert_bits_equal

    lib = F.load_library()
    lib.mgfs_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int]
    lib.mgfs_set_host.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]
    lib.mgfs_get_host.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]
    lib.mgfs_setup.argtypes = [C.c_void_p, C.POINTER(F.MgGrid), C.POINTER(F.MgParams),
                               C.POINTER(F.MgGenome), C.c_double]
    lib.mgfs_transfer_to.argtypes = [C.c_void_p]
    lib.mgfs_transfer_from.argtypes = [C.c_void_p]
    lib.mgfs_run_remodel.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    lib.mgfs_destroy.argtypes = [C.c_void_p]
    sc = remodel_scene(25, n_side=8, b=16)
    say("case: C++ facade Run2Remodelling")
    fh = C.c_void_p()
    assert lib.mgfs_create(C.byref(fh), sc.n, sc.b) == 0
    assert lib.mgfs_setup(fh, C.byref(sc.grid), C.byref(sc.params), C.byref(sc.genome), sc.dt) == 0
    for b, arr in sc.buffers.items():
        a = np.ascontiguousarray(arr)
        assert lib.mgfs_set_host(fh, b, a.ctypes.data_as(C.c_void_p), a.nbytes) == 0
    assert lib.mgfs_transfer_to(fh) == 0
    assert lib.mgfs_run_remodel(fh, 3, 4, 1) == 0
    assert lib.mgfs_transfer_from(fh) == 0
    g = FluidSystem(sc.n, sc.b)
    g.load_scene(sc)
    g.Run2PhysicalSort()
    for cycle in range(3):
        for _ in range(4):
            g.Run2InnerPhysicalLoop(sc.dt)
        g.Run2GeneAction()
        g.Run2Remodelling(4, seed=cycle)
        g.Run2PhysicalSort()
    for b in (F.FPOS, F.FVEL, F.FCONC, F.FEPIGEN, F.FELASTIDX, F.FGCELL):
        shape, dt = F.buffer_shape(b, sc.n, sc.c, sc.b)
        a = np.empty(shape, dt)
        assert lib.mgfs_get_host(fh, b, a.ctypes.data_as(C.c_void_p), a.nbytes) == 0
        assert_bits_equal(a, g.download(b), f"facade remodel {F.BUFFER_NAMES[b]}")
    lib.mgfs_destroy(fh)
