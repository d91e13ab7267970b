"""GPU bond formation (mg_initialize_bonds, the reference's initialize_bonds) against the oracle:
every buffer bit-equal after forming bonds, and the bonded state stepping on bit-equal."""
import sys

import numpy as np
import pytest

from morphogenesis_amd import _ffi as F
from morphogenesis_amd import scenes
from tests.test_gpu_parity import COMPARE, compare_all, pair

pytestmark = pytest.mark.gpu

INT_MAX = 2147483647


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bond_scene(seed, n_side=8, b=12):
    """Tissue with genes; gene 1 (solid) and the tissue genes 6, 7, 9, 10 on random subsets."""
    sc = scenes.tissue(n_side=n_side, b=b, seed=seed)
    sc.genome = scenes.example_genome()
    rng = np.random.default_rng(seed + 100)
    epi = sc.buffers[F.FEPIGEN]
    for gi, frac in ((1, 0.6), (6, 0.1), (7, 0.1), (9, 0.15), (10, 0.1)):
        epi[gi] = np.where(rng.random(sc.n) < frac, INT_MAX, 0).astype(np.uint32)
    return sc


@pytest.mark.parametrize("seed,pre", [(5, 0), (6, 30)])
def test_initialize_bonds_matches_oracle(built, seed, pre):
    sc = bond_scene(seed)
    say("case: initialize_bonds", sc.name, "seed", seed, "steps before", pre)
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
        if pre:
            sim.step(pre, sc.dt, sc.flags)
    before = o.download(F.FELASTIDX)
    g.InitializeBondsCL()
    o.InitializeBondsCL()
    compare_all(g, o, "after initialize_bonds")
    changed = np.sum(np.any(before != o.download(F.FELASTIDX), axis=2))
    assert changed > 0.3 * sc.n            # most gene-1 particles re-formed bonds
    for sim in (g, o):
        sim.step(20, sc.dt, sc.flags | (F.STEP_GRAPH if sim is g else 0))
    compare_all(g, o, "20 steps after initialize_bonds")


def test_initialize_bonds_needs_the_sorted_state(built):
    from morphogenesis_amd.fluid_system import FluidSystem

    sc = bond_scene(7, n_side=5, b=4)
    say("case: initialize_bonds before any sort is refused")
    g = FluidSystem(sc.n, sc.b)
    g.load_scene(sc)
    with pytest.raises(F.MgError):
        g.InitializeBondsCL()
    g.Run2PhysicalSort()
    g.InitializeBondsCL()


@pytest.mark.parametrize("seed,b", [(8, 32), (9, 40), (10, 64)])
def test_initialize_bonds_slot_counts(built, seed, b):
    """B = 32 takes the search that keeps each candidate's angle terms from slot to slot, B = 40
    and 64 the eight-lane search over staged bonds; all bit-equal."""
    sc = bond_scene(seed, b=b)
    say("case: initialize_bonds", sc.name, "seed", seed, "B", b)
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
        sim.step(10, sc.dt, sc.flags)
    g.InitializeBondsCL()
    o.InitializeBondsCL()
    compare_all(g, o, f"B {b}: after initialize_bonds")
    for sim in (g, o):
        sim.InitializeBondsCL()            # again, over the bonds just formed
    compare_all(g, o, f"B {b}: second initialize_bonds")
    for sim in (g, o):
        sim.step(10, sc.dt, sc.flags | (F.STEP_GRAPH if sim is g else 0))
    compare_all(g, o, f"B {b}: 10 steps after initialize_bonds")


def dense_sparse_scene(b=16, living=0.0):
    """A dense cluster (~100 candidates per bond search) beside a sparse jittered lattice."""
    from tests.helpers import mini_scene

    rng = np.random.default_rng(11)
    dense = rng.uniform(4.0, 6.5, (1200, 3))
    ax = np.arange(7.0, 10.5, 0.5)
    sparse = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    sparse += rng.uniform(-0.05, 0.05, sparse.shape)
    pos = np.concatenate([dense, sparse])
    n = len(pos)
    epi = np.zeros((F.NUM_GENES, n), np.uint32)
    for gi, frac in ((1, 0.7), (2, living), (6, 0.1), (9, 0.15)):
        epi[gi] = np.where(rng.random(n) < frac, INT_MAX, 0).astype(np.uint32)
    return mini_scene(pos, b=b, epigen=epi, genome=scenes.example_genome())


def test_initialize_bonds_dense_and_sparse_neighbourhoods(built):
    """A dense cluster (~100 candidates per search, more than the 32 a particle's lanes hold)
    beside a sparse lattice: both searches in one launch, bit-equal to the oracle."""
    from morphogenesis_amd.fluid_system import FluidSystem
    from oracle.mg_oracle import Oracle

    sc = dense_sparse_scene()
    n = sc.n
    say("case: initialize_bonds, dense cluster and sparse lattice,", n, "particles, B 16")
    g = FluidSystem(sc.n, sc.b)
    g.load_scene(sc)
    o = Oracle(sc.n, sc.b)
    o.load_scene(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
        sim.InitializeBondsCL()
    compare_all(g, o, "dense and sparse: after initialize_bonds")
    rec = o.download(F.FELASTIDX)
    assert np.sum(rec[:, :, 2] != 0) > 0.5 * n * 16 * 0.7
