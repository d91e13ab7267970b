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
