"""GPU parity of the selectable Mueller (2003) SPH kernels (MG_SPH_POLY6_SPIKY) with the FP32
oracle, bit for bit: the 32,768-particle fluid block (the CPU-runnable fluid configuration
names poly6/spiky kernels) for 100 steps, a bonded tissue with genes through the phase calls
and the fused step, a switch of kernels between graph-replayed steps, and the refusal of a
batch."""
import sys

import numpy as np
import pytest

from morphogenesis_amd import _ffi as F
from morphogenesis_amd import scenes
from tests.helpers import assert_bits_equal

pytestmark = pytest.mark.gpu

STATE = [F.FPOS, F.FVEL, F.FVEVAL, F.FFORCE, F.FPRESS, F.FDENSITY, F.FCONC, F.FEPIGEN, F.FELASTIDX,
         F.FGCELL, F.FBIN_COUNT, F.FBIN_OFFSET, F.FPARTICLE_ID]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pair(scene, mode=F.SPH_POLY6_SPIKY):
    from morphogenesis_amd.fluid_system import FluidSystem
    from oracle.mg_oracle import Oracle

    g = FluidSystem(scene.n, scene.b)
    o = Oracle(scene.n, scene.b)
    for sim in (g, o):
        sim.load_scene(scene)
        sim.set_sph_kernels(mode)
    return g, o


def compare(g, o, what, bufs=STATE):
    assert g.counts() == o.counts(), (what, g.counts(), o.counts())
    for b in bufs:
        assert_bits_equal(g.download(b), o.download(b), f"{what}: {F.BUFFER_NAMES[b]}")


def test_fluid_block_100_steps_poly6_spiky(built):
    sc = scenes.fluid_block(32, seed=1)
    say("case: fluid_block_32k, poly6/spiky kernels, 100 steps", sc.n)
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
    for k in range(4):
        g.step(25, sc.dt, sc.flags | F.STEP_GRAPH)
        o.step(25, sc.dt, sc.flags)
        compare(g, o, f"after {25 * (k + 1)} steps", [b for b in STATE if b != F.FFORCE])
    g.close()


def test_tissue_phases_and_fused_steps_poly6_spiky(built):
    sc = scenes.tissue(n_side=10, b=16, seed=21)
    say("case: tissue, poly6/spiky kernels, phases then fused steps", sc.n)
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
    for sim in (g, o):
        sim.Run2InnerPhysicalLoop(sc.dt)
        sim.Run2GeneAction()
        sim.Run2PhysicalSort()
    compare(g, o, "reference loop, one step")
    g.step(30, sc.dt, sc.flags | F.STEP_GRAPH)
    o.step(30, sc.dt, sc.flags)
    compare(g, o, "30 fused steps", [b for b in STATE if b != F.FFORCE])
    g.close()


def test_switching_kernels_between_graph_steps(built):
    sc = scenes.tissue(n_side=9, b=12, seed=22)
    say("case: switch reference -> poly6/spiky -> reference between graph steps", sc.n)
    g, o = pair(sc, F.SPH_REFERENCE)
    for sim in (g, o):
        sim.Run2PhysicalSort()
    for mode, n in ((F.SPH_REFERENCE, 6), (F.SPH_POLY6_SPIKY, 6), (F.SPH_REFERENCE, 4)):
        for sim in (g, o):
            sim.set_sph_kernels(mode)
        g.step(n, sc.dt, sc.flags | F.STEP_GRAPH)
        o.step(n, sc.dt, sc.flags)
        compare(g, o, f"mode {mode}", [b for b in STATE if b != F.FFORCE])
    g.close()


def test_batch_refuses_poly6_spiky(built):
    from morphogenesis_amd.batch import Batch
    from morphogenesis_amd.fluid_system import FluidSystem

    sc = scenes.tissue(n_side=6, b=6, seed=23)
    sims = []
    for k in range(2):
        s = FluidSystem(sc.n, sc.b)
        s.load_scene(sc)
        s.Run2PhysicalSort()
        sims.append(s)
    sims[1].set_sph_kernels(F.SPH_POLY6_SPIKY)
    with pytest.raises(F.MgError):
        Batch(sims)
    sims[1].set_sph_kernels(F.SPH_REFERENCE)
    bat = Batch(sims)
    with pytest.raises(F.MgError):
        sims[0].set_sph_kernels(F.SPH_POLY6_SPIKY)
    bat.close()
