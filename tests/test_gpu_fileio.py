"""GPU runs of scenes loaded from the reference's file formats (the demo inputs in
tests/golden/demo/) against the oracle, and exact resume from a checkpoint."""
import os
import sys

import numpy as np
import pytest

from morphogenesis_amd import _ffi as F
from morphogenesis_amd import fileio as fio
from morphogenesis_amd import scenes
from tests.helpers import assert_bits_equal
from tests.test_fileio import GENOME, PARAMS, POINTS, shuffled_ids

pytestmark = pytest.mark.gpu

ALL = fio.CHECKPOINT_BUFFERS


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same_state(x, y, what):
    assert x.counts() == y.counts(), (what, x.counts(), y.counts())
    for b in ALL:
        assert_bits_equal(x.download(b), y.download(b), f"{what}: {F.BUFFER_NAMES[b]}")


@pytest.mark.parametrize("flags", [0, F.STEP_GENES])
def test_demo_scene_from_reference_files(built, flags):
    from morphogenesis_amd.fluid_system import FluidSystem
    from oracle.mg_oracle import Oracle

    say("case: demo scene (SimParams.txt, genome.csv, 20-particle point file), flags", flags)
    g = FluidSystem.from_reference_files(PARAMS, POINTS, GENOME)
    sc = fio.load_scene(PARAMS, POINTS, GENOME)
    o = Oracle(sc.n, sc.b)
    o.load_scene(sc)
    assert (g.n, g.b, g.c) == (20, 6, 10000)
    for sim in (g, o):
        sim.Run2PhysicalSort()
    p0 = g.download(F.FPOS)[:, :3].copy()
    for chunk in range(4):           # 400 steps of m_DT
        g.step(100, sc.dt, flags)
        o.step(100, sc.dt, flags)
        same_state(g, o, f"demo after {100 * (chunk + 1)} steps")
    # the cluster moves (pressure and gravity) and stays inside the scan window
    assert g.counts() == (20, 20)
    assert np.abs(g.download(F.FPOS)[:, :3] - p0).max() > 0.1


def test_shuffled_particle_ids_bit_exact(built):
    from morphogenesis_amd.fluid_system import FluidSystem
    from oracle.mg_oracle import Oracle

    sc = shuffled_ids(scenes.tissue(n_side=9, b=12, seed=7), 3)
    say("case: tissue with particle IDs in random order", sc.name)
    g, o = FluidSystem(sc.n, sc.b), Oracle(sc.n, sc.b)
    for sim in (g, o):
        sim.load_scene(sc)
        sim.Run2PhysicalSort()
    g.step(40, sc.dt, sc.flags | F.STEP_GRAPH)
    o.step(40, sc.dt, sc.flags)
    same_state(g, o, "shuffled IDs, 40 steps")


@pytest.mark.parametrize("first", [40, 2050])
def test_checkpoint_resume_is_exact(built, tmp_path, first):
    """Save after `first` steps, run 40 more; a fresh system from the checkpoint runs the
    same 40 steps to the same bits. At first = 2050 the uninterrupted run re-homes its
    records at the start of the second call, the resumed one does not."""
    from morphogenesis_amd.fluid_system import FluidSystem
    from oracle.mg_oracle import Oracle

    sc = shuffled_ids(scenes.tissue(n_side=10, b=16, seed=4), 9)
    say("case: checkpoint resume after", first, "steps")
    a = FluidSystem(sc.n, sc.b)
    a.load_scene(sc)
    a.Run2PhysicalSort()
    a.step(first, sc.dt, sc.flags | F.STEP_GRAPH)
    path = a.save_checkpoint(str(tmp_path / "state"))
    a.step(40, sc.dt, sc.flags | F.STEP_GRAPH)
    b = FluidSystem.from_checkpoint(path)
    b.step(40, sc.dt, sc.flags | F.STEP_GRAPH)
    same_state(a, b, f"resumed after {first}")
    if first == 40:
        o = Oracle(sc.n, sc.b)
        o.load_scene(sc)
        o.Run2PhysicalSort()
        o.step(80, sc.dt, sc.flags)
        same_state(b, o, "resumed GPU run vs oracle")


def test_point_file_round_trip_through_the_gpu(built, tmp_path):
    from morphogenesis_amd.fluid_system import FluidSystem

    say("case: SavePointsCSV2 / WriteSimParams / WriteGenome of a GPU state and reload")
    g = FluidSystem.from_reference_files(PARAMS, POINTS, GENOME)
    g.Run2PhysicalSort()
    g.step(50, g.m_DT, 0)
    d = str(tmp_path)
    pts = g.SavePointsCSV2(d, 50)
    g.WriteSimParams(d)
    g.WriteGenome(d)
    # the demo cluster's pressure drives particles past PVEL_LIMIT = 3, which the reference's
    # reader rejects; the bounds check is tested on the CPU
    h = FluidSystem.from_reference_files(os.path.join(d, "SimParams.txt"), pts,
                                         os.path.join(d, "genome.csv"), check_bounds=False)
    assert np.float32(h.m_Time) == np.float32(np.float32(g.m_Time))
    for k in (F.FPARTICLE_ID, F.FELASTIDX, F.FEPIGEN, F.FMASS_RADIUS):
        assert np.array_equal(h.download(k), g.download(k)), F.BUFFER_NAMES[k]
    pg, ph = g.download(F.FPOS)[:, :3], h.download(F.FPOS)[:, :3]
    assert np.abs(pg - ph).max() <= 5e-7 * max(1.0, np.abs(pg).max())
    h.Run2PhysicalSort()
    h.step(10, h.m_DT, 0)


def test_vtp_of_a_gpu_state(built, tmp_path):
    import xml.etree.ElementTree as ET

    from morphogenesis_amd.fluid_system import FluidSystem

    say("case: SavePointsVTP2 of the demo scene after 20 steps")
    g = FluidSystem.from_reference_files(PARAMS, POINTS, GENOME)
    g.Run2PhysicalSort()
    g.step(20, g.m_DT, 0)
    p = g.SavePointsVTP2(str(tmp_path), 20)
    piece = ET.parse(p).getroot().find("PolyData/Piece")
    assert int(piece.get("NumberOfPoints")) == g.counts()[0] + 8
    pts = np.array(piece.find("Points/DataArray").text.split(), np.float32).reshape(-1, 3)
    assert np.array_equal(pts[:20], g.download(F.FPOS)[:20, :3])
    assert pts[-1].tolist() == [10.0, 20.0, 50.0]
