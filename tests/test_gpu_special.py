"""Special particles (MG_STEP_SPECIAL: externalActuation + fixedParticles) on the GPU against
the oracle, eager (time-dependent actuation) and graph-replayed (fixed particles only)."""
import sys

import pytest

from morphogenesis_amd import _ffi as F
from tests.test_gpu_parity import compare_all, pair
from tests.test_special import special_scene

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("a_f,graph,genes", [(0.01, False, False), (0.01, True, True), (0.0, True, False)])
def test_special_particles_match_oracle(built, a_f, graph, genes):
    sc, fixed, act, both = special_scene(a_f=a_f)
    say("case: special particles a_f", a_f, "graph", graph, "genes", genes)
    g, o = pair(sc)
    for sim in (g, o):
        sim.Run2PhysicalSort()
    flags = F.STEP_SPECIAL | (F.STEP_GENES if genes else 0)
    for chunk in range(3):
        g.step(20, sc.dt, flags | (F.STEP_GRAPH if graph else 0))
        o.step(20, sc.dt, flags)
        compare_all(g, o, f"special particles after {20 * (chunk + 1)} steps")


def test_batch_refuses_special(built):
    from morphogenesis_amd.batch import Batch
    from morphogenesis_amd.fluid_system import FluidSystem

    sc, *_ = special_scene()
    say("case: batch with MG_STEP_SPECIAL is refused")
    s = FluidSystem(sc.n, sc.b)
    s.load_scene(sc)
    s.Run2PhysicalSort()
    b = Batch([s])
    with pytest.raises(F.MgError):
        b.step(2, sc.dt, F.STEP_SPECIAL)
    b.close()
    s.close()
