"""Bond formation (initialize_bonds, src/fluid_system_cuda.cu:1938-2044) in the oracle, pinned by
geometric constructions whose answer is known without the scoring formula, and the acos it
uses against float64 arccos."""
import numpy as np
import pytest

from morphogenesis_amd import _ffi as F
from morphogenesis_amd import scenes
from tests.helpers import mini_scene

INT_MAX = 2147483647


def test_bond_angle_acos_within_two_ulp():
    from oracle import mg_oracle as O

    xs = np.concatenate([np.linspace(-1, 1, 40001), [-1.0, -0.5, 0.0, 0.5, 1.0, 1e-5, -1e-5, 0.4999999, 0.5000001]])
    xs = xs.astype(np.float32)
    got = np.array([O.acosf(x) for x in xs], np.float64)
    ref = np.arccos(xs.astype(np.float64))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(got - ref) / ulp) <= 2.0
    assert O.acosf(1.0) == 0.0 and abs(O.acosf(-1.0) - np.pi) < 3e-7
    assert np.isnan(O.acosf(1.0000001)) and np.isnan(O.acosf(-1.5)) and np.isnan(O.acosf(float("nan")))


def by_id(sim, ids):
    """Sorted slots of the given particle IDs."""
    pid = sim.download(F.FPARTICLE_ID)
    return [int(np.flatnonzero(pid == k)[0]) for k in ids]


def run(pos, b=4, epi_genes=None, bonds=(), ids=None):
    """Oracle over a hand-built scene; particle 0 (ID 1) carries gene 1 (and epi_genes)."""
    from oracle.mg_oracle import Oracle

    n = len(pos)
    epigen = np.zeros((F.NUM_GENES, n), np.uint32)
    epigen[1, 0] = INT_MAX
    for gi in epi_genes or ():
        epigen[gi, 0] = INT_MAX
    sc = mini_scene(pos, b=b, bonds=bonds, epigen=epigen, genome=scenes.example_genome())
    if ids is not None:
        sc.buffers[F.FPARTICLE_ID][:] = ids
    o = Oracle(sc.n, sc.b)
    o.load_scene(sc)
    o.Run2PhysicalSort()
    o.InitializeBondsCL()
    rec = o.download(F.FELASTIDX)
    slot = {k: s for k, s in zip(range(1, n + 1), by_id(o, range(1, n + 1)))} if ids is None else None
    return o, rec, slot


def partners(o, rec, slot_of_a):
    """Partner IDs of the particle at slot_of_a, per bond slot (0 = none)."""
    pid = o.download(F.FPARTICLE_ID)
    out = []
    for w in rec[slot_of_a]:
        out.append(int(pid[w[0]]) if w[0] != F.NONE else 0)
    return out


A = np.array([5.1, 5.1, 5.1])


def test_slot0_is_nearest_and_slot1_prefers_the_tetrahedral_angle():
    # A in the middle of its cell along x; B nearest (0.3 along +x); C 60 deg and D 109.47 deg
    # (cos = -1/3) from AB, both 0.6 away on the +y side, so every neighbour lies in the
    # searched cells (offsets 0..2 from A's cell)
    a = np.array([5.5, 5.1, 5.1])
    t = np.arccos(-1.0 / 3.0)
    pos = [a, a + 0.3 * np.array([1.0, 0.0, 0.0]),
           a + 0.6 * np.array([np.cos(np.radians(60)), np.sin(np.radians(60)), 0.0]),
           a + 0.6 * np.array([np.cos(t), np.sin(t), 0.0])]
    assert min(q[0] for q in pos) >= 5.0 and min(q[1] for q in pos) >= 5.0
    o, rec, slot = run(pos, b=3)
    p = partners(o, rec, slot[1])
    assert p[0] == 2            # nearest
    assert p[1] == 4            # closest to 109.5 deg from the first bond
    assert p[2] == 3            # the remaining neighbour
    # default elastin parameters of the example genome (src/fluid_system.cpp:2441-2452)
    w = rec[slot[1], 0]
    assert w[1:5].view(np.float32).tolist() == [8.0, 0.5, 1000.0, 10.0]
    assert w[5] == 2 and w[6] == 0 and w[7] == 0


def test_search_block_is_offset_upwards():
    # a neighbour 0.15 below in x sits in the cell below A's: the reference's search block is
    # gc + gridAdj (offsets 0..2), so it is never a candidate; one 0.4 above is
    pos = [A, A + np.array([-0.15, 0.0, 0.0]), A + np.array([0.4, 0.0, 0.0])]
    o, rec, slot = run(pos, b=2)
    assert partners(o, rec, slot[1]) == [3, 0]


def test_id_zero_is_never_chosen():
    pos = [A, A + np.array([0.2, 0.0, 0.0]), A + np.array([0.0, 0.5, 0.0])]
    o, rec, _ = run(pos, b=2, ids=np.array([7, 0, 9], np.uint32))
    pid = o.download(F.FPARTICLE_ID)
    a = int(np.flatnonzero(pid == 7)[0])
    got = [int(pid[w[0]]) if w[0] != F.NONE else -1 for w in rec[a]]
    assert got == [9, -1]


@pytest.mark.parametrize("genes,types", [((9,), [2, 2, 2, 2, 0]), ((6,), [1, 0, 0, 0, 0]),
                                         ((7,), [1, 0, 0, 0, 0]), ((10,), [1, 0, 0, 0, 0]),
                                         ((), [0, 0, 0, 0, 0])])
def test_bond_type_by_tissue_gene(genes, types):
    g = scenes.example_genome()
    rng = np.random.default_rng(3)
    pos = [A] + [A + rng.uniform(0.05, 0.45, 3) for _ in range(6)]
    o, rec, slot = run(pos, b=5, epi_genes=genes)
    for s, t in enumerate(types):
        w = rec[slot[1], s]
        assert w[0] != F.NONE
        assert w[1:5].view(np.float32).tolist() == [np.float32(g.param[t][k]) for k in (8, 9, 10, 11)]


def test_existing_partner_excluded_except_in_slot0():
    # A already bonded (slot 0) to its nearest B: slot 0 is re-chosen by distance (B again),
    # slot 1 may not take B
    pos = [A, A + np.array([0.2, 0.0, 0.0]), A + np.array([0.0, 0.5, 0.0])]
    o, rec, slot = run(pos, b=2, bonds=[(0, 1, 0.2, 100.0, 1.0, 2.0)])
    assert partners(o, rec, slot[1]) == [2, 3]


def test_gene1_off_forms_nothing():
    from oracle.mg_oracle import Oracle

    pos = [A, A + np.array([0.2, 0.0, 0.0])]
    sc = mini_scene(pos, b=2)
    o = Oracle(sc.n, sc.b)
    o.load_scene(sc)
    o.Run2PhysicalSort()
    before = o.download(F.FELASTIDX)
    o.InitializeBondsCL()
    assert np.array_equal(before, o.download(F.FELASTIDX))
