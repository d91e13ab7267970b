"""Special particles (SpecialParticlesCL: externalActuation + fixedParticles,
src/fluid_system_cuda.cu:3617-3705) in the oracle against their closed forms: fixed particles
keep their state, actuated ones move only in z by actuation_factor * sin(t / (period * dt))."""
import math

import numpy as np

from morphogenesis_amd import _ffi as F
from morphogenesis_amd import scenes

INT_MAX = 2147483647


def special_scene(a_f=0.01, a_p=100.0, seed=3):
    sc = scenes.tissue(n_side=6, b=8, seed=seed, genes=False)
    sc.params.actuation_factor = a_f
    sc.params.actuation_period = a_p
    epi = sc.buffers[F.FEPIGEN]
    ids = sc.buffers[F.FPARTICLE_ID]
    fixed = (ids % 7 == 0)
    act = (ids % 5 == 0)                  # IDs divisible by 35 are both: fixed wins
    epi[11] = np.where(fixed, INT_MAX, 0)
    epi[12] = np.where(act, INT_MAX, 0)
    return sc, ids[fixed], ids[act & ~fixed], ids[act & fixed]


def by_id(sim):
    ids = sim.download(F.FPARTICLE_ID)
    n = sim.counts()[0]
    order = np.argsort(ids[:n])
    return {int(k): s for k, s in zip(ids[:n][order], order)}


def test_fixed_and_actuated_particles_follow_their_closed_forms(built):
    from oracle.mg_oracle import Oracle

    sc, fixed, act, both = special_scene()
    o = Oracle(sc.n, sc.b)
    o.load_scene(sc)
    o.Run2PhysicalSort()
    s0 = by_id(o)
    p0, v0, e0 = o.download(F.FPOS), o.download(F.FVEL), o.download(F.FVEVAL)
    steps = 30
    o.step(steps, sc.dt, F.STEP_SPECIAL)
    s1 = by_id(o)
    p1, v1, e1 = o.download(F.FPOS), o.download(F.FVEL), o.download(F.FVEVAL)
    for k in list(fixed) + list(both):
        assert np.array_equal(p1[s1[k], :3], p0[s0[k], :3]) and np.array_equal(v1[s1[k]], v0[s0[k]])
        assert np.array_equal(e1[s1[k], :3], e0[s0[k], :3])
    # z displacement: sum over steps of a_f * sin(t_k / (a_p * dt)), t_k = k dt
    dz = sum(0.01 * math.sin((k * sc.dt) / (100.0 * sc.dt)) for k in range(steps))
    na = o.counts()[1]
    checked = 0
    for k in act:
        assert np.array_equal(p1[s1[k], :2], p0[s0[k], :2])
        if s1[k] < na:               # particles outside the scan window do not advance at all
            assert abs((p1[s1[k], 2] - p0[s0[k], 2]) - dz) < 2e-6
            checked += 1
    assert checked > 0.5 * len(act)
    # everybody else moved
    others = [k for k in s0 if k not in set(fixed) | set(act) | set(both)]
    moved = sum(not np.array_equal(p1[s1[k], :3], p0[s0[k], :3]) for k in others)
    assert moved > 0.9 * len(others)


def test_without_the_flag_special_genes_do_nothing(built):
    from oracle.mg_oracle import Oracle

    sc, fixed, act, both = special_scene()
    o = Oracle(sc.n, sc.b)
    o.load_scene(sc)
    o.Run2PhysicalSort()
    s0 = by_id(o)
    p0 = o.download(F.FPOS)
    o.step(10, sc.dt, 0)
    s1 = by_id(o)
    p1 = o.download(F.FPOS)
    assert sum(not np.array_equal(p1[s1[k], :3], p0[s0[k], :3]) for k in fixed) > 0.9 * len(fixed)
