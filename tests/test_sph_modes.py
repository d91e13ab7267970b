"""The selectable Mueller (2003) SPH kernels (MG_SPH_POLY6_SPIKY) in the oracle, pinned by
closed forms: the kernel constants' normalisations, an O(N^2) float64 poly6 density, the
two-particle spiky pressure and viscosity forces of Mueller, Charypar and Gross, "Particle-based
fluid simulation for interactive applications" (SCA 2003), a whole-block brute force and
momentum conservation. No GPU."""
import numpy as np
import pytest

from morphogenesis_amd import _ffi as F
from oracle.mg_oracle import Oracle
from tests.helpers import mini_scene
from tests.test_oracle_pins import jittered_block

RT64 = 5e-7  # FP64 oracle against float64 forms: the float32 downloads bound the agreement


def make(scene, precision=64):
    o = Oracle(scene.n, scene.b, precision=precision)
    o.load_scene(scene)
    o.set_sph_kernels(F.SPH_POLY6_SPIKY)
    return o


def quad(f, a, b, n=20001):
    x = np.linspace(a, b, n)
    y = f(x)
    return (b - a) / (n - 1) * (y[0] / 2 + y[1:-1].sum() + y[-1] / 2)


@pytest.mark.parametrize("sr", [1.0, 1.5])
def test_kernel_constants_normalise(sr):
    """poly6kern normalises (h^2 - r^2)^3 over the ball of radius h; |spikykern| / 3 = 15/(pi h^6)
    normalises the spiky kernel (h - r)^3, whose gradient magnitude is |spikykern| (h - r)^2;
    lapkern = -spikykern is the Laplacian constant of the viscosity kernel."""
    p = F.params_from_fluid_param(1.0, sr, 0.0, 0.0002, 0.0, (0, 0, 0), (10, 10, 10), 1, 1, 1, 0, 1,
                                  0, 0, 0, 0, 0, -2, 0, 1000, 1000, 0, 1)
    h = float(np.float32(sr))
    w6 = quad(lambda r: 4 * np.pi * float(p.poly6kern) * (h * h - r * r) ** 3 * r * r, 0.0, h)
    ws = quad(lambda r: 4 * np.pi * abs(float(p.spikykern)) / 3 * (h - r) ** 3 * r * r, 0.0, h)
    assert abs(w6 - 1) < 2e-6 and abs(ws - 1) < 2e-6, (w6, ws)
    assert float(p.lapkern) == -float(p.spikykern)
    assert float(p.vterm) == float(np.float32(np.float32(p.lapkern) * np.float32(p.pvisc)))


def poly6_density(pos, sr, params):
    """O(N^2) float64: pmass poly6kern sum over 0 < r < sr of (sr^2 - r^2)^3; an empty sum -> 1."""
    x = pos.astype(np.float64)
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    r2 = float(np.float32(sr) * np.float32(sr))
    m = (d2 < r2) & (d2 > 0)
    s = np.where(m, (r2 - d2) ** 3, 0.0).sum(1) * params.pmass * params.poly6kern
    s[s == 0] = 1.0
    return s


@pytest.mark.parametrize("sr", [1.0, 1.5])
def test_poly6_density_matches_bruteforce(built, sr):
    pos = jittered_block(7, 0.5 * sr, (5.0 * sr, 5.0 * sr, 5.0 * sr), seed=41)
    sc = mini_scene(pos, b=0, vmax=14.0 * sr, sr=sr)
    for prec, rtol in ((64, 2e-7), (32, 2e-6)):
        o = make(sc, precision=prec)
        o.Run2PhysicalSort()
        nump, na = o.counts()
        assert na == len(pos)
        o.ComputePressureCL()
        x = o.download(F.FPOS)[:na, :3]
        want = poly6_density(x, sr, sc.params)
        got = 1.0 / o.download(F.FDENSITY)[:na].astype(np.float64)
        np.testing.assert_allclose(got, want, rtol=rtol)
        press = o.download(F.FPRESS)[:na].astype(np.float64)
        np.testing.assert_allclose(press, (want - sc.params.prest_dens) * sc.params.pintstiff,
                                   rtol=1e-5, atol=1e-5 * np.abs(press).max())


def pair(r, sr, vel1=(0.0, 0.0, 0.0), **over):
    u = np.array([0.36, 0.48, 0.8])
    x0 = np.array([5.0, 5.0, 5.0]) * sr
    pos = np.array([x0, x0 + r * u])
    vel = np.array([[0.0, 0.0, 0.0], vel1])
    sc = mini_scene(pos, b=0, vmax=12.0 * sr, sr=sr, vel=vel, params_over=over)
    o = make(sc)
    o.Run2PhysicalSort()
    o.ComputePressureCL()
    o.ComputeForceCL()
    pid = o.download(F.FPARTICLE_ID)
    i0, i1 = (0, 1) if pid[0] == 1 else (1, 0)
    x = o.download(F.FPOS)[:, :3].astype(np.float64)
    f = o.download(F.FFORCE)[:, :3].astype(np.float64)
    rho_inv = o.download(F.FDENSITY).astype(np.float64)
    p = o.download(F.FPRESS).astype(np.float64)
    return sc.params, f[i0], f[i1], x[i0] - x[i1], rho_inv[i0] * rho_inv[i1], p[i0] + p[i1]


@pytest.mark.parametrize("sr", [1.0, 1.5])
def test_pair_pressure_force_is_mueller_spiky(built, sr):
    """Equal velocities: the force on particle 0 is m (p_0 + p_1) / (2 rho_0 rho_1) times the
    spiky gradient magnitude |spikykern| (h - r)^2 along x_0 - x_1, so it scales as (h - r)^2."""
    g = {}
    for r in (0.3 * sr, 0.55 * sr, 0.8 * sr):
        P, f0, f1, d, inv_rr, pp = pair(r, sr, istiff=40.0)
        assert abs(pp) > 0
        rr = np.linalg.norm(d)
        want = P.pmass * pp / 2.0 * inv_rr * abs(float(P.spikykern)) * (sr - rr * P.psimscale) ** 2 * d / rr
        np.testing.assert_allclose(f0, want, rtol=RT64, atol=RT64 * np.abs(want).max())
        np.testing.assert_allclose(f1, -want, rtol=RT64, atol=RT64 * np.abs(want).max())
        g[r] = np.linalg.norm(f0) / (P.pmass * abs(pp) * inv_rr)
    a, b = 0.3 * sr, 0.8 * sr
    np.testing.assert_allclose(g[a] / g[b], ((sr - a) / (sr - b)) ** 2, rtol=2e-5)


@pytest.mark.parametrize("sr", [1.0, 1.5])
def test_pair_viscous_force_is_mueller_laplacian(built, sr):
    """Zero pressure: the force on particle 0 is m mu (v_1 - v_0) / (rho_0 rho_1) times the
    viscosity kernel's Laplacian lapkern (h - r), linear in the relative velocity."""
    for r in (0.3 * sr, 0.7 * sr):
        prev = None
        for k in (1.0, 2.0, -3.0):
            dv = k * np.array([0.02, -0.05, 0.03], np.float32)
            P, f0, f1, d, inv_rr, pp = pair(r, sr, vel1=dv, istiff=0.0, visc=0.35)
            assert pp == 0
            rr = np.linalg.norm(d)
            want = (P.pmass * float(P.pvisc) * float(P.lapkern) * (sr - rr * P.psimscale) * inv_rr
                    * dv.astype(np.float64))
            np.testing.assert_allclose(f0, want, rtol=2 * RT64)
            np.testing.assert_allclose(f1, -want, rtol=2 * RT64)
            if prev is not None:
                np.testing.assert_allclose(f0 / k, prev, rtol=RT64)
            prev = f0 / k


def mueller_force(pos, veval, press, rho_inv, P):
    """O(N^2) float64: for 0 < r < h, m [(p_i + p_j) / 2 |spiky| (h - r)^2 (x_i - x_j) / r
    + mu lapkern (h - r) (v_j - v_i)] / (rho_i rho_j), summed over j."""
    x, v = pos.astype(np.float64), veval.astype(np.float64)
    p, ri = press.astype(np.float64), rho_inv.astype(np.float64)
    h = float(P.psmoothradius)
    out, scale = np.zeros((len(x), 3)), np.zeros((len(x), 3))
    for i in range(len(x)):
        d = x[i] - x
        dsq = (d * d).sum(1)
        hit = (dsq < float(P.r2)) & (dsq > 0)
        d, r = d[hit], np.sqrt(dsq[hit])
        c = h - r * P.psimscale
        w = ri[i] * ri[hit]
        pt = ((p[i] + p[hit]) / 2 * abs(float(P.spikykern)) * c * c * w / r)[:, None] * d
        vt = (float(P.pvisc) * float(P.lapkern) * c * w)[:, None] * (v[hit] - v[i])
        terms = P.pmass * (pt + vt)
        out[i], scale[i] = terms.sum(0), np.abs(terms).sum(0)
    return out, scale


@pytest.mark.parametrize("sr", [1.0, 1.25])
def test_block_force_matches_bruteforce_and_conserves_momentum(built, sr):
    rng = np.random.default_rng(23)
    pos = jittered_block(8, 0.5, (5.0 * sr, 5.0 * sr, 5.0 * sr), seed=19, jitter=0.08)
    vel = rng.uniform(-0.3, 0.3, pos.shape)
    sc = mini_scene(pos, b=0, vmax=14.0 * sr, sr=sr, vel=vel, params_over={"istiff": 3.0, "visc": 0.4})
    for prec, tol in ((64, 5e-7), (32, 2e-5)):
        o = make(sc, precision=prec)
        o.Run2PhysicalSort()
        nump, na = o.counts()
        o.ComputePressureCL()
        o.ComputeForceCL()
        x, ve = o.download(F.FPOS)[:na, :3], o.download(F.FVEVAL)[:na, :3]
        pr, de = o.download(F.FPRESS)[:na], o.download(F.FDENSITY)[:na]
        assert np.ptp(pr) > 0
        want, scale = mueller_force(x, ve, pr, de, sc.params)
        got = o.download(F.FFORCE)[:na, :3].astype(np.float64)
        assert np.all(np.abs(got - want) <= tol * scale + 1e-30), np.max(np.abs(got - want) / (scale + 1e-30))
        total = got.sum(0)
        assert np.all(np.abs(total) <= 1e-5 * np.abs(got).sum(0)), total


def test_mode_changes_the_step_and_defaults_to_the_reference(built):
    """The reference kernels stay the default; the mode changes density and force."""
    pos = jittered_block(6, 0.5, (5.0, 5.0, 5.0), seed=3)
    sc = mini_scene(pos, b=0, vmax=14.0)
    a = Oracle(sc.n, sc.b)
    a.load_scene(sc)
    b = make(sc, precision=32)
    for o in (a, b):
        o.Run2PhysicalSort()
        o.ComputePressureCL()
    assert not np.array_equal(a.download(F.FDENSITY), b.download(F.FDENSITY))
    with pytest.raises(F.MgError):
        a.set_sph_kernels(7)
