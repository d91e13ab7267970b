"""GPU tests of the learned coarse step: the BF16-MFMA kernels against the FP64 reference of
oracle/mlp_ref.py, the fused inference against the training forward pass, the zero network
against the plain step, and training on the GPU."""
import sys

import numpy as np
import pytest

from morphogenesis_amd import _ffi as F
from morphogenesis_amd import scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def random_data(rows, seed):
    rng = np.random.default_rng(seed)
    X = np.zeros((rows, 64), np.float32)
    X[:, :53] = rng.normal(0, 1, (rows, 53))
    T = rng.normal(0, 1, (rows, 6)).astype(np.float32)
    M = (rng.random(rows) > 0.1).astype(np.float32)
    return X, T, M


def identity_norm():
    mu = np.zeros(64, np.float32)
    inv = np.zeros(64, np.float32)
    inv[:53] = 1.0
    return mu, inv


def test_mlp_forward_matches_fp64_reference(built):
    from morphogenesis_amd.learned import Net
    from oracle import mlp_ref as R

    say("case: mlp forward vs fp64 reference")
    net = Net(seed=11)
    X, T, M = random_data(777, 1)      # not a multiple of any tile size
    mu, inv = identity_norm()
    net.set_norm(mu, inv, 1.0)
    net.load_data(X, T, M)
    y = net.forward(0, len(X))
    p = net.get_params().astype(np.float64)
    xn = R.normalise(X, mu, inv)
    ref = R.forward_bf16(p, xn)[:, :6]
    scale = np.abs(ref).max()
    assert np.abs(y - ref).max() <= 2e-3 * scale, np.abs(y - ref).max() / scale
    # against plain fp64 arithmetic the difference is the bf16 rounding, a few 1e-3
    ref64, _ = R.forward(p, xn)
    assert np.abs(y - ref64[:, :6]).max() <= 3e-2 * scale
    net.close()


def test_mlp_gradients_match_fp64_reference(built):
    from morphogenesis_amd.learned import Net
    from oracle import mlp_ref as R

    say("case: mlp gradients vs fp64 reference")
    net = Net(seed=12)
    X, T, M = random_data(3000, 2)
    mu, inv = identity_norm()
    net.set_norm(mu, inv, 1.0)
    net.load_data(X, T, M)
    loss, g = net.grad(0, len(X))
    p = net.get_params().astype(np.float64)
    ref_loss, ref_g = R.loss_and_grad(p, R.normalise(X, mu, inv), T, M)
    assert abs(loss - ref_loss) <= 2e-2 * abs(ref_loss)
    for name, (a, b) in {k: (R.unpack(g)[k], R.unpack(ref_g)[k]) for k, _ in R.SHAPES}.items():
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        assert err < 5e-2, (name, err)
    # padding rows of W1 and padding columns of W3 get exactly zero gradient
    w = R.unpack(g)
    assert np.all(w["W1"][53:] == 0) and np.all(w["W3"][:, 6:] == 0)
    net.close()


def tissue_handle(n_side=10, b=16, seed=3):
    from morphogenesis_amd.fluid_system import FluidSystem

    sc = scenes.tissue(n_side=n_side, b=b, seed=seed)
    sim = FluidSystem(sc.n, sc.b)
    sim.load_scene(sc)
    sim.Run2PhysicalSort()
    sim.step(5, sc.dt, sc.flags)
    return sc, sim


def test_fused_inference_matches_reference_and_falls_back(built):
    from morphogenesis_amd.learned import Net, _declare, features
    from oracle import mlp_ref as R

    say("case: fused inference (features + 3 MFMA layers) vs reference")
    sc, sim = tissue_handle()
    net = Net(seed=13)
    X = features(sim)
    nump, na = sim.counts()
    act = X[:na]
    assert np.isfinite(act[:, :53]).all() and np.abs(act[:, 50:53]).max() > 0
    mu = np.zeros(64, np.float32)
    inv = np.zeros(64, np.float32)
    mu[:53] = act[:, :53].mean(0)
    inv[:53] = 1.0 / np.maximum(act[:, :53].std(0), 1e-6)
    _declare(sim.lib)
    P = __import__("ctypes").POINTER(__import__("ctypes").c_float)
    DV = np.empty((sim.n, 8), np.float32)
    dt = 4 * sc.dt
    scale = np.linspace(0.2, 0.6, 6).astype(np.float32)
    net.set_norm(mu, inv, scale, in_limit=1e30, out_limit=1e30)
    sim._call("net_infer", sim.h, net.h, dt, DV.ctypes.data_as(P))
    p = net.get_params().astype(np.float64)
    ref = R.forward_bf16(p, R.normalise(X[:na], mu, inv))[:, :6] * scale
    got = DV[:na][:, [0, 1, 2, 4, 5, 6]]
    assert np.all(DV[:na, 3] == 1) and np.all(DV[:na, 7] == 0)
    assert np.abs(got - ref).max() <= 2e-3 * np.abs(ref).max() + 1e-7
    assert np.all(DV[na:] == 0)
    # the same state gives the same inputs again (the dry force pass changes no record)
    np.testing.assert_array_equal(features(sim), X)
    # tight input limits: every particle falls back to the plain step
    net.set_norm(mu, inv, scale, in_limit=1e-6, out_limit=1e30)
    sim._call("net_infer", sim.h, net.h, dt, DV.ctypes.data_as(P))
    assert np.all(DV[:na] == 0)
    assert sim.learned_fallbacks() >= na
    # a correction that would move a particle more than 1e-3 in the step: fallback
    net.set_norm(mu, inv, scale, in_limit=1e30, out_limit=1e30)
    big = float(2e-3 / np.linalg.norm(ref[:, :3], axis=1).min())
    sim._call("net_infer", sim.h, net.h, big, DV.ctypes.data_as(P))
    assert np.all(DV[:na] == 0)
    net.close()
    sim.close()


def expected_inputs(sim):
    """Inputs 0..49 of every active particle restated from the downloaded reference-layout
    buffers: up to 8 bonded neighbours, first the particle's own bond slots in slot order,
    then the bonds other particles hold to it ordered by (owner particle ID, slot), skipping
    partners outside the active range and bonds with rest length below 1e-9; (x_j - x_i,
    v_j - v_i) each; then density (1 / FDENSITY) and pressure."""
    nump, na = sim.counts()
    pos = sim.download(F.FPOS)[:, :3]
    vel = sim.download(F.FVEL)[:, :3]
    rec = sim.download(F.FELASTIDX)
    ids = sim.download(F.FPARTICLE_ID)
    dens = sim.download(F.FDENSITY)
    press = sim.download(F.FPRESS)
    partner = rec[:, :, 0]
    rest = rec[:, :, 2].view(np.float32).astype(np.float64)
    B = rec.shape[1]
    incoming = [[] for _ in range(na)]
    for k in range(na):
        for a in range(B):
            j = int(partner[k, a])
            if j < na and rest[k, a] >= 1e-9:
                incoming[j].append((int(ids[k]), a, k))
    X = np.zeros((na, 50), np.float32)
    for i in range(na):
        nb = [int(partner[i, a]) for a in range(B) if partner[i, a] < na and rest[i, a] >= 1e-9]
        nb += [k for (_, _, k) in sorted(incoming[i])]
        for q, j in enumerate(nb[:8]):
            X[i, 6 * q:6 * q + 3] = pos[j] - pos[i]
            X[i, 6 * q + 3:6 * q + 6] = vel[j] - vel[i]
        X[i, 48] = np.float32(1.0) / dens[i]
        X[i, 49] = press[i]
    return X


def test_inputs_match_numpy_restatement(built):
    """The feature gather (shared by training and inference) against a numpy restatement,
    bit for bit, on a tissue with shuffled particle IDs and some broken bonds."""
    from morphogenesis_amd.learned import features

    say("case: learned-step inputs vs numpy restatement")
    from tests.test_fileio import shuffled_ids

    sc = shuffled_ids(scenes.tissue(n_side=9, b=16, seed=8), 5)
    rng = np.random.default_rng(8)
    rec = sc.buffers[F.FELASTIDX]
    used = rec[:, :, 0] != F.NONE
    brk = used & (rng.random(used.shape) < 0.15)
    r = rec[:, :, 2].view(np.float32)
    r[brk] = 0.0
    from morphogenesis_amd.fluid_system import FluidSystem

    sim = FluidSystem(sc.n, sc.b)
    sim.load_scene(sc)
    sim.Run2PhysicalSort()
    sim.step(3, sc.dt, sc.flags)
    X = features(sim)
    nump, na = sim.counts()
    want = expected_inputs(sim)
    np.testing.assert_array_equal(X[:na, :50], want)
    assert np.count_nonzero(X[:na, :48]) > 0
    sim.close()


def test_zero_network_is_the_plain_step(built):
    from morphogenesis_amd.fluid_system import FluidSystem
    from morphogenesis_amd.learned import Net

    say("case: zero network == plain coarse step")
    sc = scenes.tissue(n_side=9, b=8, seed=4)
    a, b = FluidSystem(sc.n, sc.b), FluidSystem(sc.n, sc.b)
    for s in (a, b):
        s.load_scene(sc)
        s.Run2PhysicalSort()
    net = Net(seed=1)
    net.set_params(np.zeros(net.nparams, np.float32))
    a.step_learned(net, 6, 4 * sc.dt, sc.flags | F.STEP_GRAPH)
    a.step_learned(net, 1, 4 * sc.dt, sc.flags)
    b.step(7, 4 * sc.dt, sc.flags)
    for buf in (F.FPOS, F.FVEL, F.FVEVAL, F.FCONC, F.FGCELL):
        x, y = a.download(buf), b.download(buf)
        if x.dtype == np.float32:
            assert np.array_equal(x, y), F.BUFFER_NAMES[buf]   # +0 == -0
        else:
            assert np.array_equal(x, y), F.BUFFER_NAMES[buf]
    net.close()


def test_training_on_gpu_reduces_the_loss(built):
    from morphogenesis_amd.learned import Net, make_training_data

    say("case: training data from fine trajectory segments and Adam on the GPU")
    sc = scenes.tissue(n_side=10, b=16, seed=6)
    data = make_training_data(sc, samples=12, ratio=4, gap=2, warm=4)
    assert data.segments == 12 and data.X.shape[1] == 64
    assert len(data.X) > 0.9 * 12 * sc.n
    assert np.abs(data.T).max() > 0
    net = Net(seed=7)
    loss = net.fit(data, steps=300, batch=4096, lr=2e-3)
    say("loss", loss[:3], loss[-3:])
    assert np.all(np.isfinite(loss))
    assert loss[-20:].mean() < 0.5 * loss[:5].mean()
    net.close()


def test_learned_step_beats_the_plain_coarse_step(built):
    """The ablation: on a 4,096-particle tissue, trained on fine segments of two other tissues,
    200 learned steps of 4 dt end closer to 800 fine steps of dt than 200 plain steps of 4 dt
    (displacement-normalised position error, less than half of the plain step's), with few
    fallbacks."""
    from morphogenesis_amd.learned import Net, make_training_data, rollout_errors

    say("case: learned vs plain coarse (ablation), 4096-particle tissue")
    train = [scenes.tissue(n_side=16, b=32, seed=s) for s in (41, 42)]
    data = make_training_data(train, samples=25, ratio=4, gap=32)
    net = Net(seed=9)
    loss = net.fit(data, steps=1500, batch=8192, lr=1e-3)
    sc = scenes.tissue(n_side=16, b=32, seed=43)
    r = rollout_errors(sc, net, 200, ratio=4)
    say("loss", float(loss[:5].mean()), "->", float(loss[-50:].mean()), "errors", r["learned_rel_disp"],
        "vs coarse", r["coarse_rel_disp"], "fallbacks", r["fallbacks"])
    assert r["learned_lost"] == 0 and r["coarse_lost"] == 0
    assert r["learned_rel_disp"] < 0.5 * r["coarse_rel_disp"]
    assert r["fallbacks"] <= 0.01 * sc.n * 200
    net.close()
