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
    X[:, :50] = rng.normal(0, 1, (rows, 50))
    T = rng.normal(0, 1, (rows, 3)).astype(np.float32)
    M = (rng.random(rows) > 0.1).astype(np.float32)
    return X, T, M


def identity_norm():
    mu = np.zeros(64, np.float32)
    inv = np.zeros(64, np.float32)
    inv[:50] = 1.0
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
    ref = R.forward_bf16(p, xn)[:, :3]
    scale = np.abs(ref).max()
    assert np.abs(y - ref).max() <= 2e-3 * scale, np.abs(y - ref).max() / scale
    # against plain fp64 arithmetic the difference is the bf16 rounding, a few 1e-3
    ref64, _ = R.forward(p, xn)
    assert np.abs(y - ref64[:, :3]).max() <= 3e-2 * scale
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
    assert np.all(w["W1"][50:] == 0) and np.all(w["W3"][:, 3:] == 0)
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
    sim.ComputePressureCL()
    X = features(sim)
    nump, na = sim.counts()
    act = X[:na]
    mu = np.zeros(64, np.float32)
    inv = np.zeros(64, np.float32)
    mu[:50] = act[:, :50].mean(0)
    inv[:50] = 1.0 / np.maximum(act[:, :50].std(0), 1e-6)
    net.set_norm(mu, inv, 0.5, in_limit=1e30, out_limit=1e30)
    _declare(sim.lib)
    DV = np.empty((sim.n, 4), np.float32)
    sim._call("net_infer", sim.h, net.h, DV.ctypes.data_as(__import__("ctypes").POINTER(__import__("ctypes").c_float)))
    p = net.get_params().astype(np.float64)
    ref = R.forward_bf16(p, R.normalise(X[:na], mu, inv))[:, :3] * 0.5
    used = DV[:na, 3] != 0
    assert used.all()
    assert np.abs(DV[:na, :3] - ref).max() <= 2e-3 * np.abs(ref).max() + 1e-7
    assert np.all(DV[na:] == 0)
    # tight limits: every particle falls back to the plain step
    net.set_norm(mu, inv, 0.5, in_limit=1e-6, out_limit=1e30)
    sim._call("net_infer", sim.h, net.h, DV.ctypes.data_as(__import__("ctypes").POINTER(__import__("ctypes").c_float)))
    assert np.all(DV[:na] == 0)
    assert sim.learned_fallbacks() >= na
    net.close()
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

    say("case: training data from fine/coarse pairs and Adam on the GPU")
    sc = scenes.tissue(n_side=10, b=16, seed=6)
    data = make_training_data(sc, samples=12, ratio=4, gap=2, warm=4)
    assert data.X.shape == (12 * sc.n, 64) and data.mask.sum() > 0.9 * 12 * sc.n
    assert np.abs(data.T[data.mask != 0]).max() > 0
    net = Net(seed=7)
    loss = net.fit(data, steps=300, batch=4096, lr=2e-3)
    say("loss", loss[:3], loss[-3:])
    assert np.all(np.isfinite(loss))
    assert loss[-20:].mean() < 0.5 * loss[:5].mean()
    net.close()
