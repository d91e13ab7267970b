"""The learned coarse step: a per-particle MLP correction that lets one step of 4 dt stand in
for four steps of dt (mg_net_* / mg_step_learned in include/mg_sph.h). No counterpart in the
reference.

  net = Net(seed)
  data = make_training_data(scene, samples=200, ratio=4)      # fine vs coarse targets
  net.fit(data, steps=..., batch=..., lr=...)                  # Adam on the GPU
  sim.step_learned(net, n, 4 * dt, flags)                      # inference inside the step

Targets: for a start state s (sorted, after a pressure pass), v_fine - v_coarse per particle,
where v_fine is the velocity after `ratio` fused steps of dt and v_coarse after one fused step
of ratio * dt, both from s, matched by particle ID.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _ffi as F

IN, IN_REAL, OUT_REAL = 64, 50, 3


def _declare(lib):
    P = C.POINTER
    h = C.c_void_p
    sig = {
        "net_create": [P(h), C.c_uint64],
        "net_destroy": [h],
        "net_get_params": [h, P(C.c_float), C.c_size_t],
        "net_set_params": [h, P(C.c_float), C.c_size_t],
        "net_set_norm": [h, P(C.c_float), P(C.c_float), C.c_float, C.c_float, C.c_float],
        "net_load_data": [h, P(C.c_float), P(C.c_float), P(C.c_float), C.c_size_t],
        "net_train": [h, C.c_int, C.c_size_t, C.c_float, C.c_uint64, P(C.c_float)],
        "net_grad": [h, C.c_size_t, C.c_size_t, P(C.c_float), P(C.c_float)],
        "net_forward": [h, C.c_size_t, C.c_size_t, P(C.c_float)],
        "net_features": [h, P(C.c_float)],
        "net_infer": [h, h, P(C.c_float)],
        "step_learned": [h, h, C.c_int, C.c_float, C.c_uint],
        "learned_fallbacks": [h, P(C.c_uint64)],
    }
    for name, args in sig.items():
        fn = getattr(lib, f"mg_{name}")
        fn.argtypes = args
        fn.restype = C.c_int
    lib.mg_net_nparams.argtypes = []
    lib.mg_net_nparams.restype = C.c_size_t
    return lib


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Net:
    def __init__(self, seed: int = 1, lib=None):
        self.lib = _declare(lib or F.load_library())
        self.h = C.c_void_p()
        self._call("net_create", C.byref(self.h), int(seed))
        self.nparams = int(self.lib.mg_net_nparams())
        self.mu = np.zeros(IN, np.float32)
        self.inv_sigma = np.ones(IN, np.float32)
        self.out_scale = 1.0
        self.rows = 0

    def _call(self, name, *args):
        r = getattr(self.lib, f"mg_{name}")(*args)
        if r != F.MG_OK:
            raise F.MgError(f"mg_{name} -> {F.ERRORS.get(r, r)}: {self.lib.mg_last_error().decode()}")
        return r

    def close(self):
        if self.h:
            self.lib.mg_net_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # parameters and normalisation
    def get_params(self) -> np.ndarray:
        p = np.empty(self.nparams, np.float32)
        self._call("net_get_params", self.h, _fp(p), p.size)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._call("net_set_params", self.h, _fp(p), p.size)

    def set_norm(self, mu, inv_sigma, out_scale: float, in_limit: float = 1e30, out_limit: float = 1e30):
        self.mu = np.ascontiguousarray(mu, np.float32)
        self.inv_sigma = np.ascontiguousarray(inv_sigma, np.float32)
        self.out_scale = float(out_scale)
        self._call("net_set_norm", self.h, _fp(self.mu), _fp(self.inv_sigma), self.out_scale,
                   float(in_limit), float(out_limit))

    # training
    def load_data(self, X, T, mask):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        mask = np.ascontiguousarray(mask, np.float32)
        assert X.shape[1] == IN and T.shape[1] == OUT_REAL and len(mask) == len(X)
        self._call("net_load_data", self.h, _fp(X), _fp(T), _fp(mask), len(X))
        self.rows = len(X)

    def train(self, steps: int, batch: int, lr: float, seed: int = 0) -> np.ndarray:
        loss = np.zeros(max(steps, 1), np.float32)
        self._call("net_train", self.h, int(steps), int(batch), float(lr), int(seed), _fp(loss))
        return loss[:steps]

    def grad(self, row0: int, rows: int):
        g = np.empty(self.nparams, np.float32)
        loss = np.zeros(1, np.float32)
        self._call("net_grad", self.h, int(row0), int(rows), _fp(g), _fp(loss))
        return float(loss[0]), g

    def forward(self, row0: int, rows: int) -> np.ndarray:
        y = np.empty((rows, OUT_REAL), np.float32)
        self._call("net_forward", self.h, int(row0), int(rows), _fp(y))
        return y

    def fit(self, data: "TrainingData", steps: int = 3000, batch: int = 8192, lr: float = 1e-3,
            seed: int = 0, in_limit: float = 12.0, out_limit: float = 12.0) -> np.ndarray:
        """Normalise from the data, upload it (rows shuffled) and train with Adam."""
        m = data.mask != 0
        mu = np.zeros(IN, np.float64)
        sd = np.ones(IN, np.float64)
        mu[:IN_REAL] = data.X[m, :IN_REAL].mean(0)
        sd[:IN_REAL] = data.X[m, :IN_REAL].std(0)
        inv = np.where(sd > 1e-12, 1.0 / np.maximum(sd, 1e-30), 0.0)
        inv[IN_REAL:] = 0.0
        out_scale = float(np.sqrt(np.mean(data.T[m] ** 2))) or 1.0
        self.set_norm(mu, inv, out_scale, in_limit, out_limit)
        rng = np.random.default_rng(seed)
        order = rng.permutation(len(data.X))
        self.load_data(data.X[order], data.T[order] / out_scale, data.mask[order])
        return self.train(steps, min(batch, self.rows), lr, seed)


@dataclass
class TrainingData:
    X: np.ndarray      # [rows][64] raw inputs
    T: np.ndarray      # [rows][3] targets v_fine - v_coarse
    mask: np.ndarray   # [rows] 1 for rows to learn from


def features(sim) -> np.ndarray:
    X = np.empty((sim.n, IN), np.float32)
    _declare(sim.lib)
    sim._call("net_features", sim.h, _fp(X))
    return X


def make_training_data(scene, samples: int = 200, ratio: int = 4, gap: int = 4, warm: int = 20,
                       dt: float | None = None, lib=None) -> TrainingData:
    """`samples` start states along a fine trajectory of `scene` (one every `gap` fine
    steps after `warm`), each paired with its fine (ratio steps of dt) and coarse (one step of
    ratio * dt) outcome."""
    from .fluid_system import FluidSystem

    dt = scene.dt if dt is None else dt
    flags = scene.flags
    A = FluidSystem(scene.n, scene.b, lib=lib)
    Bsim = FluidSystem(scene.n, scene.b, lib=lib)
    A.load_scene(scene)
    Bsim.set_grid(scene.grid)
    Bsim.set_params(scene.params)
    Bsim.set_genome(scene.genome)
    A.Run2PhysicalSort()
    A.step(warm, dt, flags)
    Xs, Ts, Ms = [], [], []
    for _ in range(samples):
        A.step(gap, dt, flags)
        snap = A.TransferFromCL()
        A.ComputePressureCL()
        X = features(A)
        nump, na = A.counts()
        ids0 = A.download(F.FPARTICLE_ID)
        # coarse from the same state in the second instance
        Bsim.TransferToCL(snap)
        Bsim.Run2PhysicalSort()
        Bsim.step(1, ratio * dt, flags)
        # fine
        A.step(ratio, dt, flags)
        T = np.zeros((scene.n, OUT_REAL), np.float32)
        M = np.zeros(scene.n, np.float32)
        q = ids0[:na]
        vf, okf = _velocity_by_id(A, q)
        vc, okc = _velocity_by_id(Bsim, q)
        ok = okf & okc
        T[:na][ok] = vf[ok] - vc[ok]
        M[:na][ok] = 1.0
        Xs.append(X)
        Ts.append(T)
        Ms.append(M)
    A.close()
    Bsim.close()
    return TrainingData(np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Ms))


def _velocity_by_id(sim, query_ids):
    """Velocities of the active particles with the given IDs, and which were found."""
    nump, na = sim.counts()
    ids = sim.download(F.FPARTICLE_ID)[:na]
    v = sim.download(F.FVEL)[:na, :3]
    order = np.argsort(ids)
    ids, v = ids[order], v[order]
    k = np.clip(np.searchsorted(ids, query_ids), 0, max(len(ids) - 1, 0))
    found = (len(ids) > 0) & (ids[k] == query_ids) if len(ids) else np.zeros(len(query_ids), bool)
    return v[k], found


def positions_by_id(sim) -> tuple:
    nump, na = sim.counts()
    ids = sim.download(F.FPARTICLE_ID)[:na]
    x = sim.download(F.FPOS)[:na, :3].astype(np.float64)
    order = np.argsort(ids)
    return ids[order], x[order]


def rollout_errors(scene, net, coarse_steps: int, ratio: int = 4, dt: float | None = None, lib=None,
                   graph: bool = True) -> dict:
    """From the scene's sorted start state: `ratio * coarse_steps` fine steps of dt, and
    `coarse_steps` steps of ratio * dt with and without the network. Positions are compared by
    particle ID: rel_disp = |x - x_fine| / |x_fine - x_0| (2-norms over all particles) and
    rel_pos = |x - x_fine| / |x_fine|. Wall times are measured around the stepping calls."""
    import time

    from .fluid_system import FluidSystem

    dt = scene.dt if dt is None else dt
    flags = scene.flags | (F.STEP_GRAPH if graph else 0)
    out = {}
    sims = {}
    for name in ("fine", "learned", "coarse"):
        s = FluidSystem(scene.n, scene.b, lib=lib)
        s.load_scene(scene)
        s.Run2PhysicalSort()
        sims[name] = s
    ids0, x0 = positions_by_id(sims["fine"])
    runs = {"fine": lambda s: s.step(ratio * coarse_steps, dt, flags),
            "learned": lambda s: s.step_learned(net, coarse_steps, ratio * dt, flags),
            "coarse": lambda s: s.step(coarse_steps, ratio * dt, flags)}
    for name, run in runs.items():
        s = sims[name]
        s.sync()
        t = time.perf_counter()
        run(s)
        s.sync()
        out[f"{name}_s"] = time.perf_counter() - t
    ids_f, xf = positions_by_id(sims["fine"])
    disp = np.linalg.norm(xf - x0[np.searchsorted(ids0, ids_f)])
    for name in ("learned", "coarse"):
        ids, x = positions_by_id(sims[name])
        common, ia, ib = np.intersect1d(ids_f, ids, return_indices=True)
        d = np.linalg.norm(x[ib] - xf[ia])
        out[f"{name}_rel_disp"] = float(d / max(disp, 1e-30))
        out[f"{name}_rel_pos"] = float(d / max(np.linalg.norm(xf[ia]), 1e-30))
        out[f"{name}_lost"] = int(len(ids_f) - len(common))
    out["fallbacks"] = sims["learned"].learned_fallbacks()
    out["speedup_vs_fine"] = out["fine_s"] / out["learned_s"]
    out["coarse_steps"] = coarse_steps
    out["ratio"] = ratio
    for s in sims.values():
        s.close()
    return out
