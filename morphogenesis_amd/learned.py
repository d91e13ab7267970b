"""The learned coarse step: a per-particle MLP correction that lets one step of 4 dt stand in
for four steps of dt (mg_net_* / mg_step_learned in include/mg_sph.h). No counterpart in the
reference.

  net = Net(seed)
  data = make_training_data([scene, ...], samples=50, ratio=4)   # fine trajectory segments
  net.fit(data, steps=..., batch=..., lr=...)                      # Adam on the GPU
  sim.step_learned(net, n, 4 * dt, flags)                          # inference inside the step

The network outputs two velocity corrections for one coarse step from state s, each the value
a fine segment of four steps of dt reaches minus the plain coarse step's (v_next = v_s +
4 dt a_s, a_s the coarse step's acceleration, an input of the network):
  drift  mean_k(v_k) - v_next   (four fine steps move a particle by dt (v_1 + ... + v_4))
  veval  veval_4 - (v_s + v_next) / 2
both from the fine segment's velocities: they carry the displacement to float32 precision,
where differences of float32 positions would not resolve it. The fine segment's last velocity
v_4 differs from v_next by an amount the local state does not predict (a linear fit on the
inputs and on neighbour-acceleration sums explains under 10% of it on the 262k tissue), so the
velocity keeps the plain step's value.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _ffi as F

IN, IN_REAL, OUT, OUT_REAL = 64, 53, 16, 6


def _declare(lib):
    P = C.POINTER
    h = C.c_void_p
    sig = {
        "net_create": [P(h), C.c_uint64],
        "net_destroy": [h],
        "net_get_params": [h, P(C.c_float), C.c_size_t],
        "net_set_params": [h, P(C.c_float), C.c_size_t],
        "net_set_norm": [h, P(C.c_float), P(C.c_float), P(C.c_float), C.c_float, C.c_float],
        "net_load_data": [h, P(C.c_float), P(C.c_float), P(C.c_float), C.c_size_t],
        "net_train": [h, C.c_int, C.c_size_t, C.c_float, C.c_uint64, P(C.c_float)],
        "net_grad": [h, C.c_size_t, C.c_size_t, P(C.c_float), P(C.c_float)],
        "net_forward": [h, C.c_size_t, C.c_size_t, P(C.c_float)],
        "net_features": [h, P(C.c_float)],
        "net_infer": [h, h, C.c_float, P(C.c_float)],
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
        self.out_scale = np.ones(OUT_REAL, np.float32)
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

    def set_norm(self, mu, inv_sigma, out_scale, in_limit: float = 1e30, out_limit: float = 1e30):
        """out_scale: one scale per output (6), or one for all."""
        self.mu = np.ascontiguousarray(mu, np.float32)
        self.inv_sigma = np.ascontiguousarray(inv_sigma, np.float32)
        self.out_scale = np.ascontiguousarray(np.broadcast_to(np.asarray(out_scale, np.float32), (OUT_REAL,)))
        scale = np.zeros(OUT, np.float32)
        scale[:OUT_REAL] = self.out_scale
        self._call("net_set_norm", self.h, _fp(self.mu), _fp(self.inv_sigma), _fp(scale),
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
            seed: int = 0, in_limit: float = 6.0, out_limit: float = 12.0) -> np.ndarray:
        """Normalise from the data, upload it (rows shuffled) and train with Adam."""
        m = data.mask != 0
        mu = np.zeros(IN, np.float64)
        sd = np.ones(IN, np.float64)
        mu[:IN_REAL] = data.X[m, :IN_REAL].mean(0)
        sd[:IN_REAL] = data.X[m, :IN_REAL].std(0)
        inv = np.where(sd > 1e-12, 1.0 / np.maximum(sd, 1e-30), 0.0)
        inv[IN_REAL:] = 0.0
        # one scale per correction (drift, veval): each trains at unit RMS; a correction whose
        # targets are all zero gets scale 0, i.e. it is switched off
        rms = np.sqrt(np.mean(data.T[m].astype(np.float64) ** 2, axis=0)).reshape(2, 3).mean(1)
        out_scale = np.repeat(rms, 3).astype(np.float32)
        self.set_norm(mu, inv, out_scale, in_limit, out_limit)
        rng = np.random.default_rng(seed)
        order = rng.permutation(len(data.X))
        T = np.where(out_scale > 0, data.T[order] / np.where(out_scale > 0, out_scale, 1.0), 0.0)
        self.load_data(data.X[order], T.astype(np.float32), data.mask[order])
        return self.train(steps, min(batch, self.rows), lr, seed)


@dataclass
class TrainingData:
    X: np.ndarray      # [rows][64] raw inputs
    T: np.ndarray      # [rows][6] targets: drift and veval corrections
    mask: np.ndarray   # [rows] 1 for rows to learn from
    segments: int = 0  # fine trajectory segments the rows come from


def features(sim) -> np.ndarray:
    """Raw inputs of every slot of the sorted state (runs the pressure and a dry force pass)."""
    X = np.empty((sim.n, IN), np.float32)
    _declare(sim.lib)
    sim._call("net_features", sim.h, _fp(X))
    return X


def make_training_data(scenes, samples: int = 50, ratio: int = 4, gap: int = 80, warm: int = 0,
                       dt: float | None = None, rows_per_sample: int | None = None, seed: int = 0,
                       lib=None) -> TrainingData:
    """For each scene (one scene or a list), `samples` start states along its fine trajectory
    (one every `gap` fine steps after `warm`); from each, a fine segment of `ratio` steps of dt
    gives the target of every active particle (matched by particle ID). `rows_per_sample`
    keeps a random subset of each segment's rows."""
    from .fluid_system import FluidSystem

    if not isinstance(scenes, (list, tuple)):
        scenes = [scenes]
    rng = np.random.default_rng(seed)
    Xs, Ts, Ms = [], [], []
    segs = 0
    for scene in scenes:
        h = scene.dt if dt is None else dt
        flags = scene.flags
        A = FluidSystem(scene.n, scene.b, lib=lib)
        A.load_scene(scene)
        A.Run2PhysicalSort()
        if warm:
            A.step(warm, h, flags)
        for _ in range(samples):
            if gap:
                A.step(gap, h, flags)
            X = features(A)
            nump, na = A.counts()
            ids0 = A.download(F.FPARTICLE_ID)[:na].copy()
            v0 = A.download(F.FVEL)[:na, :3].astype(np.float64)
            a0 = X[:na, 50:53].astype(np.float64)
            vsum = np.zeros((na, 3))
            ok = np.ones(na, bool)
            for _k in range(ratio):
                A.step(1, h, flags)
                vk, found = _by_id(A, F.FVEL, ids0)
                vsum += vk
                ok &= found
            ve4, found = _by_id(A, F.FVEVAL, ids0)
            ok &= found
            vnext = v0 + ratio * h * a0
            T = np.concatenate([vsum / ratio - vnext, ve4 - 0.5 * (v0 + vnext)], axis=1)
            T = T.astype(np.float32)
            rows = np.nonzero(ok)[0]
            if rows_per_sample is not None and rows.size > rows_per_sample:
                rows = np.sort(rng.choice(rows, rows_per_sample, replace=False))
            Xs.append(X[:na][rows])
            Ts.append(T[rows])
            Ms.append(np.ones(rows.size, np.float32))
            segs += 1
        A.close()
    return TrainingData(np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Ms), segs)


def _by_id(sim, buf, query_ids):
    """xyz of a float4 buffer for the active particles with the given IDs, and which were found."""
    nump, na = sim.counts()
    ids = sim.download(F.FPARTICLE_ID)[:na]
    v = sim.download(buf)[:na, :3].astype(np.float64)
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
