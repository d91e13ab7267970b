"""FP64 reference of the learned-step MLP (forward, loss, backward) in numpy.

TEST INFRASTRUCTURE: the checker for the BF16-MFMA kernels of morphogenesis_amd/csrc/mg_mlp.hip
and for the parameter layout of mg_net (include/mg_sph.h). Pinned on its own by central finite
differences (tests/test_oracle_pins.py).

Layout of the parameter vector (fp32 on the device, row-major [in][out]):
  W1 [64][128] | b1 [128] | W2 [128][128] | b2 [128] | W3 [128][16] | b3 [16]
"""
from __future__ import annotations

import numpy as np

IN, H, OUT, OUT_REAL = 64, 128, 16, 6
SHAPES = [("W1", (IN, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, OUT)), ("b3", (OUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + k], np.float64).reshape(shp)
        o += k
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p: np.ndarray, xn: np.ndarray):
    """xn: normalised inputs [rows][64]. Returns y [rows][16] and the activations."""
    w = unpack(p)
    z1 = xn @ w["W1"] + w["b1"]
    a1 = silu(z1)
    z2 = a1 @ w["W2"] + w["b2"]
    a2 = silu(z2)
    y = a2 @ w["W3"] + w["b3"]
    return y, (xn, z1, a1, z2, a2)


def loss_and_grad(p: np.ndarray, xn: np.ndarray, t: np.ndarray, mask: np.ndarray):
    """Mean over masked rows of |y[:, :3] - t|^2 and its gradient w.r.t. p."""
    w = unpack(p)
    y, (x, z1, a1, z2, a2) = forward(p, xn)
    m = (np.asarray(mask) != 0).astype(np.float64)
    cnt = max(m.sum(), 1.0)
    e = np.zeros_like(y)
    e[:, :OUT_REAL] = (y[:, :OUT_REAL] - t) * m[:, None]
    loss = float((e ** 2).sum() / cnt)
    dy = 2.0 * e / cnt
    g = {"W3": a2.T @ dy, "b3": dy.sum(0)}
    dz2 = (dy @ w["W3"].T) * silu_grad(z2)
    g["W2"] = a1.T @ dz2
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ w["W2"].T) * silu_grad(z1)
    g["W1"] = x.T @ dz1
    g["b1"] = dz1.sum(0)
    return loss, pack(g)


def normalise(x: np.ndarray, mu: np.ndarray, inv_sigma: np.ndarray) -> np.ndarray:
    return (np.asarray(x, np.float64) - mu) * inv_sigma


def bf16_round(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float64."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(p: np.ndarray, xn: np.ndarray) -> np.ndarray:
    """The forward pass with the roundings of the BF16 kernels: weights, inputs and hidden
    activations rounded to bfloat16, sums and biases in float64."""
    w = unpack(p)
    x = bf16_round(xn)
    a1 = bf16_round(silu(x @ bf16_round(w["W1"]) + w["b1"]))
    a2 = bf16_round(silu(a1 @ bf16_round(w["W2"]) + w["b2"]))
    return a2 @ bf16_round(w["W3"]) + w["b3"]
