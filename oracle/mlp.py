"""FP64 forward / backward of the 9 -> 4x256 (SiLU) -> 13 surrogate MLP
(oracle, test infrastructure only; no reference counterpart).

``bf16=True`` rounds at the same points as the gfx950 kernels (the weights
of layers 1-4, every hidden activation and the backward gradients to BF16,
round to nearest even; the first layer's inputs and weights stay unrounded,
as the kernels run that layer in FP32; its weight gradient uses BF16 inputs,
as the kernels' backward does) and accumulates in FP64, so the kernels can
be compared to it up to accumulation order; ``bf16=False`` is the exact FP64
network, pinned in the tests by torch autograd (float64, CPU) and by finite
differences.
"""
from __future__ import annotations

import numpy as np

N_IN, WIDTH, N_OUT = 9, 256, 13
SHAPES = [(WIDTH, N_IN), (WIDTH, WIDTH), (WIDTH, WIDTH), (WIDTH, WIDTH), (N_OUT, WIDTH)]


def n_params():
    return sum(o * i + o for o, i in SHAPES)


def unflatten(p):
    out, k = [], 0
    for o, i in SHAPES:
        W = p[k:k + o * i].reshape(o, i)
        k += o * i
        b = p[k:k + o]
        k += o
        out.append((W, b))
    return out


def bf16(a):
    """Round float64/float32 values to BF16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * _sig(z)


def dsilu(z):
    s = _sig(z)
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, bf=True):
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    fb = (lambda b: b.astype(np.float32).astype(np.float64)) if bf else (lambda b: b)
    layers = unflatten(np.asarray(p, dtype=np.float64))
    h = np.asarray(x, dtype=np.float64)
    if bf:
        h = h.astype(np.float32).astype(np.float64)
    # the first layer's weight gradient reads BF16 inputs
    cache = {"h": [r(x)], "z": []}
    for l, (W, b) in enumerate(layers[:4]):
        Wl = (W.astype(np.float32).astype(np.float64) if bf else W) if l == 0 else r(W)
        z = h @ Wl.T + fb(b)
        h = r(silu(z))
        cache["z"].append(z)
        cache["h"].append(h)
    W, b = layers[4]
    y = h @ r(W).T + fb(b)
    return y, cache


def backward(p, cache, gy, bf=True):
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unflatten(np.asarray(p, dtype=np.float64))
    grads = [None] * 5
    g = r(gy)
    for l in range(4, -1, -1):
        W, _ = layers[l]
        h_in = cache["h"][l]
        grads[l] = (g.T @ h_in, g.sum(axis=0))
        if l > 0:
            gh = g @ r(W)
            g = r(gh * dsilu(cache["z"][l - 1]))
    flat = []
    for gW, gb in grads:
        flat += [gW.ravel(), gb]
    return np.concatenate(flat)
