"""NumPy evaluation of the synthetic grids (oracle, test infrastructure only).

Same definitions as the device generator in pystarc_amd/csrc/pst_grid.hip:
kind 0 - Debye-Hueckel potential sum_j q_j e^{-(d-a_j)/lambda} /
((1 + a_j/lambda) 4 pi eps0 eps_out d) with d = max(|x - r_j|, a_j, 0.5);
kind 1 - Born field (eps_out - eps_in)/(eps_out + eps_in)/(8 pi eps0 eps_out)
* sum_j a_j^3 / max(d, a_j, 0.5)^4.
Pinned in tests by the closed form of a single charge (kind 0 with a = 0 is the
reference's Yukawa monopole, pystarc/forces/gpu_batch_engine.py:407-410).
"""
from __future__ import annotations

import math

import numpy as np

from .physics import EPS0


def grid_points(shape, origin, h):
    ax = [origin[i] + h * np.arange(shape[i]) for i in range(3)]
    return np.stack(np.meshgrid(*ax, indexing="ij"), axis=-1)


def make_grid(kind, shape, origin, h, atoms, debye=7.86, eps_out=78.0, eps_in=4.0):
    X = grid_points(shape, origin, h).reshape(-1, 3)
    acc = np.zeros(len(X))
    pref0 = 1.0 / (4.0 * math.pi * EPS0 * eps_out)
    pref1 = (eps_out - eps_in) / (eps_out + eps_in) / (8.0 * math.pi * EPS0 * eps_out)
    for x, y, z, q, a in np.asarray(atoms, dtype=np.float64):
        d = np.sqrt((X[:, 0] - x) ** 2 + (X[:, 1] - y) ** 2 + (X[:, 2] - z) ** 2)
        d = np.maximum(d, max(a, 0.5))
        if kind == 0:
            acc += q * np.exp(-(d - a) / debye) / ((1.0 + a / debye) * d) * pref0
        else:
            acc += pref1 * a ** 3 / d ** 4
    return acc.reshape(shape)
