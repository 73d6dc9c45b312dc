"""Counter-based noise of the production step, restated in NumPy (oracle,
test infrastructure only).

Philox4x32-10 (Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as
easy as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies
(0xD2511F53, 0xCD9E8D57) with Weyl key increments (0x9E3779B9, 0xBB67AE85).
Counter = (trajectory lo, trajectory hi, trajectory step, stream | sub << 8),
key = (seed lo, seed hi).  Uniforms take 53 bits, (hi << 21) ^ (lo >> 11),
scaled by 2^-53; normals are Box-Muller pairs
(sqrt(-2 ln(1 - u0)) cos / sin (2 pi u1)).

Pinned by the Random123 known-answer vectors for philox4x32-10
(tests/test_oracle_production.py).
"""
from __future__ import annotations

import numpy as np

M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = np.uint64(0xFFFFFFFF)

RNG_INIT, RNG_ESC, RNG_RETQ, RNG_NOISE, RNG_BRIDGE = 1, 2, 3, 4, 5


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    c = [np.asarray(x, dtype=np.uint64) & MASK for x in (c0, c1, c2, c3)]
    c0, c1, c2, c3 = np.broadcast_arrays(*c)
    c0, c1, c2, c3 = (x.copy() for x in (c0, c1, c2, c3))
    k0 = np.uint64(int(k0) & 0xFFFFFFFF)
    k1 = np.uint64(int(k1) & 0xFFFFFFFF)
    for r in range(10):
        if r:
            k0 = np.uint64((int(k0) + W0) & 0xFFFFFFFF)
            k1 = np.uint64((int(k1) + W1) & 0xFFFFFFFF)
        p0 = np.uint64(M0) * c0
        p1 = np.uint64(M1) * c2
        n0 = (p1 >> np.uint64(32)) ^ c1 ^ k0
        n1 = p1 & MASK
        n2 = (p0 >> np.uint64(32)) ^ c3 ^ k1
        n3 = p0 & MASK
        c0, c1, c2, c3 = n0, n1, n2, n3
    return c0, c1, c2, c3


def uniform2(seed, traj, step, stream, sub):
    traj = np.asarray(traj, dtype=np.int64).astype(np.uint64)
    step = np.asarray(step, dtype=np.uint64)
    ctr3 = np.uint64(int(stream) | (int(sub) << 8))
    o0, o1, o2, o3 = philox4x32_10(traj & MASK, traj >> np.uint64(32), step, ctr3,
                                   int(seed) & 0xFFFFFFFF, (int(seed) >> 32) & 0xFFFFFFFF)
    m53 = np.uint64((1 << 53) - 1)
    a = ((o0 << np.uint64(21)) ^ (o1 >> np.uint64(11))) & m53
    b = ((o2 << np.uint64(21)) ^ (o3 >> np.uint64(11))) & m53
    s = 1.0 / 9007199254740992.0
    return a.astype(np.float64) * s, b.astype(np.float64) * s


def normal2(seed, traj, step, stream, sub):
    u0, u1 = uniform2(seed, traj, step, stream, sub)
    r = np.sqrt(-2.0 * np.log(1.0 - u0))
    ang = 2.0 * np.pi * u1
    return r * np.cos(ang), r * np.sin(ang)
