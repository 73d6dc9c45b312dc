"""Scalar physics of the batch BD path (oracle, test infrastructure only).

Restates, in plain NumPy / Python floats:

* Romberg quadrature            - pystarc/lib/numerical.py:74-97
* k_b and return probability    - pystarc/simulation/gpu_batch_simulator.py:320-404
* Wilson CI and k_on            - gpu_batch_simulator.py:61-113
* Shoemake random quaternion    - gpu_batch_simulator.py:1597-1609
* quaternion -> rotation matrix - gpu_batch_simulator.py:1611-1629
* rotation-vector update        - gpu_batch_simulator.py:1631-1660
* Stokes D_t / D_r              - pystarc/hydrodynamics/rotne_prager.py:43-64
* RPY off-diagonal tt block     - rotne_prager.py:67-124
* relative diffusion D_rel(r)   - rotne_prager.py:183-204
* Ermak-McCammon translation    - gpu_batch_simulator.py:1085-1094
* overlap rejection test        - gpu_batch_simulator.py:1102-1128
* Brownian-bridge D_eff and crossing probability
                                - gpu_batch_simulator.py:637-671, :1161-1181
"""
from __future__ import annotations

import math

import numpy as np

EPS0 = 0.000142          # e^2/(kBT*A), pystarc/global_defs/constants.py:47-49
CONV = 6.022e23 * 1e-30 / 1e-12 / 1e-3   # A^3/ps -> M^-1 s^-1, gpu_batch_simulator.py:87
KB_SI = 1.380649e-23
T_DEFAULT = 298.15
ETA_WATER = 1.002e-3
ANG_TO_M = 1.0e-10
PS_TO_S = 1.0e-12


def romberg_integrate(f, a, b, tol=1e-8, max_order=12):
    """Romberg table with absolute-difference stop (numerical.py:74-97)."""
    R = [[0.0] * (max_order + 1) for _ in range(max_order + 1)]
    h = b - a
    R[0][0] = 0.5 * h * (f(a) + f(b))
    for i in range(1, max_order + 1):
        h /= 2.0
        n_new = 2 ** i
        s = sum(f(a + (2 * k - 1) * h) for k in range(1, n_new // 2 + 1))
        R[i][0] = 0.5 * R[i - 1][0] + h * s
        for j in range(1, i + 1):
            R[i][j] = R[i][j - 1] + (R[i][j - 1] - R[i - 1][j - 1]) / (4 ** j - 1)
        if i >= 2 and abs(R[i][i] - R[i - 1][i - 1]) < tol:
            return R[i][i]
    return R[max_order][max_order]


def return_prob_and_kb(b, q_out, q_rec, q_lig, debye, D_t, use_hi=False,
                       a0=0.0, a1=0.0, kT_scale=1.0):
    """(return_prob, k_b) as gpu_batch_simulator.py:320-404 computes them.

    U(r) = q_r q_l e^{-r/debye} / (4 pi 78 eps0 r) (sdie hard-coded 78,
    :338-351); k(x) = 4 pi / int_0^{1/x} e^{U(1/s)}/D ds; return_prob =
    k(b)/k(q_out) clamped to [0, 1].  Free diffusion when |q_r q_l| < 1e-9.
    """
    eps = 78.0 * EPS0
    if abs(q_rec * q_lig) < 1e-9:
        return b / q_out, 4.0 * math.pi * D_t * b

    def U(r):
        return q_rec * q_lig / (4.0 * math.pi * eps * r) * math.exp(-r / debye)

    if use_hi:
        a2 = 0.5 * (a0 ** 2 + a1 ** 2)
        ainv = 1.0 / a0 + 1.0 / a1
        dpre = D_t / ainv

    def intgd(s):
        if s == 0.0:
            return 1.0 / D_t
        r = 1.0 / s
        v = U(r) / kT_scale
        if use_hi:
            return math.exp(v) / (dpre * (ainv - 3.0 / r + 2.0 * a2 / r ** 3))
        return math.exp(v) / D_t

    k_b = 4.0 * math.pi / romberg_integrate(intgd, 0.0, 1.0 / b, tol=1e-8)
    k_q = 4.0 * math.pi / romberg_integrate(intgd, 0.0, 1.0 / q_out, tol=1e-8)
    rp = k_b / k_q if k_q > 0 else 0.5
    return float(max(0.0, min(1.0, rp))), k_b


def norm_ppf(p):
    """Inverse standard-normal CDF by bisection on erf (z(0.975)=1.959964)."""
    lo, hi = -10.0, 10.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if 0.5 * (1.0 + math.erf(mid / math.sqrt(2.0))) < p:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def wilson_ci(n_reacted, n_escaped, confidence=0.95):
    """Wilson score interval, gpu_batch_simulator.py:70-80 (z^2/(4n^2) form)."""
    n = n_reacted + n_escaped
    if n == 0:
        return (0.0, 1.0)
    z = norm_ppf(0.5 + confidence / 2.0)
    p = n_reacted / n
    z2n = z ** 2 / n
    denom = 1.0 + z2n
    centre = (p + z2n / 2.0) / denom
    margin = z * math.sqrt(max(p * (1 - p) / n + z2n / (4.0 * n), 0)) / denom
    return (max(0.0, centre - margin), min(1.0, centre + margin))


def rate_constant(P, k_b, D_rel=0.0, r_start=0.0, r_escape=1.0):
    """k_on = CONV k_b P, or the Smoluchowski fallback (:82-96)."""
    if P == 0.0:
        return 0.0
    if k_b > 0.0:
        return CONV * k_b * P
    k_D = 4.0 * math.pi * D_rel * r_start
    beta = r_start / r_escape
    return CONV * k_D * P / (1.0 - P * (1.0 - beta))


def shoemake(u):
    """(N,3) uniforms -> (N,4) unit quaternions (w,x,y,z), :1597-1609."""
    u = np.asarray(u, dtype=np.float64)
    return np.stack([
        np.sqrt(1 - u[:, 0]) * np.sin(2 * np.pi * u[:, 1]),
        np.sqrt(1 - u[:, 0]) * np.cos(2 * np.pi * u[:, 1]),
        np.sqrt(u[:, 0]) * np.sin(2 * np.pi * u[:, 2]),
        np.sqrt(u[:, 0]) * np.cos(2 * np.pi * u[:, 2]),
    ], axis=1)


def quats_to_rotmats(q):
    """(N,4) (w,x,y,z) -> (N,3,3), :1611-1629."""
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.stack([
        1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
        2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
        2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y),
    ], axis=1).reshape(-1, 3, 3)
    return R


def apply_rotation(q, omega):
    """q <- normalize(q (x) dq(omega)), right multiplication, :1631-1660."""
    angle = np.linalg.norm(omega, axis=1, keepdims=True)
    safe = angle > 1e-10
    axis = np.where(safe, omega / np.where(safe, angle, 1.0), 0.0)
    half = angle * 0.5
    s = np.sin(half[:, 0])
    dq = np.stack([np.cos(half[:, 0]), axis[:, 0] * s, axis[:, 1] * s, axis[:, 2] * s], axis=1)
    w1, x1, y1, z1 = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    w2, x2, y2, z2 = dq[:, 0], dq[:, 1], dq[:, 2], dq[:, 3]
    qn = np.stack([
        w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
        w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
        w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
        w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2,
    ], axis=1)
    return qn / np.linalg.norm(qn, axis=1, keepdims=True)


def stokes_dt(radius, eta=ETA_WATER, T=T_DEFAULT):
    """D_t = kT/(6 pi eta a) in A^2/ps, rotne_prager.py:43-52."""
    D = KB_SI * T / (6.0 * math.pi * eta * radius * ANG_TO_M)
    return D / ANG_TO_M ** 2 * PS_TO_S


def stokes_dr(radius, eta=ETA_WATER, T=T_DEFAULT):
    """D_r = kT/(8 pi eta a^3) in rad^2/ps, rotne_prager.py:55-64."""
    return KB_SI * T / (8.0 * math.pi * eta * (radius * ANG_TO_M) ** 3) * PS_TO_S


def rpy_offdiagonal(r_vec, a, b, D_a, D_b):
    """Three-regime tt coupling block M_12 (rotne_prager.py:67-124)."""
    r = float(np.linalg.norm(r_vec))
    if r < 1e-10:
        return np.zeros((3, 3))
    rhat = r_vec / r
    outer = np.outer(rhat, rhat)
    a2, b2, r2 = a * a, b * b, r * r
    r3 = r2 * r
    if r > a + b:
        den = 8.0 * math.pi * r
        tt_I = (1.0 + (a2 + b2) / (3.0 * r2)) / den
        tt_uu = (1.0 - (a2 + b2) / r2) / den
    elif r > abs(a - b):
        am2 = (a - b) ** 2
        den = 6.0 * 32.0 * math.pi * a * b * r3
        tt_I = (16.0 * r3 * (a + b) - (am2 + 3.0 * r2) ** 2) / den
        tt_uu = 3.0 * (am2 - r2) ** 2 / den
    else:
        tt_I = 1.0 / (6.0 * math.pi * max(a, b))
        tt_uu = 0.0
    kT_over_eta = math.sqrt(D_a * 6.0 * math.pi * a * D_b * 6.0 * math.pi * b)
    return kT_over_eta * (tt_I * np.eye(3) + tt_uu * outer)


def relative_translational_diffusion(D1, D2, a1, a2, r_vec=None):
    """D_t1 + D_t2 - (2/3) tr M_12(r), floored at 1e-12 (rotne_prager.py:183-204)."""
    D0 = D1 + D2
    if r_vec is None or a1 <= 0 or a2 <= 0:
        return D0
    return max(D0 - (2.0 / 3.0) * np.trace(rpy_offdiagonal(r_vec, a1, a2, D1, D2)), 1e-12)


def em_translate(pos, D, dt, F, xi):
    """Ermak-McCammon move pos + D dt F + sqrt(2 D dt) xi
    (gpu_batch_simulator.py:1085-1094); D, dt per row."""
    D = np.asarray(D, dtype=np.float64)[:, None]
    dt = np.asarray(dt, dtype=np.float64)[:, None]
    return pos + D * dt * F + np.sqrt(2.0 * D * dt) * xi


def overlap_rejected(centroids, rec_pos, thresh=1.5):
    """True where the ligand centroid is closer than `thresh` to any receptor
    atom (gpu_batch_simulator.py:1102-1128, threshold :774)."""
    dmin = np.full(len(centroids), np.inf)
    for c0 in range(0, rec_pos.shape[0], 4096):
        dd = np.linalg.norm(centroids[:, None, :] - rec_pos[None, c0:c0 + 4096, :], axis=2)
        dmin = np.minimum(dmin, dd.min(axis=1))
    return dmin < thresh


def bridge_d_eff(D_t, a_rec, a_lig, L_rec, L_lig):
    """Diffusion coefficient of a reaction-pair distance
    (gpu_batch_simulator.py:643-660): D_t + D_rot,rec L_rec^2 + D_rot,lig
    L_lig^2, D_t split between the bodies in inverse proportion to their radii
    and D_rot,k = D_trans,k * 3 / (4 a_k^2); L = lever of the pair atom from
    its body's centre."""
    a_sum = a_rec + a_lig if (a_rec + a_lig) > 0 else 1.0
    Dtr = D_t * a_lig / a_sum
    Dtl = D_t * a_rec / a_sum
    Drr = Dtr * 3.0 / (4.0 * a_rec ** 2) if a_rec > 0 else 0.0
    Drl = Dtl * 3.0 / (4.0 * a_lig ** 2) if a_lig > 0 else 0.0
    return D_t + Drr * np.asarray(L_rec) ** 2 + Drl * np.asarray(L_lig) ** 2


def bridge_crossing_prob(x0, x1, D_dt):
    """Probability that a Brownian path from x0 to x1 over a step with
    diffusion D dt touched 0 (gpu_batch_simulator.py:1161-1181): exp(-x0 x1 /
    max(D dt, 1e-30)) where x0 > 0 and x1 > 0, else 0."""
    D_dt = np.maximum(D_dt, 1e-30)
    both = (x0 > 0) & (x1 > 0)
    with np.errstate(over="ignore"):
        return np.where(both, np.exp(-x0 * x1 / D_dt), 0.0)
