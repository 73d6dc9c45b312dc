"""Two-body Rotne-Prager-Yamakawa grand mobility and its Cholesky noise
(oracle, test infrastructure only).

The reference keeps only the translational coupling block of two spheres
(pystarc/hydrodynamics/rotne_prager.py:101-123, the three regimes of Zuk,
Wajnryb, Mizerski & Szymczak, J. Fluid Mech. 741 R5 (2014)) and uses it
through the scalar D_parallel of gpu_batch_simulator.py:1074-1084.  The
twelve-by-twelve mobility here adds the self blocks (1/(6 pi a),
1/(8 pi a^3)), the rotation-rotation and rotation-translation couplings of
the same paper, in units of 1/eta, ordered

    X = [x_rec (3), x_lig (3), theta_rec (3), theta_lig (3)]

with the receptor at the origin and the ligand at r.  Each block is built
from its tensor form (I, u u, eps.u with (eps.u)_ab = eps_abc u_c).

``aligned_factor`` is the factor the device uses: in the frame whose third
axis is u = r/|r| the matrix splits into a 4x4 block (t0x, t1x, r0y, r1y),
its sign-flipped twin (t0y, t1y, r0x, r1x) and two 2x2 blocks (t0z, t1z),
(r0z, r1z); each is Cholesky-factored and rotated back to the lab frame.
The product G G^T equals ``grand_mobility`` (tests check it), so G xi has the
covariance of the mobility.

Pins (tests/test_oracle_rpy12.py): continuity of every coefficient at
r = a0 + a1 and r = |a0 - a1|; the equal-radius overlap forms of Rotne &
Prager (1969) / Wajnryb et al. (2013), typed from the textbook, not from the
general formulas; the translational partial-overlap block against a Monte
Carlo surface average of the Oseen tensor; the reference's own D_parallel
(far regime); positive definiteness over random configurations; the
Stokes self-mobility limit far apart; numpy.linalg.cholesky.
"""
from __future__ import annotations

import math

import numpy as np


def cross_coeffs(r, a0, a1):
    """Scalar coefficients of the cross blocks (units of 1/eta).

    Returns (al, be, ga, de, x01, x10):
      tt_01 = al I + be u u
      rr_01 = ga I + de u u
      rt_ij = x_ij eps.u_ij, rotation of sphere i by a force on sphere j,
              u_ij = (r_i - r_j)/|r_i - r_j|.
    """
    s = a0 + a1
    d = abs(a0 - a1)
    pi = math.pi
    if r > s:
        sq = a0 * a0 + a1 * a1
        al = (1.0 + sq / (3.0 * r * r)) / (8.0 * pi * r)
        be = (1.0 - sq / (r * r)) / (8.0 * pi * r)
        ga = -1.0 / (16.0 * pi * r ** 3)
        de = 3.0 / (16.0 * pi * r ** 3)
        x01 = x10 = 1.0 / (8.0 * pi * r * r)
    elif r > d:
        dd = (a0 - a1) ** 2
        al = (16.0 * r ** 3 * s - (dd + 3.0 * r * r) ** 2) / (32.0 * r ** 3) / (6.0 * pi * a0 * a1)
        be = 3.0 * (dd - r * r) ** 2 / (32.0 * r ** 3) / (6.0 * pi * a0 * a1)
        A = (5.0 * r ** 6 - 27.0 * r ** 4 * (a0 * a0 + a1 * a1) + 32.0 * r ** 3 * (a0 ** 3 + a1 ** 3)
             - 9.0 * r * r * (a0 * a0 - a1 * a1) ** 2
             - (a0 - a1) ** 4 * (a0 * a0 + 4.0 * a0 * a1 + a1 * a1)) / (64.0 * r ** 3)
        B = 3.0 * (dd - r * r) ** 2 * (a0 * a0 + 4.0 * a0 * a1 + a1 * a1 - r * r) / (64.0 * r ** 3)
        ga = A / (8.0 * pi * a0 ** 3 * a1 ** 3)
        de = B / (8.0 * pi * a0 ** 3 * a1 ** 3)

        def xi(ai, aj):
            return ((ai - aj + r) ** 2 * (aj * aj + 2.0 * aj * (ai + r) - 3.0 * (ai - r) ** 2)
                    / (128.0 * pi * ai ** 3 * aj * r * r))
        x01, x10 = xi(a0, a1), xi(a1, a0)
    else:
        big = max(a0, a1)
        al, be = 1.0 / (6.0 * pi * big), 0.0
        ga, de = 1.0 / (8.0 * pi * big ** 3), 0.0
        x01 = r / (8.0 * pi * a0 ** 3) if a0 > a1 else 0.0
        x10 = r / (8.0 * pi * a1 ** 3) if a1 > a0 else 0.0
    return al, be, ga, de, x01, x10


def eps_dot(u):
    """(eps.u)_ab = eps_abc u_c, so (eps.u) v = v x u."""
    x, y, z = u
    return np.array([[0.0, z, -y], [-z, 0.0, x], [y, -x, 0.0]])


def grand_mobility(rvec, a0, a1):
    """12x12 grand mobility (1/eta units), lab frame, receptor at the origin."""
    rvec = np.asarray(rvec, dtype=np.float64)
    r = float(np.linalg.norm(rvec))
    u = rvec / r
    I = np.eye(3)
    uu = np.outer(u, u)
    E = eps_dot(u)  # u_10 = u, u_01 = -u
    al, be, ga, de, x01, x10 = cross_coeffs(r, a0, a1)
    M = np.zeros((12, 12))
    M[0:3, 0:3] = I / (6.0 * math.pi * a0)
    M[3:6, 3:6] = I / (6.0 * math.pi * a1)
    M[0:3, 3:6] = M[3:6, 0:3] = al * I + be * uu
    M[6:9, 6:9] = I / (8.0 * math.pi * a0 ** 3)
    M[9:12, 9:12] = I / (8.0 * math.pi * a1 ** 3)
    M[6:9, 9:12] = M[9:12, 6:9] = ga * I + de * uu
    rt01 = x01 * (-E)   # rotation of the receptor by a force on the ligand
    rt10 = x10 * E      # rotation of the ligand by a force on the receptor
    M[6:9, 3:6] = rt01
    M[9:12, 0:3] = rt10
    M[3:6, 6:9] = rt01.T
    M[0:3, 9:12] = rt10.T
    return M


def basis(u):
    """Orthonormal (e1, e2, u), e1 x e2 = u (Duff et al., JCGT 6(1), 2017)."""
    x, y, z = u
    sgn = math.copysign(1.0, z)
    a = -1.0 / (sgn + z)
    b = x * y * a
    e1 = np.array([1.0 + sgn * x * x * a, sgn * b, -sgn * x])
    e2 = np.array([b, sgn + y * y * a, -y])
    return e1, e2


def group_blocks(r, a0, a1):
    """Aligned-frame blocks: A (4x4), C (t_z 2x2), D (r_z 2x2)."""
    al, be, ga, de, x01, x10 = cross_coeffs(r, a0, a1)
    m0, m1 = 1.0 / (6.0 * math.pi * a0), 1.0 / (6.0 * math.pi * a1)
    p0, p1 = 1.0 / (8.0 * math.pi * a0 ** 3), 1.0 / (8.0 * math.pi * a1 ** 3)
    A = np.array([[m0, al, 0.0, -x10],
                  [al, m1, x01, 0.0],
                  [0.0, x01, p0, ga],
                  [-x10, 0.0, ga, p1]])
    Cm = np.array([[m0, al + be], [al + be, m1]])
    Dm = np.array([[p0, ga + de], [ga + de, p1]])
    return A, Cm, Dm


def aligned_factor(rvec, a0, a1):
    """Lab-frame factor G (12x12), G G^T = grand_mobility; columns = noise index.

    Noise index layout: 0-3 block A, 4-7 block B = S A S (S = diag(1,1,-1,-1)),
    8-9 block (t0z, t1z), 10-11 block (r0z, r1z).
    """
    rvec = np.asarray(rvec, dtype=np.float64)
    r = float(np.linalg.norm(rvec))
    u = rvec / r
    e1, e2 = basis(u)
    A, Cm, Dm = group_blocks(r, a0, a1)
    LA = np.linalg.cholesky(A)
    S = np.diag([1.0, 1.0, -1.0, -1.0])
    LB = S @ LA @ S
    LC = np.linalg.cholesky(Cm)
    LD = np.linalg.cholesky(Dm)
    # aligned components (t0, t1, r0, r1) x (x, y, z) as rows of a 12-vector
    # in the order [t0x t0y t0z t1x t1y t1z r0x r0y r0z r1x r1y r1z]
    Ga = np.zeros((12, 12))
    rowsA = [0, 3, 7, 10]    # t0x t1x r0y r1y
    rowsB = [1, 4, 6, 9]     # t0y t1y r0x r1x
    rowsC = [2, 5]           # t0z t1z
    rowsD = [8, 11]          # r0z r1z
    Ga[np.ix_(rowsA, range(0, 4))] = LA
    Ga[np.ix_(rowsB, range(4, 8))] = LB
    Ga[np.ix_(rowsC, range(8, 10))] = LC
    Ga[np.ix_(rowsD, range(10, 12))] = LD
    Q = np.stack([e1, e2, u], axis=1)   # aligned -> lab
    Q4 = np.kron(np.eye(4), Q)
    return Q4 @ Ga


def rotvec_quat(w):
    ang = float(np.linalg.norm(w))
    if ang > 1e-10:
        ax = w / ang
    else:
        ax = np.zeros(3)
    h = 0.5 * ang
    return np.array([math.cos(h), *(ax * math.sin(h))])


def qmul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
                     w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                     w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def qrot(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def displacement(pos, F_lig, T_lig, F_rec, T_rec, xi, dt, Df, a0, a1):
    """dX = Df dt M f + sqrt(2 Df dt) G xi, f = [F_rec, F_lig, T_rec, T_lig]."""
    M = grand_mobility(pos, a0, a1)
    G = aligned_factor(pos, a0, a1)
    f = np.concatenate([F_rec, F_lig, T_rec, T_lig])
    return Df * dt * (M @ f) + math.sqrt(2.0 * Df * dt) * (G @ np.asarray(xi, dtype=np.float64))


def move(pos, q, dX):
    """New pose in the receptor frame after both bodies moved by dX.

    Lab motion: ligand centre += dx_lig, receptor centre += dx_rec, the ligand
    turns by theta_lig and the receptor by theta_rec (both lab-frame rotation
    vectors).  Re-expressed in the turned receptor frame:
      pos' = R(theta_rec)^T (pos + dx_lig - dx_rec)
      q'   = conj(dq(theta_rec)) * dq(theta_lig) * q   (normalised)
    """
    dq_r = rotvec_quat(dX[6:9])
    dq_l = rotvec_quat(dX[9:12])
    p = np.asarray(pos, dtype=np.float64) + dX[3:6] - dX[0:3]
    p_new = qrot(dq_r).T @ p
    conj = dq_r * np.array([1.0, -1.0, -1.0, -1.0])
    qn = qmul(conj, qmul(dq_l, np.asarray(q, dtype=np.float64)))
    return p_new, qn / np.linalg.norm(qn)


def receptor_load(pos, F_lig, T_lig, torque_mode):
    """Force and torque on the receptor from the pair interaction.

    Newton's third law for the force.  Torque: zero in the reference's
    zero-torque mode (gpu_batch_engine.py:597-599); otherwise the balance of
    angular momentum about the origin, T_rec = -T_lig - pos x F_lig.
    """
    F_rec = -np.asarray(F_lig, dtype=np.float64)
    if torque_mode == 0:
        return F_rec, np.zeros(3)
    return F_rec, -np.asarray(T_lig) - np.cross(pos, F_lig)


# ---------------------------------------------------------------------------
# Batched forms (same definitions over n poses; tests check them against the
# single-pose functions above).
# ---------------------------------------------------------------------------
def cross_coeffs_v(r, a0, a1):
    r = np.asarray(r, dtype=np.float64)
    out = np.empty((6, r.shape[0]))
    for i, ri in enumerate(r):
        out[:, i] = cross_coeffs(float(ri), a0, a1)
    return out


def grand_mobility_v(rvec, a0, a1):
    rvec = np.asarray(rvec, dtype=np.float64)
    n = rvec.shape[0]
    r = np.linalg.norm(rvec, axis=1)
    u = rvec / r[:, None]
    al, be, ga, de, x01, x10 = cross_coeffs_v(r, a0, a1)
    I = np.eye(3)[None]
    uu = u[:, :, None] * u[:, None, :]
    E = np.zeros((n, 3, 3))
    E[:, 0, 1], E[:, 0, 2] = u[:, 2], -u[:, 1]
    E[:, 1, 0], E[:, 1, 2] = -u[:, 2], u[:, 0]
    E[:, 2, 0], E[:, 2, 1] = u[:, 1], -u[:, 0]
    M = np.zeros((n, 12, 12))
    M[:, 0:3, 0:3] = I / (6.0 * math.pi * a0)
    M[:, 3:6, 3:6] = I / (6.0 * math.pi * a1)
    tt = al[:, None, None] * I + be[:, None, None] * uu
    M[:, 0:3, 3:6] = tt
    M[:, 3:6, 0:3] = tt
    M[:, 6:9, 6:9] = I / (8.0 * math.pi * a0 ** 3)
    M[:, 9:12, 9:12] = I / (8.0 * math.pi * a1 ** 3)
    rr = ga[:, None, None] * I + de[:, None, None] * uu
    M[:, 6:9, 9:12] = rr
    M[:, 9:12, 6:9] = rr
    rt01 = -x01[:, None, None] * E
    rt10 = x10[:, None, None] * E
    M[:, 6:9, 3:6] = rt01
    M[:, 9:12, 0:3] = rt10
    M[:, 3:6, 6:9] = np.swapaxes(rt01, 1, 2)
    M[:, 0:3, 9:12] = np.swapaxes(rt10, 1, 2)
    return M


def aligned_factor_v(rvec, a0, a1):
    rvec = np.asarray(rvec, dtype=np.float64)
    n = rvec.shape[0]
    r = np.linalg.norm(rvec, axis=1)
    u = rvec / r[:, None]
    al, be, ga, de, x01, x10 = cross_coeffs_v(r, a0, a1)
    m0, m1 = 1.0 / (6.0 * math.pi * a0), 1.0 / (6.0 * math.pi * a1)
    p0, p1 = 1.0 / (8.0 * math.pi * a0 ** 3), 1.0 / (8.0 * math.pi * a1 ** 3)
    A = np.zeros((n, 4, 4))
    A[:, 0, 0], A[:, 1, 1], A[:, 2, 2], A[:, 3, 3] = m0, m1, p0, p1
    A[:, 0, 1] = A[:, 1, 0] = al
    A[:, 0, 3] = A[:, 3, 0] = -x10
    A[:, 1, 2] = A[:, 2, 1] = x01
    A[:, 2, 3] = A[:, 3, 2] = ga
    Cm = np.zeros((n, 2, 2))
    Cm[:, 0, 0], Cm[:, 1, 1] = m0, m1
    Cm[:, 0, 1] = Cm[:, 1, 0] = al + be
    Dm = np.zeros((n, 2, 2))
    Dm[:, 0, 0], Dm[:, 1, 1] = p0, p1
    Dm[:, 0, 1] = Dm[:, 1, 0] = ga + de
    LA = np.linalg.cholesky(A)
    S = np.array([1.0, 1.0, -1.0, -1.0])
    LB = S[None, :, None] * LA * S[None, None, :]
    LC = np.linalg.cholesky(Cm)
    LD = np.linalg.cholesky(Dm)
    Ga = np.zeros((n, 12, 12))
    Ga[:, np.ix_([0, 3, 7, 10], range(0, 4))[0], np.ix_([0, 3, 7, 10], range(0, 4))[1]] = LA
    Ga[:, np.ix_([1, 4, 6, 9], range(4, 8))[0], np.ix_([1, 4, 6, 9], range(4, 8))[1]] = LB
    Ga[:, np.ix_([2, 5], range(8, 10))[0], np.ix_([2, 5], range(8, 10))[1]] = LC
    Ga[:, np.ix_([8, 11], range(10, 12))[0], np.ix_([8, 11], range(10, 12))[1]] = LD
    Q = np.empty((n, 3, 3))
    for i in range(n):
        e1, e2 = basis(u[i])
        Q[i] = np.stack([e1, e2, u[i]], axis=1)
    G = np.empty((n, 12, 12))
    for b in range(4):
        G[:, 3 * b:3 * b + 3, :] = np.einsum("nij,njk->nik", Q, Ga[:, 3 * b:3 * b + 3, :])
    return G


def displacement_v(pos, F_lig, T_lig, F_rec, T_rec, xi, dt, Df, a0, a1):
    M = grand_mobility_v(pos, a0, a1)
    G = aligned_factor_v(pos, a0, a1)
    f = np.concatenate([F_rec, F_lig, T_rec, T_lig], axis=1)
    dt = np.asarray(dt, dtype=np.float64)
    return (Df * dt[:, None] * np.einsum("nij,nj->ni", M, f)
            + np.sqrt(2.0 * Df * dt)[:, None] * np.einsum("nij,nj->ni", G, xi))


def _rotvec_quat_v(w):
    ang = np.linalg.norm(w, axis=1)
    ax = np.where(ang[:, None] > 1e-10, w / np.where(ang > 0, ang, 1.0)[:, None], 0.0)
    h = 0.5 * ang
    return np.concatenate([np.cos(h)[:, None], ax * np.sin(h)[:, None]], axis=1)


def _qmul_v(a, b):
    w1, x1, y1, z1 = a.T
    w2, x2, y2, z2 = b.T
    return np.stack([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
                     w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                     w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2], axis=1)


def move_v(pos, q, dX):
    dq_r = _rotvec_quat_v(dX[:, 6:9])
    dq_l = _rotvec_quat_v(dX[:, 9:12])
    p = pos + dX[:, 3:6] - dX[:, 0:3]
    w, x, y, z = dq_r.T
    R = np.empty((len(p), 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - w * z); R[:, 0, 2] = 2 * (x * z + w * y)
    R[:, 1, 0] = 2 * (x * y + w * z); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - w * x)
    R[:, 2, 0] = 2 * (x * z - w * y); R[:, 2, 1] = 2 * (y * z + w * x); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    p_new = np.einsum("nji,nj->ni", R, p)
    conj = dq_r * np.array([1.0, -1.0, -1.0, -1.0])
    qn = _qmul_v(conj, _qmul_v(dq_l, q))
    return p_new, qn / np.linalg.norm(qn, axis=1, keepdims=True)


def receptor_load_v(pos, F_lig, T_lig, torque_mode):
    F_rec = -np.asarray(F_lig, dtype=np.float64)
    if torque_mode == 0:
        return F_rec, np.zeros_like(F_rec)
    return F_rec, -np.asarray(T_lig) - np.cross(pos, F_lig)
