"""Oracle (test infrastructure only) of the LMZ outer propagator.

Restates, per trajectory and in FP64 NumPy, the analytic outer-region
propagation of PySTARC's CPU simulator:
  * OuterPropagator.new_state   pystarc/simulation/outer_propagator.py:259-401
    (q-sphere at 20 x the largest molecular radius, Lamm-Schulten steps within
    the covers of the b- and q-spheres, drift + diffusion steps between them
    with the step-size rule of :340-377 kept as written, the LMZ return
    probability rate(b)/rate(q) at the q-sphere, :121-125 and :269-283)
  * step_near_absorbing_surface pystarc/simulation/step_near_surface.py:31-130
  * diffusional_rotation        pystarc/simulation/diffusional_rotation.py:87-126
    with the rotation angle for 0.25 < tau < 4 drawn by inverting the exact
    SO(3) heat-kernel CDF (the reference interpolates tabulated CDFs of the
    same distribution at tau = 0.5, 1, 2 and composes).
Random numbers: one Philox4x32-10 call per draw, keyed (trajectory, step,
RNG_OUTER, draw counter), first value of the pair, as the device does.
The trigger 1.1 b is nam_simulator.py:354-358.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erfinv

from . import philox as PH

QB_FACTOR = 1.1        # nam_simulator.py:356
RNG_OUTER = 6
LARGE = 1.0e30


class Draws:
    """Sequential draws of one outer excursion (trajectory key, its step)."""

    def __init__(self, seed, key, step):
        self.seed, self.key, self.step, self.c = int(seed), int(key), int(step), 0

    def u(self):
        a, _ = PH.uniform2(self.seed, self.key, self.step, RNG_OUTER, self.c)
        self.c += 1
        return float(a)

    def n(self):
        a, _ = PH.normal2(self.seed, self.key, self.step, RNG_OUTER, self.c)
        self.c += 1
        return float(a)


_GL = {}


def relative_rate(x, V, debye, D_of_r, D_inf, n=4000):
    """4 pi / int_0^{1/x} exp(V e^{-1/(s L)} s) / D(1/s) ds (outer_propagator.py:160-180),
    Gauss-Legendre on [0, 1/x]."""
    if n not in _GL:
        _GL[n] = np.polynomial.legendre.leggauss(n)
    g, w = _GL[n]
    s = 0.5 * (g + 1.0) / x
    r = 1.0 / s
    f = np.exp(V * np.exp(-r / debye) * s) / D_of_r(r)
    return 4.0 * math.pi / (0.5 / x * float(np.dot(w, f)))


class OuterModel:
    """Parameters of the propagator: b, q, V (q_rec q_lig / (4 pi eps kT), A),
    Debye length, relative radial diffusion D(r) (HI: dpre (ainv - 3/r +
    2 a2/r^3), else D_t), rotational diffusion of both bodies."""

    def __init__(self, b, q, V, debye, D_t, hi, a0, a1, drot0, drot1):
        self.b, self.q, self.V, self.L = float(b), float(q), float(V), float(debye)
        self.D_t, self.hi = float(D_t), bool(hi)
        self.a0, self.a1 = float(a0), float(a1)
        self.ainv = 1.0 / self.a0 + 1.0 / self.a1
        self.a2 = 0.5 * (self.a0 ** 2 + self.a1 ** 2)
        self.dpre = self.D_t / self.ainv
        self.drot0, self.drot1 = float(drot0), float(drot1)
        self.return_prob = self.rate(self.b) / self.rate(self.q)
        self.b_cover = self._cover(True)
        self.q_cover = self._cover(False)

    def D(self, r):
        if self.hi:
            return self.dpre * (self.ainv - 3.0 / r + 2.0 * self.a2 / r ** 3)
        return self.D_t + 0.0 * r

    def rate(self, x):
        return relative_rate(x, self.V, self.L, self.D, self.D_t)

    def force(self, r):
        """Radial force / kT, outward positive (:237-243)."""
        return self.V * math.exp(-r / self.L) / r * (1.0 / r + 1.0 / self.L)

    def _ts_boundary(self, rad, inner):
        """Distance at which the force stops being close to linear (:185-226)."""
        tol, L = 0.05, self.L

        def rel(r):
            return (2 * L * L + 2 * r * L + r * r) * abs(r - rad) / (r * L * (L + r)) - tol
        if inner:
            rb = 2.0 * rad
            while rel(rb) < 0.0:
                rb *= 2.0
            lo, hi = rad, rb
            fhi = rel(hi)
            while hi - lo > 1e-6 * rad:
                m = 0.5 * (lo + hi)
                fm = rel(m)
                if fm * fhi < 0.0:
                    lo = m
                else:
                    hi, fhi = m, fm
            return min(0.5 * (lo + hi), (1.0 + tol) * self.b)
        rb = 0.5 * rad
        while rel(rb) < 0.0:
            rb *= 0.5
        lo, hi = rb, rad
        while hi - lo > 1e-6 * rad:
            m = 0.5 * (lo + hi)
            if rel(m) * rel(lo) < 0.0:
                hi = m
            else:
                lo = m
        return max(0.5 * (lo + hi), (1.0 - tol) * self.q)

    def _cover(self, inner):
        """Where the Lamm-Schulten step takes over (:228-262)."""
        if inner:
            bnd = self._ts_boundary(self.b, True)
            F, L = self.force(self.b), bnd - self.b
        else:
            bnd = self._ts_boundary(self.q, False)
            F, L = -self.force(self.q), self.q - bnd

        def prob(x0):
            if x0 <= 0:
                return 1.0
            den = math.erf((x0 * F + 1.0) / 2.0) + 1.0
            return 1.0 if den == 0 else math.erfc((L - x0 * x0 * F - x0) / (2.0 * x0)) / den
        lo, hi = 0.0, L
        while hi - lo > 1e-6 * L:
            m = 0.5 * (lo + hi)
            if prob(m) > 0.01:
                hi = m
            else:
                lo = m
        x0 = 0.5 * (lo + hi)
        return self.b + x0 if inner else self.q - x0


def step_near(dr: Draws, x0, F, D):
    """(survives, new_x, time) near an absorbing wall (step_near_surface.py:31-130)."""
    bb = -F
    tau = x0 * x0 / D
    st = math.sqrt(tau)
    st2 = 2.0 * st
    bt = bb * tau
    erfmt = math.erf((x0 - bt) / st2)
    erfpt = math.erf((x0 + bt) / st2)
    psurv = min(1.0, max(0.0, 0.5 * (math.exp(bb * x0) * (erfpt - 1.0) + erfmt + 1.0)))
    if dr.u() < psurv:
        E = erfmt
        x, found = 0.0, False
        for _ in range(10000):
            iearg = min(1.0 - 1e-12, max(-1.0 + 1e-12, dr.u() * (E + 1.0) - E))
            x = 2.0 * st * float(erfinv(iearg)) - bt + x0
            if x < 0.0:
                continue
            t4 = 4.0 * tau
            p0 = math.exp(-((x - x0 + bt) ** 2) / t4)
            p1 = math.exp(bb * x0 - ((x + x0 + bt) ** 2) / t4)
            p2 = -2.0 * math.exp(-((x + bt) ** 2 + x0 * (x0 + 2.0 * (x - bt))) / t4)
            p = p0 + p1 + p2
            den = 0.5 * (erfmt + 1.0)
            pu = p0 / den if den > 1e-30 else 0.0
            if p < 0 or pu <= 0:
                continue
            found = dr.u() < p / pu
            if found:
                break
        if not found:
            x = max(x0, 0.001)
        return True, max(0.0, x), tau / D
    x02, b2 = x0 * x0, bb * bb
    if abs(bb * x0) < 0.5:
        tmax = x02 * (1.0 / 6.0 - x02 * b2 / 216.0)
    else:
        tmax = (math.sqrt(b2 * x02 + 9.0) - 3.0) / b2

    def pt(t_):
        if t_ <= 0:
            return 0.0
        return x0 * math.exp(-((x0 - bb * t_) ** 2) / (4.0 * t_)) / (2.0 * math.sqrt(math.pi * t_ ** 3))
    pmax = pt(tmax)
    if pmax <= 0:
        pmax = 1e-30
    ts = 0.0
    for _ in range(10000):
        ts = tau * dr.u()
        if pmax * dr.u() < pt(ts):
            break
    return False, 0.0, ts / D


def qmul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def angle_cdf(theta, tau, lmax=None):
    """CDF of the rotation angle after rotational diffusion for time tau
    (D_r t): the SO(3) heat kernel sum_l (2l+1) e^{-l(l+1) tau} chi_l(theta)
    on the class measure (1 - cos theta) / pi, integrated term by term."""
    if lmax is None:
        lmax = min(64, int(math.ceil(math.sqrt(40.0 / tau))) + 2)
    s = theta - math.sin(theta)
    for l in range(1, lmax + 1):
        s += (2 * l + 1) * math.exp(-l * (l + 1) * tau) * (math.sin(l * theta) / l
                                                          - math.sin((l + 1) * theta) / (l + 1))
    return s / math.pi


def rotation_angle(u, tau):
    lo, hi = 0.0, math.pi
    for _ in range(55):
        m = 0.5 * (lo + hi)
        if angle_cdf(m, tau) < u:
            lo = m
        else:
            hi = m
    return 0.5 * (lo + hi)


def diffusional_rotation(dr: Draws, tau):
    """Unit quaternion of a body diffused for tau = D_r t."""
    if tau <= 0.0:
        return np.array([1.0, 0.0, 0.0, 0.0])
    if tau <= 0.25:
        w = math.sqrt(2.0 * tau) * np.array([dr.n(), dr.n(), dr.n()])
        a = float(np.linalg.norm(w))
        if a < 1e-12:
            return np.array([1.0, 0.0, 0.0, 0.0])
        return np.concatenate([[math.cos(0.5 * a)], w / a * math.sin(0.5 * a)])
    if tau < 4.0:
        th = rotation_angle(dr.u(), tau)
        ax = np.array([dr.n(), dr.n(), dr.n()])
        nrm = float(np.linalg.norm(ax))
        if nrm < 1e-14:
            return np.array([1.0, 0.0, 0.0, 0.0])
        return np.concatenate([[math.cos(0.5 * th)], ax / nrm * math.sin(0.5 * th)])
    qv = np.array([dr.n(), dr.n(), dr.n(), dr.n()])
    return qv / np.linalg.norm(qv)


def propagate(m: OuterModel, dr: Draws, pos, q):
    """One outer excursion from pos (|pos| >= 1.1 b): (returned, pos, q)."""
    pos = np.array(pos, dtype=float)
    t = 0.0
    back = out = False
    while not (back or out):
        if dr.c > 16000000:
            out = True
            break
        r = float(np.linalg.norm(pos))
        if r <= m.b:
            pos = m.b / r * pos
            back = True
        elif r >= m.q:
            if dr.u() < m.return_prob:
                pos = m.b / r * pos
                back = True
            else:
                out = True
            t = LARGE
        elif r < m.b_cover:
            D0 = m.D(r)
            surv, nx, dt = step_near(dr, r - m.b, m.force(r), D0)
            back = not surv
            t += dt
            pos = (m.b + nx) / r * pos
        elif r > m.q_cover:
            D0 = m.D(r)
            surv, nx, dt = step_near(dr, m.q - r, -m.force(r), D0)
            if not surv:
                if dr.u() < m.return_prob:
                    pos = m.b / r * pos
                    back = True
                else:
                    out = True
                t = LARGE
            else:
                pos = (m.q - nx) / r * pos
        else:
            Fr0 = m.force(r)
            L = m.L
            D0 = m.D(r)
            # step-size rule of outer_propagator.py:340-377, as written
            Fr1 = -(m.V * math.exp(-r / L) / r) * (1.0 / r + 1.0 / L) ** 2 - 2.0 * Fr0 / r
            alpha = 0.01
            Dfac = 6.0 * math.pi * m.dpre
            if m.hi:
                rm1 = 1.0 / r
                D1 = -3.0 * D0 * rm1 - Dfac * rm1 ** 2 / math.pi
                D2 = -4.0 * D1 * rm1 + Dfac * rm1 ** 3 / math.pi
                D3 = -5.0 * D2 * rm1 - 2.0 * Dfac * rm1 ** 4 / math.pi
                if abs(Fr0) > 0 and r < 3.0 * L:
                    num = D1 + D0 * Fr0
                    den = (D0 * D3 + (D1 + 2 * Fr0 * D0) * D2 + Fr0 * D1 ** 2
                           + (3 * Fr1 + Fr0 ** 2) * D0 * D1 + Fr0 * Fr1 * D0 ** 2)
                    dtf = alpha * abs(num / den) if abs(den) > 0 else LARGE
                else:
                    den = D0 * D3 + D1 * D2
                    dtf = alpha * abs(D1 / den) if abs(den) > 0 else LARGE
            else:
                dtf = (alpha / abs(D0 * Fr1) if abs(Fr1) > 0 else LARGE) \
                    if (abs(Fr0) > 0 and r < 3.0 * L) else LARGE
            dt = min(min(m.q - r, r - m.b) ** 2 / (18.0 * D0), dtf)
            t += dt
            ur = pos / r
            pos = pos + D0 * Fr0 * ur * dt
            sr = math.sqrt(2.0 * D0 * dt)
            if m.hi:
                rm1 = 1.0 / r
                Dp = Dfac * (m.ainv / (6.0 * math.pi) - 2.0 * (rm1 - 2.0 * m.a2 * rm1 ** 3)
                             / (8.0 * math.pi))
                sp = math.sqrt(2.0 * Dp * dt)
                x_, y_, z_ = ur
                rho = math.sqrt(x_ * x_ + y_ * y_)
                if rho == 0.0:
                    ut, up = np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0])
                else:
                    ut = np.array([z_ * x_ / rho, z_ * y_ / rho, -rho])
                    up = np.array([-y_ / rho, x_ / rho, 0.0])
                pos = pos + (sr * dr.n()) * ur
                pos = pos + (sp * dr.n()) * ut
                pos = pos + (sp * dr.n()) * up
            else:
                nx_, ny_, nz_ = dr.n(), dr.n(), dr.n()
                pos = pos + sr * np.array([nx_, ny_, nz_])
    if back:
        dq0 = diffusional_rotation(dr, t * m.drot0)
        dq1 = diffusional_rotation(dr, t * m.drot1)
        qn = qmul(dq1, q)
        qn = qmul(np.array([dq0[0], -dq0[1], -dq0[2], -dq0[3]]), qn)
        q = qn / np.linalg.norm(qn)
    return back, pos, np.asarray(q, dtype=float)
