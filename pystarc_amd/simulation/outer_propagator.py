"""Host side of the LMZ outer propagator used by the batch step with
``outer="lmz"`` (the device loop is csrc/pst_outer.hpp).

Follows pystarc/simulation/outer_propagator.py:98-262: the q-sphere at 20 x
the larger molecular radius, the LMZ return probability rate(b) / rate(q)
with rate(x) = 4 pi / int_0^{1/x} exp(U(1/s)/kT) / D(1/s) ds, and the covers
of the b- and q-spheres inside which the Lamm-Schulten step takes over.  The
interaction U/kT = V e^{-r/L} / r and the radial diffusion D(r) are the batch
simulator's own (its k_b integrand), so the outer region continues the
inner one; the trigger radius is 1.1 b (nam_simulator.py:354-358).
"""
from __future__ import annotations

import math

from ..lib.numerical import romberg_integrate

QB_FACTOR = 1.1


class OuterPropagator:
    def __init__(self, b_radius, max_radius, V, debye_len, D_t, has_hi, a0=0.0, a1=0.0,
                 drot0=0.0, drot1=0.0):
        self.bradius = float(b_radius)
        self.qradius = 20.0 * float(max_radius)
        self.trigger = QB_FACTOR * self.bradius
        if self.qradius <= self.trigger:
            raise ValueError(f"q-sphere {self.qradius:.1f} A (20 x the larger molecular radius) "
                             f"does not lie beyond the trigger radius {self.trigger:.1f} A")
        self.V, self.debye_len, self.D_t = float(V), float(debye_len), float(D_t)
        self.has_hi = bool(has_hi)
        if self.has_hi and not (a0 > 0 and a1 > 0):
            raise ValueError("hydrodynamic interactions need both hydrodynamic radii")
        self.ainv = (1.0 / a0 + 1.0 / a1) if self.has_hi else 0.0
        self.a2 = 0.5 * (a0 * a0 + a1 * a1)
        self.dpre = self.D_t / self.ainv if self.has_hi else 0.0
        self.drot0, self.drot1 = float(drot0), float(drot1)
        rate_b, rate_q = self._relative_rate(self.bradius), self._relative_rate(self.qradius)
        self.return_prob = rate_b / rate_q if rate_q > 0 else 0.0
        self.bradius_cover = self._cover(True)
        self.qradius_cover = self._cover(False)

    def _D_parallel(self, r):
        if self.has_hi:
            return self.dpre * (self.ainv - 3.0 / r + 2.0 * self.a2 / r ** 3)
        return self.D_t

    def _relative_rate(self, x):
        def f(s):
            if s == 0.0:
                return 1.0 / self.D_t
            r = 1.0 / s
            return math.exp(self.V * math.exp(-r / self.debye_len) * s) / self._D_parallel(r)
        return 4.0 * math.pi / romberg_integrate(f, 0.0, 1.0 / x, tol=1e-10)

    def _radial_force(self, r):
        L = self.debye_len
        return self.V * math.exp(-r / L) / r * (1.0 / r + 1.0 / L)

    def _ts_boundary(self, rad, inner):
        # radius where the force's relative change reaches 5 % (:185-226)
        tol, L = 0.05, self.debye_len

        def rel(r):
            return (2 * L * L + 2 * r * L + r * r) * abs(r - rad) / (r * L * (L + r)) - tol
        if inner:
            rb = 2.0 * rad
            while rel(rb) < 0.0:
                rb *= 2.0
            lo, hi, fhi = rad, rb, rel(rb)
            while hi - lo > 1e-6 * rad:
                m = 0.5 * (lo + hi)
                fm = rel(m)
                if fm * fhi < 0.0:
                    lo = m
                else:
                    hi, fhi = m, fm
            return min(0.5 * (lo + hi), (1.0 + tol) * self.bradius)
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
        return max(0.5 * (lo + hi), (1.0 - tol) * self.qradius)

    def _cover(self, inner):
        # distance from the sphere where absorption within one Lamm-Schulten
        # step becomes likelier than 1 % (:228-262)
        if inner:
            F = self._radial_force(self.bradius)
            L = self._ts_boundary(self.bradius, True) - self.bradius
        else:
            F = -self._radial_force(self.qradius)
            L = self.qradius - self._ts_boundary(self.qradius, False)

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
        return self.bradius + x0 if inner else self.qradius - x0

    def fill(self, P):
        """Write the propagator's parameters into the step parameters."""
        P.lmz_on = 1
        P.lmz_hi = 1 if self.has_hi else 0
        P.lmz_trigger = self.trigger
        P.lmz_b, P.lmz_q = self.bradius, self.qradius
        P.lmz_b_cover, P.lmz_q_cover = self.bradius_cover, self.qradius_cover
        P.lmz_V, P.lmz_L = self.V, self.debye_len
        P.lmz_return_q = self.return_prob
        P.lmz_Dt, P.lmz_dpre, P.lmz_ainv, P.lmz_a2 = self.D_t, self.dpre, self.ainv, self.a2
        P.lmz_drot0, P.lmz_drot1 = self.drot0, self.drot1
