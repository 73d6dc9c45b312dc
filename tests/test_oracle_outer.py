"""Pins of the LMZ outer-propagator oracle (oracle/outer.py) against closed
forms: the rotation-angle distribution's character moments, the radial
rate integral, and the probability of returning to the b-sphere from r0,
which for any U(r) and D(r) is rate(b) / rate(r0) (= b / r0 for free
diffusion).  Also checks the product's host parameters against it."""
import math

import numpy as np
import pytest

from oracle import outer as O
from pystarc_amd.simulation.outer_propagator import OuterPropagator


@pytest.mark.parametrize("tau", [0.3, 0.7, 1.5, 3.0])
def test_angle_cdf_characters(tau):
    """E[chi_l(theta)] = (2l+1) e^{-l(l+1) tau} for the SO(3) heat kernel."""
    assert O.angle_cdf(0.0, tau) == pytest.approx(0.0, abs=1e-15)
    assert O.angle_cdf(math.pi, tau) == pytest.approx(1.0, abs=1e-12)
    th = np.linspace(0.0, math.pi, 20001)
    F = np.array([O.angle_cdf(t, tau) for t in th])
    assert (np.diff(F) >= -1e-15).all()
    mid = 0.5 * (th[1:] + th[:-1])
    dF = np.diff(F)
    for l in (1, 2, 3):
        chi = np.sin((l + 0.5) * mid) / np.sin(0.5 * mid)
        assert float((chi * dF).sum()) == pytest.approx((2 * l + 1) * math.exp(-l * (l + 1) * tau),
                                                        abs=2e-6)


@pytest.mark.parametrize("tau", [0.6, 2.0])
def test_sampled_rotations(tau):
    n = 3000
    chi = np.empty(n)
    for i in range(n):
        qv = O.diffusional_rotation(O.Draws(11, i, 3), tau)
        th = 2.0 * math.acos(min(1.0, abs(qv[0])))
        chi[i] = 1.0 + 2.0 * math.cos(th)
    exact = 3.0 * math.exp(-2.0 * tau)
    assert abs(chi.mean() - exact) < 4.0 * chi.std() / math.sqrt(n)


def test_rate_free_diffusion():
    m = O.OuterModel(b=30.0, q=200.0, V=0.0, debye=7.86, D_t=0.05, hi=False, a0=10, a1=5,
                     drot0=1e-3, drot1=1e-3)
    for x in (30.0, 45.0, 200.0):
        assert m.rate(x) == pytest.approx(4.0 * math.pi * 0.05 * x, rel=1e-12)
    assert m.return_prob == pytest.approx(30.0 / 200.0, rel=1e-12)


@pytest.mark.parametrize("V,L", [(0.0, 7.86), (-40.0, 60.0), (40.0, 60.0)])
def test_return_probability_from_r0(V, L):
    """From r0 = 50 A: 0.600 free, 0.704 attractive, 0.485 repulsive -- apart
    by more than the test's 4-sigma band, so a wrong sign or factor in the
    drift, or a wrong LMZ probability at q, fails."""
    m = O.OuterModel(b=30.0, q=150.0, V=V, debye=L, D_t=0.05, hi=False, a0=10, a1=5,
                     drot0=1e-3, drot1=1e-3)
    r0 = 50.0
    n = 600
    back = 0
    for i in range(n):
        b, _, q = O.propagate(m, O.Draws(5, i, 9), np.array([0.0, r0, 0.0]), np.array([1.0, 0, 0, 0]))
        back += b
        assert abs(np.linalg.norm(q) - 1.0) < 1e-12
    exact = m.rate(30.0) / m.rate(r0)
    if V == 0.0:
        assert exact == pytest.approx(30.0 / r0, rel=1e-10)
    sd = math.sqrt(exact * (1 - exact) / n)
    print(V, back / n, exact, sd)
    assert abs(back / n - exact) < 4 * sd


@pytest.mark.parametrize("V,hi", [(-7.1, False), (-7.1, True), (12.0, True)])
def test_host_parameters_match_the_oracle(V, hi):
    h = OuterPropagator(30.0, 9.0, V, 7.86, 0.05, hi, 12.0, 8.0, 2e-3, 5e-3)
    m = O.OuterModel(30.0, 180.0, V, 7.86, 0.05, hi, 12.0, 8.0, 2e-3, 5e-3)
    assert h.qradius == 180.0 and h.trigger == pytest.approx(33.0)
    assert h.return_prob == pytest.approx(m.return_prob, rel=1e-7)
    assert h.bradius_cover == pytest.approx(m.b_cover, rel=1e-9)
    assert h.qradius_cover == pytest.approx(m.q_cover, rel=1e-9)
    with pytest.raises(ValueError):
        OuterPropagator(30.0, 1.5, V, 7.86, 0.05, hi, 12.0, 8.0)
