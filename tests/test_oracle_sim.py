"""The oracle's whole step sequence, pinned end to end by the analytic
Debye-Smoluchowski reaction probability of two charged spheres
(examples/two_charged_spheres/analytical.py:50-70 formulas)."""
import math

import numpy as np
from scipy.integrate import quad

from oracle.physics import EPS0, wilson_ci
from oracle.simulator import OracleSim
from tests._systems import numpy_grids, oracle_params, oracle_system, two_spheres_small


def analytic_p_rxn(a, b, q, lam, qq=-1.0):
    V = qq / (4 * math.pi * 78 * EPS0)

    def S(r):
        return math.exp(V * math.exp(-r / lam) / r) / r ** 2

    def igrd(s):
        return math.exp(V * math.exp(-1 / s / lam) * s) if s > 0 else 1.0

    kb = 1.0 / quad(igrd, 0, 1 / b, limit=500)[0]
    kq = 1.0 / quad(igrd, 0, 1 / q, limit=500)[0]
    rp = kb / kq
    Ps = quad(S, b, q, limit=500)[0] / quad(S, a, q, limit=500)[0]
    return Ps / (1 - (1 - Ps) * rp)


def test_analytic_p_rxn_number():
    assert abs(analytic_p_rxn(12.5, 30.0, 60.0, 7.86) - 0.428338) < 2e-5
    # the two-charged-spheres example: exact P_rxn 0.4501
    # (examples/two_charged_spheres/README.md:18)
    assert abs(analytic_p_rxn(2.5, 10.0, 20.0, 7.828) - 0.4501) < 1e-4


def test_two_spheres_oracle_matches_analytic():
    n = 8000
    sysm = two_spheres_small(n_traj=n)
    g = numpy_grids(sysm)
    sim = OracleSim(oracle_system(sysm, g), oracle_params(sysm))
    c = sim.run()
    lo, hi = wilson_ci(c["n_reacted"], c["n_escaped"], confidence=0.95)
    P_exact = analytic_p_rxn(12.5, 30.0, 60.0, 7.86)
    assert lo <= P_exact <= hi, (c, lo, hi, P_exact)
    assert c["n_reacted"] + c["n_escaped"] >= int(0.995 * n)
