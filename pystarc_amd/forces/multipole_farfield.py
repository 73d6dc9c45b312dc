"""Receptor multipole moments for the Yukawa far-field
(pystarc/forces/multipole_farfield.py:64-110): total charge Q, dipole
p = sum q r, and quadrupole Q_ab = 1/2 sum q (3 r_a r_b - r^2 delta_ab)."""
from __future__ import annotations

import math

import numpy as np

from ..constants import EPS0


class MultipoleExpansion:
    def __init__(self, positions, charges, debye_length, sdie=78.0):
        r = np.asarray(positions, dtype=float)
        q = np.asarray(charges, dtype=float)
        self.debye = debye_length
        self.eps = sdie * EPS0
        self.four_pi_eps = 4.0 * math.pi * self.eps
        self.Q = float(q.sum())
        self.dipole = (q[:, None] * r).sum(axis=0)
        rr = np.einsum("n,na,nb->ab", q, r, r)
        self.quadrupole = 0.5 * (3.0 * rr - np.eye(3) * float((q * (r * r).sum(axis=1)).sum()))
        self.dipole_mag = float(np.linalg.norm(self.dipole))
        self.quad_mag = float(np.linalg.norm(self.quadrupole))
