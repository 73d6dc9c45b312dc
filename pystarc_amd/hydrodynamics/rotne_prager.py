"""Two-body hydrodynamics used by the batch step.

Scalar quantities of pystarc/hydrodynamics/rotne_prager.py: Stokes D_t / D_r
(:43-64), the three-regime Zuk et al. (2014) translational coupling block
(:67-124) and MobilityTensor (:127-215), plus the full 12x12 two-sphere RPY
mobility that the GPU step uses in its ``hydro_mode="rpy12"`` (no reference
counterpart; see ``rpy12_mobility``).
"""
from __future__ import annotations

import math

import numpy as np

from ..constants import ANG_TO_M, ETA_WATER, KB_SI, PS_TO_S, T_DEFAULT


def stokes_translational_diffusion(radius_ang, eta=ETA_WATER, T=T_DEFAULT):
    """kT / (6 pi eta a), A^2/ps."""
    return KB_SI * T / (6.0 * math.pi * eta * radius_ang * ANG_TO_M) / ANG_TO_M ** 2 * PS_TO_S


def stokes_rotational_diffusion(radius_ang, eta=ETA_WATER, T=T_DEFAULT):
    """kT / (8 pi eta a^3), rad^2/ps."""
    return KB_SI * T / (8.0 * math.pi * eta * (radius_ang * ANG_TO_M) ** 3) * PS_TO_S


def _tt_coeffs(r, a, b):
    """Coefficients (of I, of r_hat r_hat) of the tt coupling, viscosity units."""
    if r > a + b:
        s = a * a + b * b
        return (1.0 + s / (3.0 * r * r)) / (8.0 * math.pi * r), (1.0 - s / (r * r)) / (8.0 * math.pi * r)
    if r > abs(a - b):
        d2 = (a - b) ** 2
        den = 192.0 * math.pi * a * b * r ** 3
        return (16.0 * r ** 3 * (a + b) - (d2 + 3.0 * r * r) ** 2) / den, 3.0 * (d2 - r * r) ** 2 / den
    return 1.0 / (6.0 * math.pi * max(a, b)), 0.0


def rpy_offdiagonal(r_vec, a, b, D_a, D_b):
    """M_12 in A^2/ps, scaled by kT/eta = sqrt(6 pi a D_a 6 pi b D_b)."""
    r_vec = np.asarray(r_vec, dtype=float)
    r = float(np.linalg.norm(r_vec))
    if r < 1e-10:
        return np.zeros((3, 3))
    u = r_vec / r
    cI, cU = _tt_coeffs(r, a, b)
    scale = math.sqrt(D_a * 6.0 * math.pi * a * D_b * 6.0 * math.pi * b)
    return scale * (cI * np.eye(3) + cU * np.outer(u, u))


class MobilityTensor:
    """D_t / D_r of both molecules and their hydrodynamic radii."""

    def __init__(self, D_trans1, D_rot1, D_trans2, D_rot2, radius1=0.0, radius2=0.0, use_rpy=True):
        self.D_trans1 = D_trans1
        self.D_rot1 = D_rot1
        self.D_trans2 = D_trans2
        self.D_rot2 = D_rot2
        self.radius1 = radius1
        self.radius2 = radius2
        self.use_rpy = use_rpy

    @classmethod
    def from_radii(cls, radius1, radius2, eta=ETA_WATER, T=T_DEFAULT, use_rpy=True):
        return cls(stokes_translational_diffusion(radius1, eta, T),
                   stokes_rotational_diffusion(radius1, eta, T),
                   stokes_translational_diffusion(radius2, eta, T),
                   stokes_rotational_diffusion(radius2, eta, T),
                   radius1, radius2, use_rpy)

    def relative_translational_diffusion(self, r_vec=None):
        D0 = self.D_trans1 + self.D_trans2
        if not self.use_rpy or r_vec is None or self.radius1 <= 0 or self.radius2 <= 0:
            return D0
        M12 = rpy_offdiagonal(r_vec, self.radius1, self.radius2, self.D_trans1, self.D_trans2)
        return max(D0 - (2.0 / 3.0) * float(np.trace(M12)), 1e-12)

    def relative_rotational_diffusion(self):
        return self.D_rot1 + self.D_rot2


def rpy12_mobility(pos, a_rec, a_lig, ctx=None):
    """12x12 two-sphere grand mobility M and noise factor G (G G^T = M) per
    ligand position, evaluated on the GPU (``pst_rpy12_eval``), 1/eta units,
    order [x_rec, x_lig, theta_rec, theta_lig], receptor at the origin.
    ``pos`` (n,3) host array or device tensor; returns two (n,12,12) float64
    device tensors."""
    import torch

    from .._lib import Context, ptr
    if ctx is None:
        ctx = Context(0)
    p = torch.as_tensor(pos, dtype=torch.float64, device=f"cuda:{ctx.device}").contiguous()
    n = p.shape[0]
    M = torch.empty((n, 12, 12), dtype=torch.float64, device=p.device)
    G = torch.empty_like(M)
    ctx.check(ctx.lib.pst_rpy12_eval(ctx.h, n, ptr(p), float(a_rec), float(a_lig), ptr(M), ptr(G)),
              "pst_rpy12_eval")
    return M, G
