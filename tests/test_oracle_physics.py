"""Pins the oracle's scalar physics against known answers held by the
reference's tests (tests/test_pystarc.py) and against closed forms."""
import math

import numpy as np
import pytest
from scipy.integrate import quad

from oracle import physics as ph


# --- Romberg (reference tests:4277-4304) ------------------------------------
@pytest.mark.parametrize("n", [0, 1, 2, 3, 4, 5])
def test_romberg_power_integrals(n):
    assert abs(ph.romberg_integrate(lambda x: x ** n, 0.0, 1.0) - 1.0 / (n + 1)) < 1e-8


def test_romberg_sin():
    assert abs(ph.romberg_integrate(math.sin, 0.0, math.pi) - 2.0) < 1e-8


# --- k_b and return probability (reference tests:4246-4262) ------------------
def test_kb_two_spheres_reference_values():
    rp, kb = ph.return_prob_and_kb(10.0, 20.1, 1.0, -1.0, 7.828, 0.43371)
    assert abs(kb - 57.5) < 1.0
    assert abs(rp - 0.52) < 0.03


def test_kb_matches_scipy_quadrature():
    D, lam = 0.130992, 7.86
    V = -1.0 / (4 * math.pi * 78 * ph.EPS0)

    def igrd(s):
        return math.exp(V * math.exp(-1 / s / lam) * s) / D if s > 0 else 1 / D

    kb_q = 4 * math.pi / quad(igrd, 0, 1 / 30.0, limit=500)[0]
    rp, kb = ph.return_prob_and_kb(30.0, 60.0, 1.0, -1.0, lam, D)
    assert abs(kb - kb_q) / kb_q < 1e-7


def test_config1_analytic_numbers():
    """Two spheres R 10 + 2, a = 12.5, b = 30, q = 60, lambda = 7.86:
    k_b(30) = 49.423, return_prob = 0.500406 (analytic, mu = 0.243)."""
    rp, kb = ph.return_prob_and_kb(30.0, 60.0, 1.0, -1.0, 7.86, 0.130992)
    assert abs(kb - 49.423) < 2e-3
    assert abs(rp - 0.500406) < 2e-5


def test_free_diffusion_fallback():
    rp, kb = ph.return_prob_and_kb(30.0, 60.0, 0.0, -1.0, 7.86, 0.2)
    assert rp == 0.5 and abs(kb - 4 * math.pi * 0.2 * 30) < 1e-12


def test_stokes_D():
    # D_rel of two 1.005 A spheres (reference tests:3992-3995), SI viscosity
    assert abs(2 * ph.stokes_dt(1.005) - 0.43371) < 0.002
    assert abs(ph.stokes_dt(10.0) / ph.stokes_dt(20.0) - 2.0) < 1e-12
    assert abs(ph.stokes_dr(10.0) / ph.stokes_dr(20.0) - 8.0) < 1e-12


# --- Wilson interval (reference GPUBatchResult form) ---------------------------
def test_wilson_closed_form():
    lo, hi = ph.wilson_ci(40, 60)
    z = 1.959963984540054
    n, p = 100, 0.4
    c = (p + z * z / (2 * n)) / (1 + z * z / n)
    m = z * math.sqrt(p * (1 - p) / n + z * z / (4 * n * n)) / (1 + z * z / n)
    assert abs(lo - (c - m)) < 1e-9 and abs(hi - (c + m)) < 1e-9
    assert ph.wilson_ci(0, 0) == (0.0, 1.0)


def test_norm_ppf():
    assert abs(ph.norm_ppf(0.975) - 1.959963984540054) < 1e-9


# --- quaternions ------------------------------------------------------------------
def test_rotmat_orthonormal_and_known():
    rng = np.random.default_rng(1)
    q = ph.shoemake(rng.random((50, 3)))
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0)
    R = ph.quats_to_rotmats(q)
    assert np.allclose(R @ np.swapaxes(R, 1, 2), np.eye(3), atol=1e-12)
    assert np.allclose(np.linalg.det(R), 1.0)
    # 90 degrees about z maps x to y
    qz = np.array([[math.cos(math.pi / 4), 0, 0, math.sin(math.pi / 4)]])
    assert np.allclose(ph.quats_to_rotmats(qz)[0] @ [1, 0, 0], [0, 1, 0])


def test_apply_rotation_right_multiplies():
    q0 = np.array([[1.0, 0, 0, 0]])
    q1 = ph.apply_rotation(q0, np.array([[0, 0, math.pi / 2]]))
    assert np.allclose(ph.quats_to_rotmats(q1)[0] @ [1, 0, 0], [0, 1, 0])
    # body-frame composition: R(q (x) dq) = R(q) R(dq)
    rng = np.random.default_rng(2)
    q = ph.shoemake(rng.random((10, 3)))
    w = rng.normal(size=(10, 3))
    qn = ph.apply_rotation(q, w)
    dq = ph.apply_rotation(np.tile([1.0, 0, 0, 0], (10, 1)), w)
    assert np.allclose(ph.quats_to_rotmats(qn),
                       ph.quats_to_rotmats(q) @ ph.quats_to_rotmats(dq), atol=1e-12)
    # tiny angles leave q unchanged
    assert np.allclose(ph.apply_rotation(q, np.zeros((10, 3))), q)


def test_shoemake_uniform_mean_rotation():
    rng = np.random.default_rng(3)
    R = ph.quats_to_rotmats(ph.shoemake(rng.random((20000, 3))))
    assert np.abs(R.mean(axis=0)).max() < 0.02


# --- RPY tt block (rotne_prager.py:67-124) -------------------------------------
def test_rpy_offdiag_symmetric_and_continuous():
    a, b = 17.0, 15.0
    Da, Db = ph.stokes_dt(a), ph.stokes_dt(b)
    for r in (1.0, 5.0, 20.0, 40.0):
        M = ph.rpy_offdiagonal(np.array([r, 0.3, 0.1]), a, b, Da, Db)
        assert np.allclose(M, M.T)
    for r0 in (a + b, abs(a - b)):
        u = np.array([0.0, 0.0, 1.0])
        M1 = ph.rpy_offdiagonal(u * (r0 - 1e-9), a, b, Da, Db)
        M2 = ph.rpy_offdiagonal(u * (r0 + 1e-9), a, b, Da, Db)
        assert np.allclose(M1, M2, rtol=1e-6, atol=1e-12)
    assert np.allclose(ph.rpy_offdiagonal(np.zeros(3), a, b, Da, Db), 0)


def test_relative_diffusion_limits():
    D1, D2 = ph.stokes_dt(17.0), ph.stokes_dt(15.0)
    far = ph.relative_translational_diffusion(D1, D2, 17.0, 15.0, np.array([1e6, 0, 0]))
    assert abs(far - (D1 + D2)) / (D1 + D2) < 1e-4
    near = ph.relative_translational_diffusion(D1, D2, 17.0, 15.0, np.array([32.0, 0, 0]))
    assert near < D1 + D2
