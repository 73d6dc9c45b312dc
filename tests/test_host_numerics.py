"""Host-side numerics of the package (CPU): Romberg quadrature
(pystarc_amd/lib/numerical.py, used for k_b and the return probability),
the receptor multipole moments of the Yukawa far-field
(pystarc_amd/forces/multipole_farfield.py) and the Stokes / RPY
hydrodynamics (pystarc_amd/hydrodynamics/rotne_prager.py), each pinned by
closed forms, limits and invariants rather than by its own formula."""
import math

import numpy as np
import pytest

from pystarc_amd.forces.multipole_farfield import MultipoleExpansion
from pystarc_amd.hydrodynamics.rotne_prager import (MobilityTensor, _tt_coeffs, rpy_offdiagonal,
                                                    stokes_rotational_diffusion,
                                                    stokes_translational_diffusion)
from pystarc_amd.lib.numerical import romberg_integrate


@pytest.mark.parametrize("f,a,b,exact", [
    (lambda x: x ** 4, 0.0, 1.0, 0.2),
    (math.sin, 0.0, math.pi, 2.0),
    (math.exp, 0.0, 1.0, math.e - 1.0),
    (lambda x: 1.0 / (1.0 + x * x), 0.0, 1.0, math.pi / 4),
])
def test_romberg_closed_forms(f, a, b, exact):
    assert romberg_integrate(f, a, b) == pytest.approx(exact, abs=1e-9)


def test_romberg_is_exact_for_cubics():
    # one Richardson step is Simpson's rule, exact for cubics
    v = romberg_integrate(lambda x: 4 * x ** 3 - 3 * x * x + 2 * x - 1, -1.0, 2.0)
    F = lambda x: x ** 4 - x ** 3 + x * x - x             # antiderivative
    assert v == pytest.approx(F(2.0) - F(-1.0), abs=1e-12)


def test_romberg_k_b_of_a_sphere_pair():
    # k_b = 4 pi D / int_b^inf dr / r^2 = 4 pi D b for free diffusion; with
    # u = 1/r the integral is int_0^{1/b} du = 1/b
    b, D = 60.0, 0.0205
    inv = romberg_integrate(lambda u: 1.0, 0.0, 1.0 / b)
    assert 4 * math.pi * D / inv == pytest.approx(4 * math.pi * D * b, rel=1e-12)


def moments(pos, q):
    return MultipoleExpansion(np.asarray(pos, float), np.asarray(q, float), 7.86, 78.0)


def test_point_charge_at_origin_has_no_higher_moments():
    m = moments([[0.0, 0.0, 0.0]], [2.5])
    assert m.Q == 2.5
    assert np.allclose(m.dipole, 0) and np.allclose(m.quadrupole, 0)


def test_pure_dipole():
    d = np.array([0.3, -1.2, 2.0])
    m = moments([d / 2, -d / 2], [1.5, -1.5])
    assert m.Q == 0.0
    assert np.allclose(m.dipole, 1.5 * d)
    assert np.allclose(m.quadrupole, 0, atol=1e-14)     # symmetric pair: no quadrupole


def test_linear_quadrupole_textbook_values():
    # +q at +-a z, -2q at the origin: Theta_zz = 2 q a^2, Theta_xx = Theta_yy = -q a^2
    q, a = 0.7, 1.9
    m = moments([[0, 0, a], [0, 0, -a], [0, 0, 0]], [q, q, -2 * q])
    assert m.Q == pytest.approx(0.0, abs=1e-15)
    assert np.allclose(m.dipole, 0)
    assert np.allclose(np.diag(m.quadrupole), [-q * a * a, -q * a * a, 2 * q * a * a])
    assert np.allclose(m.quadrupole - np.diag(np.diag(m.quadrupole)), 0)


def test_quadrupole_is_symmetric_traceless_and_rotates_as_a_tensor():
    rng = np.random.default_rng(5)
    pos = rng.normal(scale=8.0, size=(200, 3))
    q = rng.normal(size=200)
    m = moments(pos, q)
    assert np.allclose(m.quadrupole, m.quadrupole.T)
    assert abs(np.trace(m.quadrupole)) < 1e-10 * np.abs(m.quadrupole).max()
    # random rotation: p -> R p, Theta -> R Theta R^T
    A = rng.normal(size=(3, 3))
    R, _ = np.linalg.qr(A)
    mr = moments(pos @ R.T, q)
    assert np.allclose(mr.dipole, R @ m.dipole)
    assert np.allclose(mr.quadrupole, R @ m.quadrupole @ R.T)


def test_neutral_dipole_does_not_depend_on_the_origin():
    rng = np.random.default_rng(6)
    pos = rng.normal(size=(50, 3))
    q = rng.normal(size=50)
    q -= q.mean()                                        # net charge zero
    shift = np.array([10.0, -4.0, 3.0])
    assert np.allclose(moments(pos + shift, q).dipole, moments(pos, q).dipole)


# --- host hydrodynamics (pystarc_amd/hydrodynamics/rotne_prager.py) -----------


def test_stokes_reference_value_and_scaling():
    # D_rel of two 1.005 A spheres is 0.43371 A^2/ps (reference tests:3992-3995)
    m = MobilityTensor.from_radii(1.005, 1.005)
    assert m.relative_translational_diffusion() == pytest.approx(0.43371, abs=5e-5)
    # kT/(6 pi eta a) over kT/(8 pi eta a^3) = 4 a^2 / 3
    for a in (1.0, 15.0, 17.0):
        ratio = stokes_translational_diffusion(a) / stokes_rotational_diffusion(a)
        assert ratio == pytest.approx(4 * a * a / 3, rel=1e-12)
    assert stokes_translational_diffusion(10.0) == pytest.approx(
        2 * stokes_translational_diffusion(20.0), rel=1e-12)


def test_rpy_coupling_limits():
    a, b = 17.0, 15.0
    # far apart: the Oseen tensor, both coefficients 1/(8 pi r)
    cI, cU = _tt_coeffs(1e6, a, b)
    assert cI == pytest.approx(1 / (8 * math.pi * 1e6), rel=1e-9)
    assert cU == pytest.approx(1 / (8 * math.pi * 1e6), rel=1e-9)
    # continuous where the regimes meet (contact and full overlap)
    for r0 in (a + b, abs(a - b)):
        lo, hi = _tt_coeffs(r0 * (1 - 1e-10), a, b), _tt_coeffs(r0 * (1 + 1e-10), a, b)
        assert lo == pytest.approx(hi, rel=1e-6, abs=1e-14)
    # one sphere inside the other: the larger sphere's self mobility
    assert _tt_coeffs(0.5, a, b) == (1 / (6 * math.pi * a), 0.0)
    # symmetric, and reduces relative diffusion at contact
    Da, Db = stokes_translational_diffusion(a), stokes_translational_diffusion(b)
    M = rpy_offdiagonal([20.0, 25.0, 3.0], a, b, Da, Db)
    assert np.allclose(M, M.T)
    mt = MobilityTensor.from_radii(a, b)
    assert mt.relative_translational_diffusion([32.0, 0, 0]) < Da + Db
    assert mt.relative_translational_diffusion([1e7, 0, 0]) == pytest.approx(Da + Db, rel=1e-5)
