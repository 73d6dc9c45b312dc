"""Pins of the two-body RPY grand mobility oracle (oracle/rpy12.py)."""
import math

import numpy as np
import pytest

from oracle import rpy12
from pystarc_amd.hydrodynamics.rotne_prager import _tt_coeffs

PAIRS = [(17.0, 15.0), (15.0, 17.0), (1.0, 5.0), (3.0, 3.0)]


def _coeffs(r, a0, a1):
    return np.array(rpy12.cross_coeffs(r, a0, a1))


@pytest.mark.parametrize("a0,a1", PAIRS)
def test_continuity_at_regime_boundaries(a0, a1):
    h = 1e-9
    for rb in (a0 + a1, abs(a0 - a1)):
        if rb <= 0:
            continue
        lo, hi = _coeffs(rb - h, a0, a1), _coeffs(rb + h, a0, a1)
        scale = np.abs(_coeffs(rb, a0, a1)).max()
        np.testing.assert_allclose(lo, hi, rtol=0, atol=1e-6 * scale)


def test_equal_radius_textbook_forms():
    """Rotne & Prager (1969) / Wajnryb et al. (2013), equal spheres."""
    a = 3.0
    for r in (0.7, 2.0, 4.5, 5.9):  # overlapping, r < 2a
        al, be, ga, de, x01, x10 = rpy12.cross_coeffs(r, a, a)
        assert al == pytest.approx((1 - 9 * r / (32 * a)) / (6 * math.pi * a), rel=1e-13)
        assert be == pytest.approx((3 * r / (32 * a)) / (6 * math.pi * a), rel=1e-13)
        assert ga == pytest.approx((1 - 27 * r / (32 * a) + 5 * r ** 3 / (64 * a ** 3))
                                   / (8 * math.pi * a ** 3), rel=1e-12)
        assert de == pytest.approx((9 * r / (32 * a) - 3 * r ** 3 / (64 * a ** 3))
                                   / (8 * math.pi * a ** 3), rel=1e-12)
        xr = (r / a - 3 * r * r / (8 * a * a)) / (16 * math.pi * a * a)
        assert x01 == pytest.approx(xr, rel=1e-12) and x10 == pytest.approx(xr, rel=1e-12)
    for r in (6.5, 10.0, 40.0):  # separated
        al, be, ga, de, x01, x10 = rpy12.cross_coeffs(r, a, a)
        assert al == pytest.approx((1 + 2 * a * a / (3 * r * r)) / (8 * math.pi * r), rel=1e-13)
        assert be == pytest.approx((1 - 2 * a * a / (r * r)) / (8 * math.pi * r), rel=1e-13)
        assert ga == pytest.approx(-1 / (16 * math.pi * r ** 3), rel=1e-13)
        assert de == pytest.approx(3 / (16 * math.pi * r ** 3), rel=1e-13)
        assert x01 == pytest.approx(1 / (8 * math.pi * r * r), rel=1e-13)


@pytest.mark.parametrize("a0,a1", PAIRS)
def test_tt_block_matches_reference_restatement(a0, a1):
    """The translational coupling equals the reference's three-regime block
    (rotne_prager.py:101-123, restated in the host mirror)."""
    for r in np.linspace(0.05, 3.0 * (a0 + a1), 37):
        al, be = rpy12.cross_coeffs(r, a0, a1)[:2]
        cI, cU = _tt_coeffs(r, a0, a1)
        assert al == pytest.approx(cI, rel=1e-12, abs=1e-15)
        assert be == pytest.approx(cU, rel=1e-12, abs=1e-15)


def test_far_regime_matches_reference_d_parallel():
    """D_par = Df (m0 + m1 - 2 (al + be)) is gpu_batch_simulator.py:1074-1084."""
    Df, a0, a1 = 1.0 / 0.243, 17.0, 15.0
    a2 = 0.5 * (a0 * a0 + a1 * a1)
    for r in (33.0, 40.0, 90.0):
        al, be = rpy12.cross_coeffs(r, a0, a1)[:2]
        ours = Df * (1 / (6 * math.pi * a0) + 1 / (6 * math.pi * a1) - 2 * (al + be))
        ref = (Df * (1 / (6 * math.pi * a0) + 1 / (6 * math.pi * a1))
               - 2 * Df * (2 / r - (4.0 / 3.0) * a2 / r ** 3) / (8 * math.pi))
        assert ours == pytest.approx(ref, rel=1e-13)


def _oseen_mc(r, a0, a1, n=1_500_000, seed=5):
    """Surface average of the Oseen tensor (I + w w/|w|^2)/(8 pi |w|) over a
    point of sphere 0 (at the origin) and a point of sphere 1 (at r e_z)."""
    rng = np.random.default_rng(seed)
    acc = np.zeros((3, 3))
    done = 0
    while done < n:
        k = min(250_000, n - done)
        p = rng.standard_normal((k, 3))
        q = rng.standard_normal((k, 3))
        p *= a0 / np.linalg.norm(p, axis=1, keepdims=True)
        q *= a1 / np.linalg.norm(q, axis=1, keepdims=True)
        w = q - p
        w[:, 2] += r
        d = np.linalg.norm(w, axis=1)
        wh = w / d[:, None]
        acc += (np.eye(3)[None] + wh[:, :, None] * wh[:, None, :]).__truediv__(
            8 * math.pi * d[:, None, None]).sum(axis=0)
        done += k
    return acc / n


@pytest.mark.parametrize("r", [40.0, 26.0, 20.0, 1.0])
def test_tt_against_surface_average_of_oseen_tensor(r):
    """Zuk et al. define the tt coupling as this double surface average."""
    a0, a1 = 17.0, 15.0
    mc = _oseen_mc(r, a0, a1)
    al, be = rpy12.cross_coeffs(r, a0, a1)[:2]
    exact = np.diag([al, al, al + be])
    assert np.abs(mc - exact).max() < 3e-3 * (al + abs(be))


@pytest.mark.parametrize("a0,a1", PAIRS)
def test_positive_definite_all_regimes(a0, a1):
    rng = np.random.default_rng(7)
    for r in np.concatenate([np.linspace(0.01, 2.5 * (a0 + a1), 120), [abs(a0 - a1), a0 + a1]]):
        if r <= 0:
            continue
        u = rng.standard_normal(3)
        M = rpy12.grand_mobility(u / np.linalg.norm(u) * r, a0, a1)
        assert np.allclose(M, M.T, rtol=0, atol=1e-18)
        w = np.linalg.eigvalsh(M)
        assert w.min() > 0, (r, w.min())


def test_far_apart_is_stokes_self_mobility():
    a0, a1 = 17.0, 15.0
    M = rpy12.grand_mobility([0.0, 3e7, 4e7], a0, a1)
    D = np.diag([1 / (6 * math.pi * a0)] * 3 + [1 / (6 * math.pi * a1)] * 3
                + [1 / (8 * math.pi * a0 ** 3)] * 3 + [1 / (8 * math.pi * a1 ** 3)] * 3)
    off = M - D
    assert np.abs(off[:6, :6]).max() < 1e-6 * D[0, 0]
    assert np.abs(off[6:, 6:]).max() < 1e-9 * D[6, 6]


def test_far_field_rotation_by_stokeslet():
    """A force F on the receptor turns the far ligand at F x u / (8 pi r^2)."""
    a0, a1, r = 2.0, 1.0, 50.0
    u = np.array([0.6, 0.0, 0.8])
    M = rpy12.grand_mobility(r * u, a0, a1)
    F = np.array([0.3, -1.2, 0.5])
    f = np.concatenate([F, np.zeros(9)])
    w_lig = (M @ f)[9:12]
    np.testing.assert_allclose(w_lig, np.cross(F, u) / (8 * math.pi * r * r), rtol=1e-12)


@pytest.mark.parametrize("a0,a1", PAIRS)
def test_aligned_factor_reproduces_mobility(a0, a1):
    rng = np.random.default_rng(11)
    for r in np.linspace(0.05, 2.5 * (a0 + a1), 41):
        u = rng.standard_normal(3)
        rv = u / np.linalg.norm(u) * r
        M = rpy12.grand_mobility(rv, a0, a1)
        G = rpy12.aligned_factor(rv, a0, a1)
        np.testing.assert_allclose(G @ G.T, M, rtol=0, atol=1e-13 * np.abs(M).max())
    # poles: u = +-e_z
    for s in (1.0, -1.0):
        rv = np.array([0.0, 0.0, s * 0.7 * (a0 + a1)])
        M = rpy12.grand_mobility(rv, a0, a1)
        G = rpy12.aligned_factor(rv, a0, a1)
        np.testing.assert_allclose(G @ G.T, M, rtol=0, atol=1e-13 * np.abs(M).max())


def test_blocks_cholesky_matches_lapack():
    A, Cm, Dm = rpy12.group_blocks(24.0, 17.0, 15.0)
    for X in (A, Cm, Dm):
        L = np.linalg.cholesky(X)
        assert np.allclose(np.triu(L, 1), 0)
        np.testing.assert_allclose(L @ L.T, X, rtol=0, atol=1e-15 * np.abs(X).max())


def test_move_frame_change():
    pos = np.array([10.0, -4.0, 21.0])
    q = np.array([0.5, 0.5, -0.5, 0.5])
    p1, q1 = rpy12.move(pos, q, np.zeros(12))
    np.testing.assert_allclose(p1, pos, rtol=1e-15)
    np.testing.assert_allclose(q1, q, rtol=1e-15)
    # receptor turns by theta about z: the ligand appears turned by -theta
    th = 0.3
    dX = np.zeros(12)
    dX[8] = th
    p2, q2 = rpy12.move(pos, q, dX)
    c, s = math.cos(th), math.sin(th)
    Rz = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    np.testing.assert_allclose(p2, Rz.T @ pos, rtol=1e-14)
    np.testing.assert_allclose(rpy12.qrot(q2), Rz.T @ rpy12.qrot(q), atol=1e-14)
    # a joint translation of both bodies changes nothing
    dX = np.zeros(12)
    dX[0:3] = dX[3:6] = [1.0, 2.0, -3.0]
    p3, _ = rpy12.move(pos, q, dX)
    np.testing.assert_allclose(p3, pos, rtol=1e-14)
    # the same rotation of both bodies leaves the relative orientation
    dX = np.zeros(12)
    dX[6:9] = dX[9:12] = [0.1, -0.2, 0.05]
    _, q4 = rpy12.move(pos, q, dX)
    np.testing.assert_allclose(rpy12.qrot(q4), rpy12.qrot(q), atol=1e-14)


def test_receptor_load():
    F = np.array([1.0, 2.0, 3.0])
    T = np.array([0.1, 0.0, -0.2])
    pos = np.array([20.0, 0.0, 5.0])
    Fr, Tr = rpy12.receptor_load(pos, F, T, 0)
    np.testing.assert_array_equal(Fr, -F)
    np.testing.assert_array_equal(Tr, 0.0)
    Fr, Tr = rpy12.receptor_load(pos, F, T, 1)
    # total torque about the origin vanishes
    np.testing.assert_allclose(Tr + T + np.cross(pos, F), 0.0, atol=1e-14)
