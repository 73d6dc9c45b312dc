"""Pins of the production-step oracle: Philox4x32-10 known answers, the
batched RPY forms, and the whole per-trajectory step against the analytic
two-sphere reaction probability."""
import math

import numpy as np
import pytest

from oracle import philox as PH
from oracle import rpy12
from oracle.physics import wilson_ci
from oracle.production import ProductionOracle
from tests._systems import numpy_grids, oracle_params, oracle_system, two_spheres_small
from tests.test_oracle_sim import analytic_p_rxn


@pytest.mark.parametrize("ctr,key,out", [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
])
def test_philox_known_answers(ctr, key, out):
    """Random123 known-answer vectors for philox4x32-10."""
    got = PH.philox4x32_10(*ctr, *key)
    assert tuple(int(x) for x in got) == out


def test_philox_uniform_and_normal_moments():
    t = np.arange(200_000)
    u0, u1 = PH.uniform2(123, t, 7, PH.RNG_NOISE, 0)
    assert 0.0 <= u0.min() and u0.max() < 1.0
    assert abs(u0.mean() - 0.5) < 3e-3 and abs(u1.var() - 1 / 12) < 2e-3
    z0, z1 = PH.normal2(123, t, 7, PH.RNG_NOISE, 1)
    z = np.concatenate([z0, z1])
    assert abs(z.mean()) < 6e-3 and abs(z.var() - 1.0) < 1e-2
    assert abs(np.corrcoef(z0, z1)[0, 1]) < 1e-2
    # distinct streams / steps / subs give distinct numbers
    a, _ = PH.uniform2(123, t[:1000], 7, PH.RNG_NOISE, 0)
    b, _ = PH.uniform2(123, t[:1000], 7, PH.RNG_NOISE, 1)
    c, _ = PH.uniform2(123, t[:1000], 8, PH.RNG_NOISE, 0)
    d, _ = PH.uniform2(123, t[:1000], 7, PH.RNG_ESC, 0)
    assert len({*a, *b, *c, *d}) == 4000


def test_batched_rpy_matches_single_pose():
    rng = np.random.default_rng(3)
    n = 40
    u = rng.standard_normal((n, 3))
    rv = u / np.linalg.norm(u, axis=1, keepdims=True) * rng.uniform(0.5, 70, (n, 1))
    F, T = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    xi = rng.standard_normal((n, 12))
    dt = rng.uniform(0.1, 5.0, n)
    FR, TR = rpy12.receptor_load_v(rv, F, T, 1)
    dX = rpy12.displacement_v(rv, F, T, FR, TR, xi, dt, 4.1, 17.0, 15.0)
    pn, qn = rpy12.move_v(rv, q, dX)
    for i in range(n):
        fr, tr = rpy12.receptor_load(rv[i], F[i], T[i], 1)
        d1 = rpy12.displacement(rv[i], F[i], T[i], fr, tr, xi[i], dt[i], 4.1, 17.0, 15.0)
        np.testing.assert_allclose(dX[i], d1, rtol=0, atol=1e-13 * np.abs(d1).max())
        p1, q1 = rpy12.move(rv[i], q[i], d1)
        np.testing.assert_allclose(pn[i], p1, rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(qn[i], q1, rtol=1e-13, atol=1e-13)


def test_rpy_noise_covariance_by_sampling():
    """sqrt(2 Df dt) G xi has covariance 2 Df dt M."""
    rv = np.array([[3.0, -20.0, 12.0]])
    n = 400_000
    rng = np.random.default_rng(9)
    G = rpy12.aligned_factor(rv[0], 17.0, 15.0)
    X = rng.standard_normal((n, 12)) @ G.T
    C = X.T @ X / n
    M = rpy12.grand_mobility(rv[0], 17.0, 15.0)
    assert np.abs(C - M).max() < 0.02 * np.abs(M).max()


def test_production_oracle_two_spheres_matches_analytic():
    """The per-trajectory step with Philox draws gives the Debye-Smoluchowski
    reaction probability of two charged spheres."""
    sysm = two_spheres_small(n_traj=8000)
    g = numpy_grids(sysm)
    o = ProductionOracle(oracle_system(sysm, g), oracle_params(sysm))
    c = o.run()
    lo, hi = wilson_ci(c["n_reacted"], c["n_escaped"], confidence=0.95)
    P_exact = analytic_p_rxn(12.5, 30.0, 60.0, 7.86)
    assert lo <= P_exact <= hi, (c, lo, hi, P_exact)
    assert c["n_max_steps"] == 0
    assert (o.out_steps > 0).all()


def test_production_oracle_start_distribution():
    sysm = two_spheres_small(n_traj=20000)
    g = numpy_grids(sysm)
    o = ProductionOracle(oracle_system(sysm, g), oracle_params(sysm))
    r = np.linalg.norm(o.pos, axis=1)
    np.testing.assert_allclose(r, 30.0, rtol=1e-14)
    u = o.pos / r[:, None]
    assert np.abs(u.mean(axis=0)).max() < 0.03
    np.testing.assert_allclose(np.linalg.norm(o.q, axis=1), 1.0, rtol=1e-14)
    # Shoemake: uniform rotations -> E[w^2] = 1/4
    assert abs((o.q[:, 0] ** 2).mean() - 0.25) < 0.01


def test_production_oracle_shards_reproduce_the_whole_run():
    """Trajectories first .. first+count-1 run as a shard (random keys offset
    by first) are those trajectories of the whole run: same poses, outcomes
    and step counts after the same number of steps."""
    whole_sys = two_spheres_small(n_traj=60)
    g = numpy_grids(whole_sys)
    whole = ProductionOracle(oracle_system(whole_sys, g), oracle_params(whole_sys))
    parts = []
    for first, count in ((0, 25), (25, 35)):
        s = two_spheres_small(n_traj=count)
        parts.append(ProductionOracle(oracle_system(s, g), oracle_params(s), traj_base=first))
    for _ in range(120):
        whole.step()
        for p in parts:
            p.step()
    pos = np.concatenate([p.pos for p in parts])
    q = np.concatenate([p.q for p in parts])
    st = np.concatenate([p.status for p in parts])
    ns = np.concatenate([p.out_steps for p in parts])
    assert np.array_equal(pos, whole.pos) and np.array_equal(q, whole.q)
    assert np.array_equal(st, whole.status) and np.array_equal(ns, whole.out_steps)
    assert (whole.status != 0).sum() > 0
    # and a shard is not simply the whole run's first trajectories re-labelled
    s = two_spheres_small(n_traj=35)
    unshifted = ProductionOracle(oracle_system(s, g), oracle_params(s))
    assert not np.array_equal(unshifted.pos, whole.pos[25:])
