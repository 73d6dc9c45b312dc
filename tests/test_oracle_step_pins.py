"""Pins of the oracle's per-step functions, each against something other than
its own formula: the reference tests' known answers, the physical meaning
of each criterion, an independent model (Stokes, the two-sphere RPY
mobility) or brute force (Brownian paths, rigid-body Monte Carlo).

  adaptive dt      pystarc/simulation/gpu_batch_simulator.py:961-1067
  HI diffusion     :740-751, :1074-1084
  Ermak-McCammon   :1085-1094
  overlap          :1102-1128
  bridge D_eff     :637-671
  bridge crossing  :1161-1181
  Yukawa V_fac     gpu_batch_engine.py:161-175
"""
import math

import numpy as np
import pytest

from oracle import forces as OF
from oracle import physics as ph
from oracle import rpy12
from oracle.simulator import OracleSim

D_REF = 0.43371   # D_rel of two 1.005 A spheres (reference tests:3992-3995)


def make_sim(D_t=D_REF, b=10.0, r_esc=20.0, q_rec=0.0, q_lig=0.0, debye=7.828, hi=False,
             a_rec=1.0, a_lig=1.0, max_dt=0.0, min_core_dt=0.0, rec_pos=None):
    """An OracleSim on a minimal system (one receptor atom, one ligand atom)
    whose constructor derives every step constant."""
    system = {"lig_body": np.zeros((1, 3)), "lig_q": np.array([q_lig]),
              "rec_pos": np.zeros((1, 3)) if rec_pos is None else rec_pos, "pairs": [],
              "n_needed": -1, "q_rec": q_rec, "q_lig": q_lig, "D_t": D_t, "D_r": 0.01,
              "a_rec": a_rec, "a_lig": a_lig, "engine": None, "debye": debye}
    params = {"n_trajectories": 4, "seed": 1, "r_start": b, "r_escape": r_esc,
              "max_steps": 10, "hydrodynamic_interactions": hi, "minimum_core_dt": min_core_dt,
              "max_dt": max_dt, "overlap_check": True, "gpu_force_batch": 0, "kT_scale": 1.0}
    return OracleSim(system, params)


# --- adaptive dt ---------------------------------------------------------------------
def test_dt_edge_reference_value_and_three_sigma_meaning():
    """At 0.5 A from a boundary with D = 0.43371 the edge criterion gives
    0.032 ps (reference tests:4177-4187, whose trigger radius is the edge the
    batch step uses as r_escape); it is the dt at which three standard
    deviations of the step reach the boundary."""
    s = make_sim(b=10.0, r_esc=20.0)
    dt = float(s.adaptive_dt(np.array([19.5]))[0])
    assert abs(dt - 0.032) < 0.005
    assert 3.0 * math.sqrt(2.0 * D_REF * dt) == pytest.approx(0.5, rel=1e-12)
    # nearer the b-sphere side of the outer zone
    dt2 = float(s.adaptive_dt(np.array([12.0]))[0])
    assert 3.0 * math.sqrt(2.0 * D_REF * dt2) == pytest.approx(2.0, rel=1e-12)


def test_dt_edge_only_in_the_outer_zone():
    """Inside 1.1 b the edge criterion is off and the 10 % rule holds: the
    RMS step (one axis) is a tenth of the separation.  At r = 10.5 the edge
    rule would give 0.032 ps."""
    s = make_sim(b=10.0, r_esc=20.0)
    for r in (10.2, 10.5, 10.99):
        dt = float(s.adaptive_dt(np.array([r]))[0])
        assert math.sqrt(2.0 * D_REF * dt) == pytest.approx(0.1 * r, rel=1e-12)
        assert dt > 1.0


def test_dt_edge_vanishes_at_the_escape_radius():
    s = make_sim(b=10.0, r_esc=20.0)
    assert float(s.adaptive_dt(np.array([20.0]))[0]) < 0.001


def _yukawa_force(V0, lam, r, h=1e-4):
    U = lambda x: V0 * math.exp(-x / lam) / x
    return -(U(r + h) - U(r - h)) / (2 * h)


@pytest.mark.parametrize("r", [3.0, 5.0, 8.0, 14.0, 20.0])
def test_dt_force_is_the_force_change_criterion(r):
    """dt_force = 0.01 / |D dF/dr| for the screened Coulomb force between the
    net charges (reference comment :969-970), with dF/dr taken by finite
    differences of the potential.  The reference's closed form differs from
    the exact derivative in one term, -V0 e^{-r/l} / l^2 where the exact
    derivative has -V0 e^{-r/l} / (r l^2); that term is kept (a quirk)."""
    q1, q2, lam = 5.0, -5.0, 7.828
    s = make_sim(b=40.0, r_esc=80.0, q_rec=q1, q_lig=q2, debye=lam)
    V0 = q1 * q2 / (4 * math.pi * 78.0 * ph.EPS0)
    h = 1e-3
    dF = (_yukawa_force(V0, lam, r + h) - _yukawa_force(V0, lam, r - h)) / (2 * h)
    quirk = V0 * math.exp(-r / lam) * (1.0 / (r * lam ** 2) - 1.0 / lam ** 2)
    dt_force = 0.01 / abs(D_REF * (dF + quirk))
    dt_pair = 0.005 * r * r / D_REF
    assert dt_force < dt_pair
    assert float(s.adaptive_dt(np.array([r]))[0]) == pytest.approx(dt_force, rel=2e-6)


def test_dt_force_off_beyond_three_debye_lengths_and_without_charge():
    s = make_sim(b=40.0, r_esc=80.0, q_rec=5.0, q_lig=-5.0, debye=7.828)
    r = 3 * 7.828 + 0.5
    assert math.sqrt(2 * D_REF * float(s.adaptive_dt(np.array([r]))[0])) == \
        pytest.approx(0.1 * r, rel=1e-12)
    s0 = make_sim(b=40.0, r_esc=80.0, q_rec=0.0, q_lig=-5.0)
    assert math.sqrt(2 * D_REF * float(s0.adaptive_dt(np.array([5.0]))[0])) == \
        pytest.approx(0.5, rel=1e-12)


def test_dt_cap_then_floor():
    """max_dt caps and minimum_core_dt floors, the floor applied last
    (:1062-1067)."""
    r = np.array([10.5])   # dt_pair = 1.27 ps
    assert float(make_sim(max_dt=0.5).adaptive_dt(r)[0]) == 0.5
    assert float(make_sim(min_core_dt=2.0).adaptive_dt(r)[0]) == 2.0
    assert float(make_sim(max_dt=0.05, min_core_dt=0.2).adaptive_dt(r)[0]) == 0.2


# --- hydrodynamic D_parallel ------------------------------------------------------------
@pytest.mark.parametrize("r", [9.0, 11.0, 20.0, 60.0])
def test_hi_diffusion_is_the_radial_rpy_relative_mobility(r):
    """With HI the relative diffusion along the line of centres is
    (kT/eta) u.(M_00 + M_11 - 2 M_01).u of the two-sphere RPY mobility in
    its separated regime (r > a0 + a1), kT/eta = 1/0.243."""
    a0, a1 = 5.0, 3.5
    s = make_sim(hi=True, a_rec=a0, a_lig=a1)
    u = np.array([0.6, 0.0, 0.8])
    M = rpy12.grand_mobility(r * u, a0, a1)
    Mrel = M[0:3, 0:3] + M[3:6, 3:6] - 2.0 * M[0:3, 3:6]
    expect = (1.0 / 0.243) * u @ Mrel @ u
    assert float(s.diffusion(np.array([r]))[0]) == pytest.approx(expect, rel=1e-12)


def test_hi_diffusion_far_limit_and_no_hi():
    a0, a1 = 5.0, 3.5
    s = make_sim(hi=True, a_rec=a0, a_lig=a1)
    stokes = (1.0 / 0.243) * (1 / (6 * math.pi * a0) + 1 / (6 * math.pi * a1))
    assert float(s.diffusion(np.array([1e7]))[0]) == pytest.approx(stokes, rel=1e-6)
    assert np.all(make_sim().diffusion(np.array([3.0, 30.0])) == D_REF)


# --- Ermak-McCammon translation --------------------------------------------------------
def test_em_translate_reference_numbers():
    """Drift -0.00396 A and noise scale 0.4169 A at D = 0.43371, F = -0.04561,
    dt = 0.2 (reference tests:4141-4151)."""
    p0 = np.zeros((1, 3))
    drift = ph.em_translate(p0, [D_REF], [0.2], np.array([[-0.04561, 0, 0]]), np.zeros((1, 3)))
    assert abs(drift[0, 0] - (-0.00396)) < 1e-4
    noise = ph.em_translate(p0, [D_REF], [0.2], np.zeros((1, 3)), np.array([[1.0, 0, 0]]))
    assert abs(noise[0, 0] - 0.4169) < 1e-3


def test_em_translate_moments():
    rng = np.random.default_rng(3)
    n = 200_000
    F = np.tile([[0.3, -0.2, 0.1]], (n, 1))
    X = ph.em_translate(np.zeros((n, 3)), np.full(n, 0.5), np.full(n, 0.4), F,
                        rng.standard_normal((n, 3)))
    np.testing.assert_allclose(X.mean(axis=0), 0.5 * 0.4 * F[0], atol=4 * math.sqrt(0.4 / n))
    np.testing.assert_allclose(X.var(axis=0), 2 * 0.5 * 0.4, rtol=0.01)


# --- overlap rejection ------------------------------------------------------------------
def test_overlap_threshold_hand_placed_atoms():
    rec = np.array([[0.0, 0.0, 0.0], [5.0, 0.0, 0.0]])
    c = np.array([[1.49, 0, 0], [1.51, 0, 0], [3.49, 0, 0], [3.51, 0, 0], [0, 1.2, 0.9]])
    # distances to the nearest atom: 1.49, 1.51, 1.51, 1.49, 1.5 (not below)
    assert ph.overlap_rejected(c, rec).tolist() == [True, False, False, True, False]
    # many receptor atoms (the 4,096-atom blocking)
    rng = np.random.default_rng(0)
    big = rng.uniform(-30, 30, size=(9000, 3))
    cc = rng.uniform(-30, 30, size=(50, 3))
    d = np.linalg.norm(cc[:, None, :] - big[None, :, :], axis=2).min(axis=1)
    assert np.array_equal(ph.overlap_rejected(cc, big), d < 1.5)


# --- Brownian bridge ---------------------------------------------------------------------
def _bridge_hits(x0, x1, Ddt, n_paths, K, rng):
    """Fraction of Brownian bridges from x0 to x1 (variance 2 D dt per unit
    time) whose path reaches 0, on K substeps, with the Broadie-Glasserman-Kou
    continuity correction (barrier moved by 0.5826 sigma_sub) for the
    continuous path between substeps."""
    sig = math.sqrt(2.0 * Ddt / K)
    shift = 0.5826 * sig
    t = np.arange(1, K + 1) / K
    hits = 0
    for c0 in range(0, n_paths, 1000):
        m = min(1000, n_paths - c0)
        S = np.cumsum(rng.standard_normal((m, K)) * sig, axis=1)
        B = S - t[None, :] * S[:, -1:]
        path = x0 + (x1 - x0) * t[None, :] + B
        hits += int((path.min(axis=1) < shift).sum())
    return hits / n_paths


@pytest.mark.parametrize("x0,x1,D,dt", [(0.3, 0.5, 0.4, 0.5), (0.8, 0.2, 0.13, 2.0),
                                        (0.05, 1.0, 0.43371, 0.2)])
def test_bridge_crossing_probability_against_brownian_paths(x0, x1, D, dt):
    rng = np.random.default_rng(11)
    n = 20000
    sim = _bridge_hits(x0, x1, D * dt, n, 2000, rng)
    p = float(ph.bridge_crossing_prob(np.array([x0]), np.array([x1]), np.array([D * dt]))[0])
    assert abs(sim - p) < 4.0 * math.sqrt(p * (1 - p) / n) + 0.005, (sim, p)


def test_bridge_crossing_zero_when_an_end_is_inside():
    p = ph.bridge_crossing_prob(np.array([-0.1, 0.2, 0.0]), np.array([0.3, -0.2, 0.4]),
                                np.array([0.1, 0.1, 0.1]))
    assert p.tolist() == [0.0, 0.0, 0.0]


def test_bridge_d_eff_is_stokes_translation_plus_rotation():
    """D_eff = D_t + D_r,rec L_rec^2 + D_r,lig L_lig^2 with the Stokes
    rotational coefficients kT/(8 pi eta a^3) of each body, when D_t is the
    Stokes relative translation kT/(6 pi eta)(1/a_rec + 1/a_lig)."""
    a_r, a_l = 17.0, 3.5
    D_t = ph.stokes_dt(a_r) + ph.stokes_dt(a_l)
    L_r, L_l = np.array([0.0, 12.0, 5.0]), np.array([4.0, 0.0, 2.5])
    got = ph.bridge_d_eff(D_t, a_r, a_l, L_r, L_l)
    expect = D_t + ph.stokes_dr(a_r) * L_r ** 2 + ph.stokes_dr(a_l) * L_l ** 2
    np.testing.assert_allclose(got, expect, rtol=1e-12)


def _rotvec_matrix(w):
    th = np.linalg.norm(w, axis=1)[:, None, None]
    k = w / np.maximum(np.linalg.norm(w, axis=1, keepdims=True), 1e-300)
    K = np.zeros((len(w), 3, 3))
    K[:, 0, 1], K[:, 0, 2], K[:, 1, 2] = -k[:, 2], k[:, 1], -k[:, 0]
    K[:, 1, 0], K[:, 2, 0], K[:, 2, 1] = k[:, 2], -k[:, 1], k[:, 0]
    return np.eye(3)[None] + np.sin(th) * K + (1 - np.cos(th)) * (K @ K)


def test_bridge_d_eff_against_rigid_body_monte_carlo():
    """Both bodies translate and turn by independent Stokes diffusion over a
    short dt; the pair distance between a receptor atom at lever L_rec and a
    ligand atom at lever L_lig, both levers across the pair axis, then
    spreads with variance 2 D_eff dt."""
    a_r, a_l, L_r, L_l, X = 6.0, 2.0, 4.0, 1.5, 9.0
    Dtr, Dtl = ph.stokes_dt(a_r), ph.stokes_dt(a_l)
    Drr, Drl = ph.stokes_dr(a_r), ph.stokes_dr(a_l)
    dt = 1e-3
    n = 400_000
    rng = np.random.default_rng(5)
    lr = np.array([0.0, L_r, 0.0])          # receptor atom about the receptor centre (origin)
    cl = np.array([X, L_r, -L_l])           # ligand centre
    ll = np.array([0.0, 0.0, L_l])          # ligand atom about the ligand centre
    Rr = _rotvec_matrix(rng.standard_normal((n, 3)) * math.sqrt(2 * Drr * dt))
    Rl = _rotvec_matrix(rng.standard_normal((n, 3)) * math.sqrt(2 * Drl * dt))
    tr = rng.standard_normal((n, 3)) * math.sqrt(2 * Dtr * dt)
    tl = rng.standard_normal((n, 3)) * math.sqrt(2 * Dtl * dt)
    rec_atom = tr + Rr @ lr
    lig_atom = cl + tl + Rl @ ll
    d = np.linalg.norm(lig_atom - rec_atom, axis=1)
    D_mc = d.var() / (2 * dt)
    D_eff = float(ph.bridge_d_eff(Dtr + Dtl, a_r, a_l, L_r, L_l))
    assert D_mc == pytest.approx(D_eff, rel=0.02)


# --- Yukawa prefactor (replaces arithmetic on literals) --------------------------------
def test_yukawa_vfac_and_potential_reference_values():
    """V_fac = -7.1847 for q_rec q_lig = -1 and phi(10 A) = -0.200268 at
    lambda = 7.828 (reference tests:4027-4034), from the oracle's far-field."""
    ff = OF.farfield_params(-1.0, 7.828)
    assert abs(ff["V_fac"] - (-7.1847)) < 0.01
    _, E, _, _ = OF.yukawa(np.array([[[10.0, 0.0, 0.0]]]), np.array([1.0]), ff)
    assert abs(E[0] - (-0.200268)) < 1e-3


# --- state-machine dt refinement ----------------------------------------------------------
def _sm_oracle(min_core_rxn_dt=0.0, min_core_dt=0.0):
    from oracle.production import ProductionOracle
    system = {"lig_body": np.zeros((1, 3)), "lig_q": np.array([0.0]), "rec_pos": np.zeros((1, 3)),
              "pairs": [(0, 0, 2.0)], "n_needed": 1, "q_rec": 0.0, "q_lig": 0.0, "D_t": D_REF,
              "D_r": 0.01, "a_rec": 1.0, "a_lig": 1.0, "engine": None, "debye": 7.828,
              "sm_reactions": [("free", "bound")], "first_state": "free"}
    params = {"n_trajectories": 3, "seed": 1, "r_start": 10.0, "r_escape": 20.0, "max_steps": 10,
              "hydrodynamic_interactions": False, "minimum_core_dt": min_core_dt, "max_dt": 0.0,
              "overlap_check": False, "gpu_force_batch": 0, "kT_scale": 1.0,
              "minimum_core_reaction_dt": min_core_rxn_dt}
    return ProductionOracle(system, params)


def test_sm_dt_refinement_twelve_sigma_rule():
    """Near a reachable reaction surface dt shrinks until twelve standard
    deviations of the step equal the distance to it (:1014-1060); an
    unreachable surface or a far one leaves dt alone."""
    o = _sm_oracle()
    run = np.array([0, 1, 2])
    r = np.array([5.0, 5.0, 5.0])
    old_d = np.array([[2.3], [12.0], [2.3]])
    o.sm_state[2] = 1                       # 'bound': no reaction leaves it
    dt = o._sm_dt(r, run, old_d)
    assert 12.0 * math.sqrt(2 * D_REF * dt[0]) == pytest.approx(0.3, rel=1e-12)
    base = float(o.adaptive_dt(r[:1])[0])
    assert dt[1] == base and dt[2] == base


def test_sm_dt_reaction_floor_then_core_floor():
    run = np.array([0])
    r = np.array([5.0])
    old_d = np.array([[2.3]])
    assert float(_sm_oracle(min_core_rxn_dt=0.05)._sm_dt(r, run, old_d)[0]) == 0.05
    # the core floor is applied after the refinement
    assert float(_sm_oracle(min_core_rxn_dt=0.05, min_core_dt=0.2)._sm_dt(r, run, old_d)[0]) == 0.2
