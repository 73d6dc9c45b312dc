"""LMZ outer propagator in the production step (outer="lmz") against the
per-trajectory oracle (the same Philox draws in the same order), and the
two-sphere reaction probability against the Debye-Smoluchowski result."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import outer as O  # noqa: E402
from oracle.physics import wilson_ci  # noqa: E402
from oracle.production import ProductionOracle  # noqa: E402
from pystarc_amd import systems as S  # noqa: E402
from pystarc_amd.simulation.gpu_batch_simulator import GPUBatchSimulator  # noqa: E402
from tests._systems import (dx_grids, numpy_grids, oracle_params, oracle_system,  # noqa: E402
                            small_protein_ligand, two_spheres_small)
from tests.test_oracle_sim import analytic_p_rxn  # noqa: E402


def _model(sim):
    h = sim.outer_propagator
    return O.OuterModel(h.bradius, h.qradius, h.V, h.debye_len, h.D_t, h.has_hi,
                        float(sim.mob.radius1 or 1.0), float(sim.mob.radius2 or 1.0),
                        h.drot0, h.drot1)


@pytest.mark.parametrize("name", ["two_spheres", "protein_ligand_hi"])
def test_lmz_production_matches_oracle(name):
    # K stays short of the 99.5 % early stop, whose step moves nobody
    if name == "two_spheres":
        sysm, K = two_spheres_small(n_traj=300), 150
    else:
        sysm, K = small_protein_ligand(n_traj=128), 150
        sysm.params.hydrodynamic_interactions = True
    N = int(sysm.params.n_trajectories)
    g = numpy_grids(sysm)
    eng = S.make_engine(sysm, dx_grids(g))
    sim = GPUBatchSimulator(sysm.mol1, sysm.mol2, sysm.mob, sysm.pathway_set, sysm.params, eng,
                            n_slots=N, precision="fp64", outer="lmz")
    drv = sim.production_driver(record=True)
    try:
        k = drv.advance(K)
        pos, q, st, ns = drv.state()
        out, steps = drv.outcomes()
        ctr = drv.counters()
    finally:
        drv.close()
    assert not ctr["stopped"]
    m = _model(sim)
    assert m.return_prob == pytest.approx(sim.outer_propagator.return_prob, rel=1e-7)
    orc = ProductionOracle(oracle_system(sysm, g, chunk=0), oracle_params(sysm), outer=m)
    for _ in range(k):
        orc.step()
    run_g, run_o = out == 0, orc.status == 0
    fin = ~run_g & ~run_o
    same = (out[fin] == orc.status[fin]) & (steps[fin] == orc.out_steps[fin])
    n_bad = int((~same).sum()) + int((run_g != run_o).sum())
    print(f"{name}: {k} steps, returns {ctr['n_returns']} (oracle {orc.n_returns}), escaped "
          f"{ctr['n_escaped']}, reacted {ctr['n_reacted']}, mismatches {n_bad}")
    assert n_bad == 0
    assert ctr["n_returns"] == orc.n_returns and ctr["n_returns"] > 0
    assert ctr["n_escaped"] == int((orc.status == 2).sum()) and ctr["n_escaped"] > 0
    both = run_g & run_o
    assert both.sum() > 0
    np.testing.assert_allclose(pos[both], orc.pos[both], rtol=0, atol=1e-7)
    np.testing.assert_allclose(q[both], orc.q[both], rtol=0, atol=1e-7)


def test_lmz_two_spheres_reaction_probability():
    """Two spheres (+1 / -1 e, contact 12.5 A, b = 30 A): with the outer region
    carried by the propagator the reaction probability is the exact
    Debye-Smoluchowski one (any escape radius gives the same value)."""
    n = 100_000
    sysm = two_spheres_small(n_traj=n)
    eng = S.make_engine(sysm, dx_grids(numpy_grids(sysm)))
    res = GPUBatchSimulator(sysm.mol1, sysm.mol2, sysm.mob, sysm.pathway_set, sysm.params, eng,
                            outer="lmz", n_slots=32768).run()
    assert res.n_max_steps <= 0.005 * n + 1      # left running by the 99.5 % early stop
    lo, hi = wilson_ci(res.n_reacted, res.n_escaped, confidence=0.95)
    P = analytic_p_rxn(12.5, 30.0, 60.0, 7.86)
    assert abs(P - analytic_p_rxn(12.5, 30.0, 400.0, 7.86)) < 1e-6
    print(f"lmz: P = {res.reaction_probability:.4f} [{lo:.4f}, {hi:.4f}], exact {P:.4f}, "
          f"returns {res.sim_data['n_returns_total']}")
    assert lo <= P <= hi


def test_lmz_needs_production_mode():
    sysm = two_spheres_small(n_traj=10)
    eng = S.make_engine(sysm, dx_grids(numpy_grids(sysm)))
    with pytest.raises(ValueError):
        GPUBatchSimulator(sysm.mol1, sysm.mol2, sysm.mob, sysm.pathway_set, sysm.params, eng,
                          rng="numpy", outer="lmz")
