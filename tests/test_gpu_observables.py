"""Device-collected observables of the production step (record level 2)
against the production oracle's restatement of the reference's sim_data
(gpu_batch_simulator.py:576-620, :1255-1340): per-trajectory records and the
run-wide histograms, on the same seed, after K steps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle.production import ProductionOracle  # noqa: E402
from pystarc_amd import systems as S  # noqa: E402
from pystarc_amd.simulation.gpu_batch_simulator import GPUBatchSimulator  # noqa: E402
from tests._systems import (dx_grids, numpy_grids, oracle_params, oracle_system,  # noqa: E402
                            small_protein_protein, two_spheres_small)


def _case(name):
    if name == "two_spheres":
        return two_spheres_small(n_traj=500), "scalar", "reference", 300
    return small_protein_protein(n_traj=128), "rpy12", "physical", 150


@pytest.mark.parametrize("name", ["two_spheres", "protein_protein"])
def test_observables_match_oracle(name):
    sysm, hydro, torque, K = _case(name)
    N = int(sysm.params.n_trajectories)
    g = numpy_grids(sysm)
    eng = S.make_engine(sysm, dx_grids(g), torque_mode=torque)
    sim = GPUBatchSimulator(sysm.mol1, sysm.mol2, sysm.mob, sysm.pathway_set, sysm.params, eng,
                            n_slots=N, precision="fp64", hydro=hydro)
    drv = sim.production_driver(record="observables")
    try:
        k = drv.advance(K)
        out, steps = drv.outcomes()
        sd = drv.observables()
    finally:
        drv.close()
    orc = ProductionOracle(oracle_system(sysm, g, chunk=0), oracle_params(sysm), torque_mode=torque,
                           hydro=hydro, observe=True)
    for _ in range(k):
        orc.step()
    tr = orc.observables()
    assert np.array_equal(out, orc.status)
    np.testing.assert_allclose(sd["start_pos"], tr[:, 0:3], rtol=0, atol=1e-12)
    np.testing.assert_allclose(sd["start_q"], tr[:, 3:7], rtol=0, atol=1e-12)
    np.testing.assert_allclose(sd["min_dist"], tr[:, 7], rtol=0, atol=1e-8)
    assert np.array_equal(sd["step_at_min"], tr[:, 8].astype(np.int64))
    np.testing.assert_allclose(sd["total_time_ps"], tr[:, 16], rtol=1e-12, atol=0)
    assert np.array_equal(sd["n_returns"], tr[:, 17].astype(np.int64))
    assert np.array_equal(sd["bb_triggered"], tr[:, 18].astype(np.int64))
    rx = out == 1
    np.testing.assert_allclose(sd["encounter_pos"], tr[rx, 19:22], rtol=0, atol=1e-8)
    assert np.array_equal(sd["encounter_n_pairs"], tr[rx, 26].astype(np.int64))
    es = out == 2
    np.testing.assert_allclose(sd["near_miss_pos"], tr[es, 9:12], rtol=0, atol=1e-8)
    assert np.array_equal(sd["radial_counts"], orc.h_rad)
    assert np.array_equal(sd["angular_counts"], orc.h_ang)
    assert np.array_equal(sd["milestone_flux_out"], orc.h_ms_out)
    assert np.array_equal(sd["milestone_flux_in"], orc.h_ms_in)
    assert np.array_equal(sd["trans_matrix"], orc.h_trans)
    assert np.array_equal(sd["contact_pair_counts"], orc.h_contact)
    assert sd["contact_total_steps"] == orc.h_contact_steps == orc.total_steps
    assert rx.sum() > 0 and sd["radial_counts"].sum() == orc.total_steps


def test_run_with_observables_sim_data():
    sysm = two_spheres_small(n_traj=3000)
    eng = S.make_engine(sysm, dx_grids(numpy_grids(sysm)))
    res = GPUBatchSimulator(sysm.mol1, sysm.mol2, sysm.mob, sysm.pathway_set, sysm.params, eng,
                            record="observables").run()
    sd = res.sim_data
    for key in ("start_pos", "min_dist", "step_at_min", "total_time_ps", "n_returns",
                "bb_triggered", "encounter_pos", "near_miss_pos", "radial_counts",
                "angular_counts", "milestone_flux_out", "trans_matrix", "contact_pair_counts"):
        assert key in sd, key
    assert sd["radial_counts"].sum() == res.total_steps == sd["trans_matrix"].sum()
    assert len(sd["encounter_traj"]) == res.n_reacted
    # every trajectory that moved has a closest approach; only reactions carry
    # the bridge flag (set, as in the reference, for every reaction decided
    # in the bridge phase of a step)
    moved = sd["n_steps"] > 0
    assert (sd["min_dist"][moved] < 1e29).all()
    assert (sd["bb_triggered"][sd["outcome"] != 1] == 0).all()
    assert (sd["encounter_n_pairs"] <= 1).all()
    assert (sd["total_time_ps"][sd["outcome"] > 0] > 0).all()
