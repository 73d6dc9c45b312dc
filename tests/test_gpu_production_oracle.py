"""Production kernels (Philox noise, refill-free slot pool) against the
per-trajectory production oracle (oracle/production.py) on the same seed.

After K device steps every still-running trajectory must sit at the oracle's
pose (FP64 forces, so poses agree to 1e-8 A / 1e-8 in the quaternion), and
every finished one must carry the oracle's outcome and step count.  Covers the
scalar step (with and without HI D_parallel) and the two-body RPY move with
zero and physical torques.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle.production import ProductionOracle  # noqa: E402
from pystarc_amd import systems as S  # noqa: E402
from pystarc_amd.simulation.gpu_batch_simulator import GPUBatchSimulator  # noqa: E402
from tests._systems import (dx_grids, numpy_grids, oracle_params, oracle_system,  # noqa: E402
                            small_protein_ligand, small_protein_protein,
                            state_machine_protein_protein, two_spheres_small)


def _case(name):
    if name == "two_spheres":
        return two_spheres_small(n_traj=600), "scalar", "reference", 400
    if name == "two_spheres_shard":
        # trajectories 200 .. 499 of a run, as one part of a split run
        return two_spheres_small(n_traj=300), "scalar", "reference", 400
    if name == "protein_ligand_hi":
        s = small_protein_ligand(n_traj=128)
        s.params.hydrodynamic_interactions = True
        return s, "scalar", "reference", 150
    if name == "protein_protein_rpy12":
        return small_protein_protein(n_traj=128), "rpy12", "physical", 150
    if name == "protein_protein_rpy12_zero_torque":
        return small_protein_protein(n_traj=128, seed=10), "rpy12", "reference", 150
    if name == "state_machine":
        return state_machine_protein_protein(n_traj=128), "scalar", "reference", 300
    raise KeyError(name)


@pytest.mark.parametrize("name", ["two_spheres", "two_spheres_shard", "protein_ligand_hi",
                                  "protein_protein_rpy12", "protein_protein_rpy12_zero_torque",
                                  "state_machine"])
def test_production_matches_oracle_per_trajectory(name):
    sysm, hydro, torque, K = _case(name)
    N = int(sysm.params.n_trajectories)
    base = 200 if name.endswith("_shard") else 0
    g = numpy_grids(sysm)
    eng = S.make_engine(sysm, dx_grids(g), torque_mode=torque)
    sim = GPUBatchSimulator(sysm.mol1, sysm.mol2, sysm.mob, sysm.pathway_set, sysm.params, eng,
                            n_slots=N, precision="fp64", hydro=hydro,
                            traj_range=(base, N) if base else None)
    drv = sim.production_driver(record=True)
    try:
        k = drv.advance(K)
        pos, q, st, ns = drv.state()
        out, steps = drv.outcomes()
        ctr = drv.counters()
        sm = drv.completed_reactions()
    finally:
        drv.close()
    # chunk=0: reverse Born evaluated for every trajectory.  The production
    # kernel culls per trajectory exactly (a receptor atom farther than the
    # ligand grid's corner radius cannot contribute); the reference's
    # batch-wide skip (gpu_batch_engine.py:633-639) compares against the x
    # extent only and depends on which trajectories share a chunk.
    orc = ProductionOracle(oracle_system(sysm, g, chunk=0), oracle_params(sysm),
                           torque_mode=torque, hydro=hydro, traj_base=base)
    assert abs(orc.return_prob - sim._return_prob) < 1e-12
    for _ in range(k):
        orc.step()
    run_g = out == 0
    run_o = orc.status == 0
    fin = ~run_g & ~run_o
    same = (out[fin] == orc.status[fin]) & (steps[fin] == orc.out_steps[fin])
    n_bad = int((~same).sum()) + int((run_g != run_o).sum())
    print(f"{name}: {k} steps, running {int(run_g.sum())}, finished {int(fin.sum())}, "
          f"mismatches {n_bad}, counters {ctr}")
    assert n_bad == 0
    assert fin.sum() > 0
    both = run_g & run_o
    assert both.sum() > 0
    assert (ns[both] == k).all()
    np.testing.assert_allclose(pos[both], orc.pos[both], rtol=0, atol=1e-8)
    np.testing.assert_allclose(q[both], orc.q[both], rtol=0, atol=1e-8)
    assert ctr["total_steps"] == orc.total_steps
    assert ctr["n_reacted"] == int((orc.status == 1).sum())
    assert ctr["n_escaped"] == int((orc.status == 2).sum())
    if orc.sm is not None:
        print("state machine:", sm)
        assert [r["n_fired_discrete"] for r in sm] == orc.sm_fired.tolist()
        assert [r["n_fired_bridge"] for r in sm] == orc.sm_fired_bridge.tolist()
        assert [r["unique_trajectories"] for r in sm] == [len(x) for x in orc.sm_sets]
        assert sum(r["n_fired"] for r in sm) > 0


def test_shards_reproduce_the_whole_run():
    """Two parts of a run (trajectory ranges, same seed) on the device give
    each trajectory the outcome and step count it has in the whole run."""
    sysm = small_protein_ligand(n_traj=500)
    sysm.params.max_steps = 20000
    eng = S.make_engine(sysm, dx_grids(numpy_grids(sysm)))

    def run(rng_range):
        return GPUBatchSimulator(sysm.mol1, sysm.mol2, sysm.mob, sysm.pathway_set, sysm.params,
                                 eng, record=True, n_slots=256, traj_range=rng_range).run()

    whole = run(None)
    parts = [run(r) for r in ((0, 180), (180, 320))]
    out = np.concatenate([p.sim_data["outcome"] for p in parts])
    ns = np.concatenate([p.sim_data["n_steps"] for p in parts])
    wo, wn = whole.sim_data["outcome"], whole.sim_data["n_steps"]
    done = (wo == 1) | (wo == 2)
    both = done & ((out == 1) | (out == 2))
    assert both.sum() >= 0.98 * len(wo)
    assert np.array_equal(out[both], wo[both]) and np.array_equal(ns[both], wn[both])
    assert (wo[both] == 1).sum() > 0
