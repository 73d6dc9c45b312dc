"""Run control of the production loop on the device: every trajectory's own
max_steps budget with refilled slots, checkpoint / resume equal to the
uninterrupted run, the reference's stall rule, path / energetics snapshots
and live convergence reports.

Reference: pystarc/simulation/gpu_batch_simulator.py:548-575 (resume),
:1342-1356 (99.5 % stop, stall rule), :1318-1340 (snapshots), :753-761 and
:1358-1372 (convergence milestones), :1400-1420 and :1505-1510 (checkpoint).
"""
from types import SimpleNamespace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from pystarc_amd import systems as S  # noqa: E402
from pystarc_amd.constants import CONV  # noqa: E402
from pystarc_amd.simulation.gpu_batch_simulator import GPUBatchSimulator  # noqa: E402
from tests._systems import dx_grids, numpy_grids, small_protein_ligand, two_spheres_small  # noqa: E402


def sim(sysm, eng, **kw):
    return GPUBatchSimulator(sysm.mol1, sysm.mol2, sysm.mob, sysm.pathway_set, sysm.params, eng,
                             **kw)


@pytest.fixture(scope="module")
def spheres():
    sysm = two_spheres_small(n_traj=2000)
    return sysm, S.make_engine(sysm, dx_grids(numpy_grids(sysm)))


def test_every_trajectory_gets_its_own_step_budget(spheres):
    """2,000 trajectories in 256 refilled slots, max_steps 60: each one ends
    (none is left unstarted or cut short by a global step cap), the ones that
    hit the cap ran exactly 60 of their own steps."""
    sysm, eng = spheres
    sysm.params.max_steps = 60
    try:
        r = sim(sysm, eng, n_slots=256, record=True, graph_steps=8).run()
    finally:
        sysm.params.max_steps = 1_000_000
    out, ns = r.sim_data["outcome"], r.sim_data["n_steps"]
    print(f"budget: outcomes {np.bincount(out, minlength=4)}", flush=True)
    assert np.all(out > 0)
    assert np.all(ns <= 60)
    assert np.all(ns[out == 3] == 60)
    assert (out == 3).sum() > 100 and (out == 1).sum() > 0
    assert r.n_max_steps == (out == 3).sum()


def test_stall_rule_stops_early_and_keeps_outcomes(spheres):
    """With a stall check after every step the run stops at the first check
    past 95 % done where fewer than 2 trajectories finished; every trajectory
    finished by then has the fate and step count of the full run."""
    sysm, eng = spheres
    full = sim(sysm, eng, record=True).run()
    st = sim(sysm, eng, record=True, stall_interval=1).run()
    N = sysm.params.n_trajectories
    print(f"stall: {st.sim_data['stop_reason']} at {st.sim_data['bd_steps']} steps, "
          f"{st.n_completed} done; full {full.sim_data['bd_steps']} steps", flush=True)
    assert st.sim_data["stop_reason"] == "stall"
    assert st.n_completed > 0.95 * N and st.n_max_steps > 0
    assert st.sim_data["bd_steps"] < full.sim_data["bd_steps"]
    o1, o2 = full.sim_data["outcome"], st.sim_data["outcome"]
    fin = o2 > 0
    assert np.array_equal(o1[fin], o2[fin])
    assert np.array_equal(full.sim_data["n_steps"][fin], st.sim_data["n_steps"][fin])


def test_resume_from_checkpoint_equals_uninterrupted_run(tmp_path):
    """Interrupted after 150 device steps with checkpoint_interval 40, the run
    resumed from checkpoint.npz ends with every trajectory's outcome, step
    count and observables equal to those of the uninterrupted run; the file
    is removed when the run completes."""
    sysm = small_protein_ligand(n_traj=400)
    eng = S.make_engine(sysm, dx_grids(numpy_grids(sysm)))
    p = sysm.params
    p.checkpoint_interval = 40
    p._output_cfg = SimpleNamespace(save_interval=25, full_paths=True, energetics=True)
    kw = dict(n_slots=128, record="observables", graph_steps=8)
    p._work_dir = str(tmp_path / "straight")
    ref = sim(sysm, eng, **kw).run()
    assert not (tmp_path / "straight" / "checkpoint.npz").exists()
    p._work_dir = str(tmp_path / "resumed")
    part = sim(sysm, eng, **kw).run(max_device_steps=150)
    ck = tmp_path / "resumed" / "checkpoint.npz"
    assert ck.exists() and part.sim_data["bd_steps"] == 150
    assert part.n_completed < ref.n_completed
    res = sim(sysm, eng, **kw).run()
    print(f"resume: ref {ref.n_reacted}/{ref.n_escaped} in {ref.sim_data['bd_steps']} steps, "
          f"resumed {res.n_reacted}/{res.n_escaped} in {res.sim_data['bd_steps']}", flush=True)
    assert res.sim_data["resumed"] and not ck.exists()
    assert (res.n_reacted, res.n_escaped, res.total_steps) == (ref.n_reacted, ref.n_escaped,
                                                               ref.total_steps)
    assert res.sim_data["bd_steps"] == ref.sim_data["bd_steps"]
    for k in ("outcome", "n_steps", "min_dist", "step_at_min", "n_returns", "radial_counts",
              "trans_matrix", "angular_counts", "milestone_flux_out"):
        assert np.array_equal(res.sim_data[k], ref.sim_data[k]), k
    assert len(res.sim_data["path_steps"]) == len(ref.sim_data["path_steps"])
    for a, b in zip(res.sim_data["path_steps"], ref.sim_data["path_steps"]):
        assert np.array_equal(a, b)
    assert [c["milestone"] for c in res.sim_data["convergence"]] == \
        [c["milestone"] for c in ref.sim_data["convergence"]]


def test_snapshots_and_convergence_reports(spheres):
    sysm, eng = spheres
    p = sysm.params
    p._output_cfg = SimpleNamespace(save_interval=7, full_paths=True, energetics=True)
    p.convergence_interval = 20
    try:
        r = sim(sysm, eng, n_slots=512, record="observables", graph_steps=8).run()
    finally:
        del p._output_cfg
        del p.convergence_interval
    paths, en = r.sim_data["path_steps"], r.sim_data["energy_steps"]
    assert len(paths) == len(en) > 10
    for k, (a, b) in enumerate(zip(paths, en)):
        # loop index 0, 7, 14, ... of the reference (after moves 1, 8, 15, ...)
        assert np.all(a[:, 1] == 7 * k) and np.all(b[:, 1] == 7 * k)
        assert np.array_equal(a[:, 0], b[:, 0])
        assert np.all(np.diff(a[:, 0]) > 0)     # trajectory order, as the reference lists them
        assert np.all((a[:, 0] >= 0) & (a[:, 0] < p.n_trajectories))
        r_c = np.linalg.norm(a[:, 2:5], axis=1)
        assert np.all(r_c > 0) and np.all(r_c < 1.2 * p.r_escape)
        assert np.all(np.sum(a[:, 5:8] ** 2, axis=1) <= 1 + 1e-12)
        assert np.all(b[:, 5] > 0)
    conv = r.sim_data["convergence"]
    assert [c["pct"] for c in conv] == [20.0, 40.0, 60.0, 80.0]
    N = p.n_trajectories
    for c in conv:
        assert c["milestone"] <= c["done"] < c["milestone"] + 0.1 * N
        assert c["k_on"] == pytest.approx(CONV * r.k_b * c["P_rxn"])
        assert c["P_rxn"] == pytest.approx(c["n_reacted"] / (c["n_reacted"] + c["n_escaped"]))
