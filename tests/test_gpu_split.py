"""Split force evaluation of the production step (gather modes "split" and
"split_zpair"): the terms on the ligand atoms over the grid-path list, then
the reverse Born over the same trajectories re-sorted by the ligand-frame
direction of the receptor (k_rev_sort), applied to the first pass's F and T.

With the same lanes per trajectory the two passes perform the fused
kernel's operations in the same order, so a production run under the split
evaluation is the direct kernel's run trajectory for trajectory (FP64: every
outcome, step count and final pose identical; FP32: the same, or within FP32
rounding of a term where the compiler fuses a multiply-add differently in
the part kernels).  A system without a ligand Born grid has no reverse pass:
the request runs the direct kernel.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from pystarc_amd import systems as S  # noqa: E402
from pystarc_amd.simulation.gpu_batch_simulator import GPUBatchSimulator  # noqa: E402
from tests._systems import (dx_grids, numpy_grids, small_protein_ligand,  # noqa: E402
                            small_protein_protein, two_spheres_small)


def _system(name):
    if name == "protein_ligand":
        return small_protein_ligand(n_traj=1500)
    return small_protein_protein(n_traj=1500)


def _run(sysm, eng, gather, hydro, steps=None):
    sim = GPUBatchSimulator(sysm.mol1, sysm.mol2, sysm.mob, sysm.pathway_set, sysm.params, eng,
                            record=True, gather=gather, hydro=hydro, n_slots=1500)
    if steps is None:
        return sim.run(), None
    drv = sim.production_driver(record=True)
    try:
        drv.advance(steps)
        return drv.outcomes(), drv.state()
    finally:
        drv.close()


@pytest.mark.parametrize("name,hydro", [("protein_ligand", "scalar"), ("protein_protein", "rpy12")])
@pytest.mark.parametrize("torque", ["reference", "physical"])
@pytest.mark.parametrize("G", [16, 64])
def test_split_production_identical_to_direct_fp64(name, hydro, torque, G):
    sysm = _system(name)
    sysm.params.max_steps = 20000
    eng = S.make_engine(sysm, dx_grids(numpy_grids(sysm)), torque_mode=torque, force_group=G)
    a, _ = _run(sysm, eng, "direct", hydro)
    assert eng.last_eval() == ("direct", G)
    b, _ = _run(sysm, eng, "split", hydro)
    assert eng.last_eval() == ("split", G)
    assert (a.n_reacted, a.n_escaped, a.total_steps) == (b.n_reacted, b.n_escaped, b.total_steps)
    assert np.array_equal(a.sim_data["outcome"], b.sim_data["outcome"])
    assert np.array_equal(a.sim_data["n_steps"], b.sim_data["n_steps"])
    assert a.n_reacted > 0 and a.n_escaped > 0


@pytest.mark.parametrize("name,hydro", [("protein_ligand", "scalar"), ("protein_protein", "rpy12")])
@pytest.mark.parametrize("torque", ["reference", "physical"])
@pytest.mark.parametrize("G", [16, 64])
def test_split_zpair_steps_match_zpair_fp32(name, hydro, torque, G):
    """40 production steps from the same start under the z-pair kernel and
    the split z-pair evaluation: poses agree to FP32 force rounding."""
    sysm = _system(name)
    eng = S.make_engine(sysm, dx_grids(numpy_grids(sysm)), precision="fp32", torque_mode=torque,
                        force_group=G)
    (oa, _), (pa, qa, sa, na) = _run(sysm, eng, "zpair", hydro, steps=40)
    assert eng.last_eval() == ("zpair", G)
    (ob, _), (pb, qb, sb, nb) = _run(sysm, eng, "split_zpair", hydro, steps=40)
    assert eng.last_eval() == ("split_zpair", G)
    same = np.array_equal(pa, pb) and np.array_equal(qa, qb)
    dp = float(np.abs(pa - pb).max())
    dq = float(np.abs(qa - qb).max())
    print(f"{name} {torque} G={G}: identical {same}, max |dpos| {dp:.3e} A, max |dq| {dq:.3e}",
          flush=True)
    assert np.array_equal(sa, sb) and np.array_equal(na, nb)
    assert dp < 1e-5 and dq < 1e-6


def test_split_without_ligand_born_grid_runs_direct():
    sysm = two_spheres_small(n_traj=500)
    sysm.params.max_steps = 5000
    eng = S.make_engine(sysm, dx_grids(numpy_grids(sysm)))
    sim = GPUBatchSimulator(sysm.mol1, sysm.mol2, sysm.mob, sysm.pathway_set, sysm.params, eng,
                            gather="split", n_slots=500)
    sim.run()
    assert eng.last_eval()[0] == "direct"
