"""Brick-binned force evaluation (count / scan / scatter / LDS-tile gather /
reduce) against the direct kernel and the FP64 oracle.

The binned pipeline forms every per-atom term with the direct kernel's own
arithmetic and sums them in the same lane order, so for the same lanes per
trajectory its FP64 forces, torques and energies equal the direct kernel's
bit for bit (FP32: to FP32 rounding, the compiler may fuse a term's
multiply-adds differently in the two kernels); FP64 trajectories of the
production step are then identical too.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle.physics import shoemake  # noqa: E402
from pystarc_amd import systems as S  # noqa: E402
from pystarc_amd.simulation.gpu_batch_simulator import GPUBatchSimulator  # noqa: E402
from tests._systems import (dx_grids, numpy_grids, small_protein_ligand,  # noqa: E402
                            small_protein_protein, two_spheres_small)
from tests.test_gpu_forces import assert_close, oracle_forces  # noqa: E402


def _poses(n, rng, rmin, rmax):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return d * rng.uniform(rmin, rmax, size=(n, 1)), shoemake(rng.random((n, 3)))


def _system(name):
    if name == "protein_ligand":
        return small_protein_ligand(), (5.0, 26.0)
    return small_protein_protein(), (8.0, 30.0)


@pytest.mark.parametrize("name", ["protein_ligand", "protein_protein"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("torque", ["reference", "physical"])
@pytest.mark.parametrize("G", [16, 64])
def test_binned_equals_direct_bit_for_bit(name, prec, torque, G):
    sysm, (rmin, rmax) = _system(name)
    g = numpy_grids(sysm)
    eng = S.make_engine(sysm, dx_grids(g), precision=prec, torque_mode=torque, force_group=G)
    eng.set_ligand(sysm.mol2.positions_array() - sysm.mol2.centroid(), sysm.mol2.charges_array())
    rng = np.random.default_rng(31)
    c, q = _poses(777, rng, rmin, rmax)
    eng.set_gather("direct")
    F0, T0, E0 = eng.forces_pose(c, q)
    assert eng.last_eval() == ("direct", G)
    eng.set_gather("binned", max_traj=1024)
    F1, T1, E1 = eng.forces_pose(c, q)
    assert eng.last_eval() == ("binned", G)
    if prec == "fp64":
        assert torch.equal(F0, F1) and torch.equal(T0, T1) and torch.equal(E0, E1)
    else:
        # FP32 terms: the compiler may fuse the multiply-adds of one term
        # differently in the two kernels; the results agree to FP32 rounding
        for a, b in ((F0, F1), (T0, T1), (E0, E1)):
            scale = float(a.abs().max()) + 1e-30
            assert float((a - b).abs().max()) <= 1e-6 * scale
    assert float(F0.abs().sum()) > 0


def test_binned_matches_oracle_fp64():
    sysm, (rmin, rmax) = _system("protein_protein")
    g = numpy_grids(sysm)
    eng = S.make_engine(sysm, dx_grids(g), precision="fp64")
    eng.set_ligand(sysm.mol2.positions_array() - sysm.mol2.centroid(), sysm.mol2.charges_array())
    eng.set_gather("binned", max_traj=512)
    rng = np.random.default_rng(32)
    c, q = _poses(300, rng, rmin, rmax)
    F, T, E = eng.forces_pose(c, q)
    assert eng.last_eval()[0] == "binned"
    F0, T0, E0, lab, R, sf, se = oracle_forces(sysm, g, c, q, chunk=0)
    assert_close(F.cpu().numpy(), F0, sf, 1e-12, 1e-3, "binned F")
    assert_close(E.cpu().numpy()[:, None], E0[:, None], se, 1e-12, 1e-3, "binned E")


def test_binned_request_falls_back_without_pair_grid():
    """Two spheres carry one FP64 grid: no float2 pair, the direct kernel runs."""
    sysm = two_spheres_small(n_traj=10)
    eng = S.make_engine(sysm, dx_grids(numpy_grids(sysm)))
    eng.set_ligand(sysm.mol2.positions_array() - sysm.mol2.centroid(), sysm.mol2.charges_array())
    eng.set_gather("binned", max_traj=64)
    c, q = _poses(50, np.random.default_rng(3), 12.0, 40.0)
    eng.forces_pose(c, q)
    assert eng.last_eval()[0] == "direct"


@pytest.mark.parametrize("name,hydro", [("protein_ligand", "scalar"), ("protein_protein", "rpy12")])
def test_binned_production_identical_to_direct(name, hydro):
    sysm, _ = _system(name)
    sysm.params.n_trajectories = 1500
    sysm.params.max_steps = 20000
    eng = S.make_engine(sysm, dx_grids(numpy_grids(sysm)))
    res = {}
    for mode in ("direct", "binned"):
        sim = GPUBatchSimulator(sysm.mol1, sysm.mol2, sysm.mob, sysm.pathway_set, sysm.params, eng,
                                record=True, gather=mode, hydro=hydro, n_slots=512)
        res[mode] = sim.run()
        assert eng.last_eval()[0] == mode
    a, b = res["direct"], res["binned"]
    assert (a.n_reacted, a.n_escaped, a.total_steps) == (b.n_reacted, b.n_escaped, b.total_steps)
    assert np.array_equal(a.sim_data["outcome"], b.sim_data["outcome"])
    assert np.array_equal(a.sim_data["n_steps"], b.sim_data["n_steps"])


@pytest.mark.parametrize("name", ["protein_ligand", "protein_protein"])
@pytest.mark.parametrize("torque", ["reference", "physical"])
@pytest.mark.parametrize("G", [16, 64])
def test_zpair_equals_direct(name, torque, G):
    """The z-pair launch fetches the same node values with wider loads: the
    FP32 forces equal the direct kernel's to FP32 rounding, the FP64 forces
    to FP64 rounding."""
    sysm, (rmin, rmax) = _system(name)
    g = numpy_grids(sysm)
    eng = S.make_engine(sysm, dx_grids(g), precision="fp32", torque_mode=torque, force_group=G)
    eng.set_ligand(sysm.mol2.positions_array() - sysm.mol2.centroid(), sysm.mol2.charges_array())
    c, q = _poses(777, np.random.default_rng(41), rmin, rmax)
    eng.set_gather("direct")
    F0, T0, E0 = eng.forces_pose(c, q)
    eng.set_gather("zpair")
    F1, T1, E1 = eng.forces_pose(c, q)
    assert eng.last_eval() == ("zpair", G)
    for a, b in ((F0, F1), (T0, T1), (E0, E1)):
        scale = float(a.abs().max()) + 1e-30
        assert float((a - b).abs().max()) <= 1e-6 * scale
    assert float(F0.abs().sum()) > 0
    e64 = S.make_engine(sysm, dx_grids(g), precision="fp64", torque_mode=torque, force_group=G)
    e64.set_ligand(sysm.mol2.positions_array() - sysm.mol2.centroid(), sysm.mol2.charges_array())
    e64.set_gather("direct")
    F0, T0, E0 = e64.forces_pose(c, q)
    e64.set_gather("zpair")
    F1, T1, E1 = e64.forces_pose(c, q)
    assert e64.last_eval() == ("zpair", G)
    for a, b in ((F0, F1), (T0, T1), (E0, E1)):
        scale = float(a.abs().max()) + 1e-30
        assert float((a - b).abs().max()) <= 1e-13 * scale
