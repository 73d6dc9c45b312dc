"""GPU force operator (libpystarc_hip through the C ABI) against the FP64
oracle on identical inputs.

Tolerances (stated per the precision contract):
  FP64 path: |F_gpu - F_oracle| <= 1e-12 * (|F_oracle| + 1e-3 * sum_i |f_i|)
  FP32 path: |F_gpu - F_oracle| <= 1e-5  * (|F_oracle| + 5e-2 * sum_i |f_i|)
where sum_i |f_i| is the sum of per-atom force magnitudes (the scale of the
cancellation in a net force of a near-neutral molecule).  The FP32 floor,
5e-7 * sum_i |f_i|, is a few FP32 ulps of the per-atom terms: FP32 per-atom
arithmetic cannot resolve a net force below it whatever the accumulation.
Energies use the same form with per-atom |e_i|.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import forces as OF  # noqa: E402
from oracle import grids as OG  # noqa: E402
from oracle.physics import quats_to_rotmats, shoemake  # noqa: E402
from pystarc_amd import _lib  # noqa: E402
from pystarc_amd import systems as S  # noqa: E402
from tests._systems import (dx_grids, numpy_grids, oracle_engine, small_protein_ligand,  # noqa: E402
                            small_protein_protein, two_spheres_small)


def poses(n, rng, rmin, rmax):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c = d * rng.uniform(rmin, rmax, size=(n, 1))
    q = shoemake(rng.random((n, 3)))
    return c, q


def engine_for(system, g, **kw):
    return S.make_engine(system, dx_grids(g), **kw)


def oracle_forces(system, g, c, q, torque_mode="reference", chunk=None):
    eng = oracle_engine(system, g, chunk=chunk)
    body = system.mol2.positions_array() - system.mol2.centroid()
    R = quats_to_rotmats(q)
    lab = OF.place_ligand(R, c, body)
    F, T, E = OF.engine_forces(eng, lab, system.mol2.charges_array(), R, c, torque_mode)
    # per-atom magnitude scale for tolerances
    f1, e1, Fa, Ea = OF.eval_batch(lab, system.mol2.charges_array(), eng["elec"], 0.0, False,
                                   eng["ff"], per_atom=True)
    f2, e2, Fb, Eb = OF.eval_batch(lab, system.mol2.charges_array(), eng["born"], eng["alpha"], True,
                                   eng["ff"], per_atom=True)
    scale_f = np.linalg.norm(Fa, axis=2).sum(1) + np.linalg.norm(Fb, axis=2).sum(1)
    scale_e = np.abs(Ea).sum(1) + np.abs(Eb).sum(1)
    return F, T, E, lab, R, scale_f, scale_e


def assert_close(got, ref, scale, rel, frac, what):
    err = np.linalg.norm(np.atleast_2d((got - ref).T).T.reshape(len(ref), -1), axis=1)
    mag = np.linalg.norm(np.asarray(ref).reshape(len(ref), -1), axis=1)
    tol = rel * (mag + frac * scale) + 1e-300
    bad = np.where(err > tol)[0]
    assert len(bad) == 0, f"{what}: {len(bad)} bad, worst {np.max(err / tol):.3g}x tol; idx {bad[:5]}"


def test_grid_generator_matches_numpy():
    sysm = small_protein_ligand()
    ctx = _lib.Context(0)
    gd = S.build_grids(sysm, ctx)
    gn = numpy_grids(sysm)
    for role in ("elec", "born", "lig_born"):
        for a, (b, _, _) in zip(gd[role], gn[role]):
            ref = b.astype(np.float64)
            assert np.allclose(a.data.astype(np.float64), ref, rtol=2e-6,
                               atol=1e-6 * np.abs(ref).max()), role


@pytest.mark.parametrize("prec,rel,frac", [("fp64", 1e-12, 1e-3), ("fp32", 1e-5, 5e-2)])
def test_positions_mode_protein_ligand(prec, rel, frac):
    sysm = small_protein_ligand()
    g = numpy_grids(sysm)
    eng = engine_for(sysm, g, precision=prec)
    rng = np.random.default_rng(11)
    c, q = poses(300, rng, 6.0, 26.0)
    F0, T0, E0, lab, R, sf, se = oracle_forces(sysm, g, c, q)
    F, T, E = eng(lab, sysm.mol2.charges_array(), R_matrices=R, centroids=c)
    F, T, E = F.cpu().numpy(), T.cpu().numpy(), E.cpu().numpy()
    assert_close(F, F0, sf, rel, frac, f"F {prec}")
    assert_close(E[:, None], E0[:, None], se, rel, frac, f"E {prec}")
    assert np.all(T == 0)


def test_pose_mode_equals_positions_mode():
    sysm = small_protein_ligand()
    g = numpy_grids(sysm)
    eng = engine_for(sysm, g)
    body = sysm.mol2.positions_array() - sysm.mol2.centroid()
    eng.set_ligand(body, sysm.mol2.charges_array())
    rng = np.random.default_rng(12)
    c, q = poses(200, rng, 6.0, 26.0)
    F0, T0, E0, lab, R, sf, se = oracle_forces(sysm, g, c, q, chunk=0)
    Fp, Tp, Ep = eng.forces_pose(c, q)
    # the exact per-trajectory cull never drops a contributing receptor atom,
    # so it equals the oracle without any skip
    assert_close(Fp.cpu().numpy(), F0, sf, 1e-12, 1e-3, "pose F")
    assert_close(Ep.cpu().numpy()[:, None], E0[:, None], se, 1e-12, 1e-3, "pose E")


def test_reference_chunk_skip():
    sysm = small_protein_ligand()
    g = numpy_grids(sysm)
    eng = engine_for(sysm, g)
    rng = np.random.default_rng(13)
    c, q = poses(64, rng, 20.0, 30.0)   # all far: the reference skips the chunk
    F0, _, _, lab, R, sf, _ = oracle_forces(sysm, g, c, q)
    F, _, _ = eng(lab, sysm.mol2.charges_array(), R_matrices=R, centroids=c)
    assert_close(F.cpu().numpy(), F0, sf, 1e-12, 1e-3, "chunk skip")


def test_physical_torque_matches_oracle():
    sysm = small_protein_ligand()
    g = numpy_grids(sysm)
    eng = engine_for(sysm, g, torque_mode="physical")
    rng = np.random.default_rng(14)
    c, q = poses(100, rng, 6.0, 16.0)
    F0, T0, E0, lab, R, sf, _ = oracle_forces(sysm, g, c, q, torque_mode="physical", chunk=0)
    F, T, E = eng(lab, sysm.mol2.charges_array(), R_matrices=R, centroids=c, rev_skip="exact")
    lever = np.linalg.norm(lab - c[:, None, :], axis=2).max(1)
    assert_close(T.cpu().numpy(), T0, sf * lever, 1e-12, 1e-3, "torque")
    assert_close(F.cpu().numpy(), F0, sf, 1e-12, 1e-3, "F physical")


def test_two_spheres_grid_and_yukawa():
    sysm = two_spheres_small()
    g = numpy_grids(sysm)
    eng = engine_for(sysm, g)
    rng = np.random.default_rng(15)
    c, q = poses(500, rng, 12.5, 60.0)   # both inside the grid and in the far-field
    F0, T0, E0, lab, R, sf, se = oracle_forces(sysm, g, c, q)
    F, T, E = eng(lab, sysm.mol2.charges_array(), R_matrices=R, centroids=c)
    assert_close(F.cpu().numpy(), F0, sf, 1e-12, 1e-3, "two spheres F")
    assert_close(E.cpu().numpy()[:, None], E0[:, None], se, 1e-12, 1e-3, "two spheres E")


@pytest.mark.parametrize("prec,rel,frac", [("fp64", 1e-12, 1e-3), ("fp32", 1e-5, 5e-2)])
def test_protein_protein_bidirectional_born(prec, rel, frac):
    sysm = small_protein_protein()
    g = numpy_grids(sysm)
    eng = engine_for(sysm, g, precision=prec)
    rng = np.random.default_rng(16)
    c, q = poses(128, rng, 8.0, 22.0)
    F0, T0, E0, lab, R, sf, se = oracle_forces(sysm, g, c, q)
    F, T, E = eng(lab, sysm.mol2.charges_array(), R_matrices=R, centroids=c)
    # reverse-Born per-atom scale joins the tolerance
    eng_o = oracle_engine(sysm, g)
    _, fa = OF.born_reverse(R, c, eng_o["rec_pos"], eng_o["rec_q"], eng_o["lig_born"][0],
                            eng_o["alpha"], per_atom=True)
    sf = sf + np.linalg.norm(fa, axis=2).sum(1)
    assert_close(F.cpu().numpy(), F0, sf, rel, frac, f"pp F {prec}")


@pytest.mark.parametrize("prec,rel,frac,lanes", [("fp64", 1e-12, 1e-3, 16), ("fp64", 1e-12, 1e-3, 64),
                                                 ("fp32", 1e-5, 5e-2, 32)])
def test_pose_mode_spatially_ordered_ligand(prec, rel, frac, lanes):
    """A 120-atom ligand is held on the device in spatial order: pose-mode
    forces, lever-arm torques and energies still equal the oracle's sums in
    input order (and every lane count walks the same set of atoms)."""
    sysm = small_protein_protein()
    assert len(sysm.mol2.charges_array()) > 64
    g = numpy_grids(sysm)
    eng = engine_for(sysm, g, precision=prec, torque_mode="physical", force_group=lanes)
    body = sysm.mol2.positions_array() - sysm.mol2.centroid()
    eng.set_ligand(body, sysm.mol2.charges_array())
    rng = np.random.default_rng(18)
    c, q = poses(150, rng, 8.0, 22.0)
    F0, T0, E0, lab, R, sf, se = oracle_forces(sysm, g, c, q, torque_mode="physical", chunk=0)
    eng_o = oracle_engine(sysm, g)
    _, fa = OF.born_reverse(R, c, eng_o["rec_pos"], eng_o["rec_q"], eng_o["lig_born"][0],
                            eng_o["alpha"], per_atom=True)
    sf = sf + np.linalg.norm(fa, axis=2).sum(1)
    Fp, Tp, Ep = eng.forces_pose(c, q)
    lever = np.linalg.norm(lab - c[:, None, :], axis=2).max(1)
    assert_close(Fp.cpu().numpy(), F0, sf, rel, frac, f"ordered F {prec}")
    assert_close(Tp.cpu().numpy(), T0, sf * lever, rel, frac, f"ordered T {prec}")
    assert_close(Ep.cpu().numpy()[:, None], E0[:, None], se, rel, frac, f"ordered E {prec}")


@pytest.mark.parametrize("lanes", [1, 4, 16, 64])
def test_force_group_sizes_agree(lanes):
    sysm = small_protein_ligand()
    g = numpy_grids(sysm)
    eng = engine_for(sysm, g, force_group=lanes)
    rng = np.random.default_rng(17)
    c, q = poses(77, rng, 6.0, 26.0)
    F0, _, E0, lab, R, sf, se = oracle_forces(sysm, g, c, q)
    F, _, E = eng(lab, sysm.mol2.charges_array(), R_matrices=R, centroids=c)
    assert_close(F.cpu().numpy(), F0, sf, 1e-12, 1e-3, f"G={lanes}")


def test_bad_arguments_raise():
    sysm = small_protein_ligand()
    g = numpy_grids(sysm)
    eng = engine_for(sysm, g)
    with pytest.raises(RuntimeError):
        eng.ctx.check(eng.ctx.lib.pst_set_force_group(eng.ctx.h, 3), "pst_set_force_group")
    with pytest.raises(ValueError):
        eng(np.zeros((4, 3)), np.zeros(3))
