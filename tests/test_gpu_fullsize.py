"""Force parity at the workloads' full sizes: the protein-ligand system
(3,220 / 18 atoms, 161^3 and 65^3 grids), the protein pair (1,700 / 1,400
atoms, three 193^3 grids) and the large receptor with 16 ligands (4,700 atoms
on 257^3 grids, every ligand through the grouped force kernel in one launch).
Grids are evaluated on the device, copied to the host and handed to the
FP64 oracle; 64 poses per system (near contact to beyond the grids) are
compared per force, torque (both torque definitions) and energy.

Bars (as tests/test_gpu_forces.py): FP64 within 1e-12 relative of the
oracle plus 1e-3 of the per-atom magnitude sum; FP32 within 1e-5 plus 5e-2 of
it.  The measured FP32 errors are printed per system and quantity.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import forces as OF  # noqa: E402
from oracle.physics import quats_to_rotmats, shoemake  # noqa: E402
from pystarc_amd import systems as S  # noqa: E402
from pystarc_amd._lib import Context  # noqa: E402
from pystarc_amd.forces.gpu_batch_engine import forces_pose_group  # noqa: E402
from tests._systems import oracle_engine  # noqa: E402
from tests.test_gpu_forces import assert_close  # noqa: E402

BARS = {"fp64": (1e-12, 1e-3), "fp32": (1e-5, 5e-2)}


def _np(grids):
    return {k: [(np.asarray(g.data), np.asarray(g.origin), float(np.asarray(g.delta)[0, 0]))
                for g in v] for k, v in grids.items()}


def _poses(rng, n, rmin, rmax):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return d * rng.uniform(rmin, rmax, size=(n, 1)), shoemake(rng.random((n, 3)))


def _oracle(system, g, c, q, torque_mode):
    eng = oracle_engine(system, g, chunk=0)
    body = system.mol2.positions_array() - system.mol2.centroid()
    R = quats_to_rotmats(q)
    lab = OF.place_ligand(R, c, body)
    lq = system.mol2.charges_array()
    F, T, E = OF.engine_forces(eng, lab, lq, R, c, torque_mode)
    _, _, Fa, Ea = OF.eval_batch(lab, lq, eng["elec"], 0.0, False, eng["ff"], per_atom=True)
    _, _, Fb, Eb = OF.eval_batch(lab, lq, eng["born"], eng["alpha"], True, eng["ff"], per_atom=True)
    sf = np.linalg.norm(Fa, axis=2).sum(1) + np.linalg.norm(Fb, axis=2).sum(1)
    se = np.abs(Ea).sum(1) + np.abs(Eb).sum(1)
    # lever arm scale of the torque: the force scale times the ligand radius
    st = sf * (np.linalg.norm(body, axis=1).max() + 1.0)
    return F, T, E, sf, st, se


def _errors(name, got, ref, scale):
    err = np.linalg.norm(np.atleast_2d((got - ref).T).T.reshape(len(ref), -1), axis=1)
    mag = np.linalg.norm(np.asarray(ref).reshape(len(ref), -1), axis=1)
    big = mag > 1e-3 * max(mag.max(), 1e-300)
    rel = err[big] / mag[big] if big.any() else np.zeros(1)
    print(f"  {name}: max rel {rel.max():.2e}, rms rel {np.sqrt((rel ** 2).mean()):.2e}, "
          f"max err / per-atom scale {np.max(err / (scale + 1e-300)):.2e}", flush=True)


def _check(label, eng_prec, system, g, F, T, E, c, q, torque_mode):
    F0, T0, E0, sf, st, se = _oracle(system, g, c, q, torque_mode)
    rel, frac = BARS[eng_prec]
    print(f"{label} {eng_prec} torque={torque_mode}:", flush=True)
    _errors("F", F, F0, sf)
    _errors("E", E[:, None], E0[:, None], se)
    assert_close(F, F0, sf, rel, frac, f"{label} F")
    assert_close(E[:, None], E0[:, None], se, rel, frac, f"{label} E")
    if torque_mode == "physical":
        _errors("T", T, T0, st)
        assert_close(T, T0, st, rel, frac, f"{label} T")
    else:
        assert np.all(T == 0.0)


def _system_grids(system):
    ctx = Context(0)
    grids = S.build_grids(system, ctx)
    ctx.close()
    return grids


@pytest.fixture(scope="module")
def c2():
    s = S.protein_ligand(n_traj=100)
    return s, _system_grids(s)


@pytest.fixture(scope="module")
def c3():
    s = S.protein_protein(n_traj=100)
    return s, _system_grids(s)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("torque_mode", ["reference", "physical"])
def test_protein_ligand_full_size(c2, prec, torque_mode):
    system, grids = c2
    g = _np(grids)
    eng = S.make_engine(system, grids, precision=prec, torque_mode=torque_mode)
    eng.set_ligand(system.mol2.positions_array() - system.mol2.centroid(),
                   system.mol2.charges_array())
    rng = np.random.default_rng(21)
    c, q = _poses(rng, 64, 24.0, 60.0)
    F, T, E = (x.cpu().numpy() for x in eng.forces_pose(c, q))
    _check("protein-ligand", prec, system, g, F, T, E, c, q, torque_mode)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("torque_mode", ["reference", "physical"])
def test_protein_pair_full_size(c3, prec, torque_mode):
    system, grids = c3
    g = _np(grids)
    eng = S.make_engine(system, grids, precision=prec, torque_mode=torque_mode)
    eng.set_ligand(system.mol2.positions_array() - system.mol2.centroid(),
                   system.mol2.charges_array())
    rng = np.random.default_rng(31)
    c, q = _poses(rng, 64, 34.0, 90.0)
    F, T, E = (x.cpu().numpy() for x in eng.forces_pose(c, q))
    _check("protein-pair", prec, system, g, F, T, E, c, q, torque_mode)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_multi_ligand_grouped_kernel_full_size(prec):
    """All 16 ligands of the 4,700-atom receptor through one grouped launch
    (pst_forces_pose_group) against the oracle, ligand by ligand."""
    ls = S.multi_ligand(n_traj=100)
    ctx = Context(0)
    grids = [S.build_grids(s, ctx) for s in ls]
    ctx.close()
    e0 = S.make_engine(ls[0], grids[0], precision=prec, torque_mode="physical")
    engines = [e0] + [S.make_engine(s, gr, precision=prec, torque_mode="physical", receptor_from=e0)
                      for s, gr in zip(ls[1:], grids[1:])]
    rng = np.random.default_rng(41)
    cs, qs = [], []
    for s, e in zip(ls, engines):
        e.set_ligand(s.mol2.positions_array() - s.mol2.centroid(), s.mol2.charges_array())
        c, q = _poses(rng, 64, 26.0, 75.0)
        cs.append(c)
        qs.append(q)
    outs = forces_pose_group(engines, cs, qs)
    rec = _np(grids[0])
    for k, (s, gr, (F, T, E)) in enumerate(zip(ls, grids, outs)):
        g = {"elec": rec["elec"], "born": rec["born"], "lig_born": _np(gr)["lig_born"]}
        _check(f"multi-ligand {k} ({s.meta['atoms']} atoms)", prec, s, g, F.cpu().numpy(),
               T.cpu().numpy(), E.cpu().numpy(), cs[k], qs[k], "physical")


def test_protein_pair_rpy12_full_size_production():
    """The protein pair at full size with the two-body RPY 12x12 move: a
    production run of 20,000 trajectories keeps poses finite and unit
    quaternions, and ends (the 99.5 % stop) with both outcomes."""
    from pystarc_amd.simulation.gpu_batch_simulator import GPUBatchSimulator
    system = S.protein_protein(n_traj=20000)
    grids = _system_grids(system)
    eng = S.make_engine(system, grids, precision="fp32")
    sim = GPUBatchSimulator(system.mol1, system.mol2, system.mob, system.pathway_set,
                            system.params, eng, hydro="rpy12", n_slots=20000)
    drv = sim.production_driver(record=True)
    try:
        k = drv.advance(3000, graph_steps=10)
        pos, quat, st, ns = drv.state()
        out, steps = drv.outcomes()
        ctr = drv.counters()
    finally:
        drv.close()
    print(f"rpy12 full size: {k} steps, reacted {ctr['n_reacted']}, escaped {ctr['n_escaped']}, "
          f"running {(out == 0).sum()}", flush=True)
    assert ctr["stopped"] and k < 3000
    assert ctr["n_reacted"] + ctr["n_escaped"] >= int(0.995 * 20000)
    run = st == 0
    assert np.all(np.isfinite(pos[run])) and np.all(np.isfinite(quat[run]))
    np.testing.assert_allclose(np.linalg.norm(quat[run], axis=1), 1.0, atol=1e-12)
    assert ctr["n_reacted"] > 0 and ctr["n_escaped"] > 0
    assert np.all(steps[out > 0] > 0)
