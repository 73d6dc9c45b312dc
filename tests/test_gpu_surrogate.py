"""The learned surrogate in the production step on the protein-protein
system: its force error against the FP64 grid path (its reference), the audit
inside the step, k_on against the plain path's Wilson interval, and the
throughput it gains over the plain path."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle.physics import wilson_ci  # noqa: E402
from pystarc_amd import systems as S  # noqa: E402
from pystarc_amd._lib import Context  # noqa: E402
from pystarc_amd.simulation.gpu_batch_simulator import GPUBatchSimulator  # noqa: E402
from pystarc_amd.surrogate import build_surrogate  # noqa: E402


@pytest.fixture(scope="module")
def pp():
    sysm = S.protein_protein(n_traj=30000, seed=3013)
    ctx = Context(0)
    grids = S.build_grids(sysm, ctx)
    ctx.close()
    return sysm, grids


@pytest.fixture(scope="module")
def trained(pp):
    sysm, grids = pp
    eng = S.make_engine(sysm, grids)
    # a surrogate whose in-step force error (~6-8 %) is what the bench's
    # training reaches; a loosely fitted one (~17 %) shifts k_on by ~2 standard
    # errors of the 30,000-trajectory estimate
    s = build_surrogate(eng, sysm, n_train=4_000_000, epochs=24, batch=8192, max_rows=32768,
                        log=print, lr=3e-3)
    return eng, s


AUDIT_TOL = 0.25   # stated force tolerance: relative RMS against the FP64 grid path


def test_surrogate_force_error(trained):
    eng, s = trained
    assert s.val_error < AUDIT_TOL, s.val_error


def test_surrogate_k_on_and_gain(pp, trained):
    sysm, grids = pp
    eng, s = trained
    kw = dict(n_slots=30000, graph_steps=8)
    t0 = time.time()
    plain = GPUBatchSimulator(sysm.mol1, sysm.mol2, sysm.mob, sysm.pathway_set, sysm.params,
                              eng, precision="fp64", **kw).run()
    t1 = time.time()
    eng32 = S.make_engine(sysm, grids, precision="fp32")
    surr = GPUBatchSimulator(sysm.mol1, sysm.mol2, sysm.mob, sysm.pathway_set, sysm.params,
                             eng32, surrogate=s, audit_tol=AUDIT_TOL, **kw).run()
    t2 = time.time()
    audit = surr.sim_data["surrogate_audit"]
    print(f"plain fp64 {plain.steps_per_sec:.3e} traj-steps/s P={plain.reaction_probability:.4f}; "
          f"surrogate+fp32 {surr.steps_per_sec:.3e} P={surr.reaction_probability:.4f}; "
          f"audit rel rms {audit['rel_rms']:.3e}")
    assert audit["fallback_step"] is None
    assert audit["rel_rms"] < AUDIT_TOL
    lo, hi = wilson_ci(plain.n_reacted, plain.n_escaped, 0.95)
    assert lo <= surr.reaction_probability <= hi
    assert surr.steps_per_sec > plain.steps_per_sec


def test_surrogate_torque_audited_in_physical_mode():
    """With the physical torque the surrogate applies its predicted torque;
    the in-step audit then compares torques too (pst_prod_audit), and both
    the validation and the audited torque errors are stated."""
    from tests._systems import dx_grids, numpy_grids, small_protein_ligand
    sysm = small_protein_ligand(n_traj=4000)
    g = dx_grids(numpy_grids(sysm))
    e64 = S.make_engine(sysm, g, torque_mode="physical")
    s = build_surrogate(e64, sysm, n_train=400_000, epochs=10, batch=4096, max_rows=4096,
                        shell=4.0)
    assert s.use_torque and s.val_torque_error is not None
    e32 = S.make_engine(sysm, g, precision="fp32", torque_mode="physical")
    res = GPUBatchSimulator(sysm.mol1, sysm.mol2, sysm.mob, sysm.pathway_set, sysm.params, e32,
                            surrogate=s, n_slots=4000).run()
    a = res.sim_data["surrogate_audit"]
    print(f"physical torque: validation force {s.val_error:.3e}, torque {s.val_torque_error:.3e}; "
          f"audit force {a['rel_rms']:.3e}, torque {a['torque_rel_rms']:.3e}", flush=True)
    assert a["n"] > 0 and a["torque_rel_rms"] is not None
    assert a["fallback_step"] is None
    assert a["torque_rel_rms"] < AUDIT_TOL and s.val_torque_error < AUDIT_TOL
