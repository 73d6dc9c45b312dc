"""The learned surrogate on the full-size protein-small-ligand system
(3,220 / 18 atoms, 161^3 / 65^3 grids): trained on the FP64 grid path's
forces, it stays within the stated tolerance in the step's audit and keeps
the reaction probability inside the plain FP64 path's Wilson 95 % interval
on the same seeds (no additive slack)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle.physics import wilson_ci  # noqa: E402
from pystarc_amd import systems as S  # noqa: E402
from pystarc_amd._lib import Context  # noqa: E402
from pystarc_amd.simulation.gpu_batch_simulator import AUDIT_TOL, GPUBatchSimulator  # noqa: E402
from pystarc_amd.surrogate import build_surrogate  # noqa: E402

N = 40_000


def test_protein_ligand_surrogate_keeps_k_on():
    sysm = S.protein_ligand(n_traj=N, seed=2012)
    ctx = Context(0)
    grids = S.build_grids(sysm, ctx)
    ctx.close()
    e64 = S.make_engine(sysm, grids)
    s = build_surrogate(e64, sysm, n_train=4_000_000, epochs=24, batch=8192, max_rows=N,
                        lr=3e-3, log=print)
    print(f"validation force error {s.val_error:.4f}", flush=True)
    assert s.val_error < AUDIT_TOL
    kw = dict(n_slots=N, graph_steps=8)
    plain = GPUBatchSimulator(sysm.mol1, sysm.mol2, sysm.mob, sysm.pathway_set, sysm.params,
                              e64, precision="fp64", **kw).run()
    e32 = S.make_engine(sysm, grids, precision="fp32")
    surr = GPUBatchSimulator(sysm.mol1, sysm.mol2, sysm.mob, sysm.pathway_set, sysm.params,
                             e32, surrogate=s, **kw).run()
    audit = surr.sim_data["surrogate_audit"]
    lo, hi = wilson_ci(plain.n_reacted, plain.n_escaped, 0.95)
    print(f"plain P={plain.reaction_probability:.5f} [{lo:.5f}, {hi:.5f}], surrogate "
          f"P={surr.reaction_probability:.5f}, audit {audit['rel_rms']:.4f}, "
          f"{surr.steps_per_sec / plain.steps_per_sec:.2f}x", flush=True)
    assert audit["fallback_step"] is None and audit["n"] > 0
    assert audit["rel_rms"] < AUDIT_TOL
    assert plain.n_reacted + plain.n_escaped > 0.99 * N
    assert lo <= surr.reaction_probability <= hi
    s.close()
