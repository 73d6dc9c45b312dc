"""Config-5 generator: one receptor, a batch of ligands of uneven size."""
import numpy as np

from pystarc_amd import systems as S


def test_multi_ligand_spec():
    ls = S.multi_ligand(n_traj=1000)
    assert len(ls) == 16
    assert all(s.mol1 is ls[0].mol1 for s in ls)
    assert len(ls[0].mol1.atoms) == 4700
    atoms = [s.meta["atoms"] for s in ls]
    assert min(atoms) >= 12 and max(atoms) <= 64 and len(set(atoms)) > 4
    for s in ls:
        q = s.mol2.charges_array()
        assert abs(q.sum() - s.meta["net"]) < 1e-9 and s.meta["net"] in (-2, -1, 0, 1, 2)
        assert len(s.pathway_set.reactions[0].criteria.pairs) == 3
        assert s.pathway_set.reactions[0].criteria.n_needed == 2
        lig = [g for g in s.grid_specs if g.role == "lig_born"]
        assert len(lig) == 1 and lig[0].shape == (65, 65, 65) and lig[0].spacing == 0.4
    rec = [g for g in ls[0].grid_specs if g.role in ("elec", "born")]
    assert [g.shape for g in rec] == [(257, 257, 257)] * 2
    assert all(len(s.grid_specs) == 1 for s in ls[1:])
    # deterministic from the seeds
    again = S.multi_ligand(n_traj=1000)
    assert [s.meta["atoms"] for s in again] == atoms
    np.testing.assert_array_equal(again[3].mol2.positions_array(), ls[3].mol2.positions_array())
