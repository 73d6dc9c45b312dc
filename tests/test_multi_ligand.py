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


def test_receptor_interior_has_no_room_for_a_ligand_centroid():
    """Probe the receptor on a 0.2 A grid: inside 2.5 A below its surface
    every point lies within the 1.5 A overlap radius of some atom, so the
    overlap rule keeps ligand centroids out of the interior."""
    from scipy.spatial import cKDTree
    rec = S.multi_ligand(n_traj=10, n_lig=1)[0].mol1.positions_array()
    rec = rec - rec.mean(axis=0)
    R = np.linalg.norm(rec, axis=1).max()
    ax = np.arange(-R, R + 0.2, 0.2)
    G = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), axis=-1).reshape(-1, 3)
    inner = G[np.linalg.norm(G, axis=1) < R - 2.5]
    d, _ = cKDTree(rec).query(inner)
    assert len(inner) > 1_000_000
    assert d.max() < 1.5
    # and the atoms stay apart
    assert cKDTree(rec).query(rec, k=2)[0][:, 1].min() > 1.9
