"""The synthetic workloads (pystarc_amd/systems.py) against the
configurations they stand for: atom counts, net charges, sizes, grids,
contact criteria and geometry of configs 1-3.  These are the shapes the
bench and the parity tests run, so a drift here would change what is
measured (CPU only; the grids' values are evaluated on the device and are
not built here)."""
import numpy as np
import pytest

from pystarc_amd import systems as S


def net(mol):
    return float(np.sum(mol.charges_array()))


def radius(mol):
    p = mol.positions_array()
    return float(np.linalg.norm(p - p.mean(0), axis=1).max())


def test_two_spheres_config():
    s = S.two_spheres(n_traj=100)
    assert net(s.mol1) == pytest.approx(1.0) and net(s.mol2) == pytest.approx(-1.0)
    assert (s.params.r_start, s.params.r_escape) == (30.0, 60.0)
    assert s.mob.radius1 == 10.0 and s.mob.radius2 == 2.0
    (rxn,) = s.pathway_set.reactions
    assert rxn.criteria.pairs[0].distance_cutoff == pytest.approx(12.5)
    (g,) = s.grid_specs
    assert g.shape == (97, 97, 97) and g.spacing == 0.5 and g.dtype == "float64"
    assert g.debye == pytest.approx(7.86)


def test_protein_ligand_config():
    s = S.protein_ligand(n_traj=100)
    assert len(s.mol1.atoms) == 3220 and len(s.mol2.atoms) == 18
    assert net(s.mol1) == pytest.approx(6.0, abs=1e-9)
    assert net(s.mol2) == pytest.approx(1.0, abs=1e-9)
    # inside the 22 x 20 x 18 A ellipsoid (semi-axes, plus the lattice jitter)
    p = s.mol1.positions_array() - s.mol1.positions_array().mean(0)
    assert np.all(((p / np.array([22.2, 20.2, 18.2])) ** 2).sum(1) <= 1.0 + 1e-9)
    r = s.mol1.radii_array()
    assert r.min() >= 1.4 and r.max() <= 2.0
    (rxn,) = s.pathway_set.reactions
    assert len(rxn.criteria.pairs) == 3 and rxn.criteria.n_needed == 2
    assert all(pr.distance_cutoff == pytest.approx(4.0) for pr in rxn.criteria.pairs)
    assert s.params.r_start == 40.0
    shapes = {g.role: (g.shape, g.spacing, g.dtype) for g in s.grid_specs}
    assert shapes["elec"] == ((161,) * 3, 0.5, "float32")
    assert shapes["born"] == ((161,) * 3, 0.5, "float32")
    assert shapes["lig_born"] == ((65,) * 3, 0.4, "float32")


def test_protein_protein_config():
    s = S.protein_protein(n_traj=100)
    assert len(s.mol1.atoms) == 1700 and len(s.mol2.atoms) == 1400
    assert net(s.mol1) == pytest.approx(2.0, abs=1e-9)
    assert net(s.mol2) == pytest.approx(-6.0, abs=1e-9)
    assert radius(s.mol1) <= 17.0 + 0.5 and radius(s.mol2) <= 15.0 + 0.5
    assert s.mob.radius1 == 17.0 and s.mob.radius2 == 15.0
    (rxn,) = s.pathway_set.reactions
    assert len(rxn.criteria.pairs) == 5 and rxn.criteria.n_needed == 2
    assert all(pr.distance_cutoff == pytest.approx(5.5) for pr in rxn.criteria.pairs)
    assert s.params.r_start == 60.0 and s.params.r_escape == 120.0
    assert s.params.hydrodynamic_interactions
    roles = sorted(g.role for g in s.grid_specs)
    assert roles == ["born", "elec", "lig_born"]          # bidirectional Born
    assert all(g.shape == (193,) * 3 and g.spacing == 0.5 and g.dtype == "float32"
               for g in s.grid_specs)
    # the lattice keeps atoms apart (jittered cubic lattice, spacing ~2.2 A)
    s1, s2 = s.meta["lattice_spacing"]
    assert 1.8 < s1 < 2.6 and 1.8 < s2 < 2.6


def test_builders_are_seeded():
    a = S.protein_protein(n_traj=10)
    b = S.protein_protein(n_traj=10)
    assert np.array_equal(a.mol2.positions_array(), b.mol2.positions_array())
    assert np.array_equal(a.mol1.charges_array(), b.mol1.charges_array())
