"""Input / output formats either side of the batch step: PQR, rxns.xml and
results.json.  Fixtures in tests/golden/formats are data taken from the
reference's examples (two charged spheres, barnase-barstar, thrombin-
thrombomodulin), excerpted; expected values are read off the files."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from pystarc_amd.pipeline.geometry import (parse_rxns_xml_criteria, parse_rxns_xml_reaction_groups,
                                           pathway_set_from_rxns)
from pystarc_amd.pipeline.output_writer import results_dict, write_results_json
from pystarc_amd.simulation.gpu_batch_simulator import GPUBatchResult
from pystarc_amd.structures.molecules import Atom, Molecule
from pystarc_amd.structures.pqr_io import parse_pqr, parse_pqr_records, write_pqr

F = Path(__file__).parent / "golden" / "formats"


def test_two_spheres_pqr():
    rec = parse_pqr(F / "two_spheres_receptor.pqr")
    lig = parse_pqr(F / "two_spheres_ligand.pqr")
    assert len(rec.atoms) == 1 and len(lig.atoms) == 1
    assert rec.total_charge() == 1.0 and lig.total_charge() == -1.0
    assert rec.atoms[0].radius == 1.0 and rec.atoms[0].residue_name == "FOO"
    np.testing.assert_array_equal(rec.centroid(), [0.0, 0.0, 0.0])


def test_barnase_pqr_columns():
    recs = parse_pqr_records(F / "barnase_receptor_head.pqr")
    assert len(recs) == 28
    a = recs[0]
    assert (a.serial, a.name, a.resname, a.resid) == (1, "N", "VAL", 1)
    assert (a.x, a.y, a.z, a.charge, a.radius, a.element) == (16.783, 48.812, 26.447, 0.0577, 1.55, "N")
    lig = parse_pqr(F / "barnase_ligand_head.pqr")
    assert lig.atoms[0].name == "N" and lig.atoms[0].residue_index == 109
    assert lig.atoms[-1].name == "OXT" and lig.atoms[-1].charge == -0.8132


def test_pqr_whitespace_fallback_and_skips(tmp_path):
    p = tmp_path / "mixed.pqr"
    p.write_text("REMARK   generated\n"
                 "ATOM      1  U   FOO    1     0.0  0.0  0.0   1.0 1.0\n"
                 "HETATM    2  C1  LIGA A   7  1.5 -2.25 3.0 -0.5 1.7 C\n"
                 "ATOM      3  O   WAT   12    4.0   5.0   6.0  -0.8  1.4\n"
                 "TER\n"
                 "END\n")
    r = parse_pqr_records(p)
    assert [x.record_type for x in r] == ["ATOM", "HETATM", "ATOM"]
    assert (r[0].chain, r[0].resid, r[0].charge, r[0].radius) == ("", 1, 1.0, 1.0)
    assert (r[1].chain, r[1].resid, r[1].x, r[1].y, r[1].element) == ("A", 7, 1.5, -2.25, "C")
    assert (r[2].resname, r[2].resid, r[2].z, r[2].charge) == ("WAT", 12, 6.0, -0.8)


def test_pqr_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    mol = Molecule(name="m", atoms=[Atom(index=i, name="CA", residue_name="ALA", residue_index=i + 1,
                                         chain="B", x=float(x), y=float(y), z=float(z),
                                         charge=float(q), radius=float(r))
                                    for i, (x, y, z, q, r) in enumerate(np.round(np.column_stack(
                                        [rng.uniform(-99, 99, (50, 3)), rng.uniform(-1, 1, 50),
                                         rng.uniform(1, 2, 50)]), 3))])
    write_pqr(mol, tmp_path / "m.pqr")
    back = parse_pqr(tmp_path / "m.pqr")
    np.testing.assert_allclose(back.positions_array(), mol.positions_array(), atol=5e-4)
    np.testing.assert_allclose(back.charges_array(), mol.charges_array(), atol=5e-5)
    assert [a.chain for a in back.atoms] == ["B"] * 50


def test_rxns_two_spheres_one_based():
    pairs, nn = parse_rxns_xml_criteria(F / "two_spheres_rxns.xml")
    assert nn == 1
    assert [(p.rec_index, p.lig_index, p.cutoff) for p in pairs] == [(0, 0, 2.5)]
    ps = pathway_set_from_rxns(F / "two_spheres_rxns.xml")
    c = ps.reactions[0].criteria
    assert c.n_needed == 1 and c.pairs[0].mol1_atom_index == 0 and c.pairs[0].distance_cutoff == 2.5


def test_rxns_thrombin_and_barnase():
    pairs, nn = parse_rxns_xml_criteria(F / "thrombin_rxns.xml")
    assert nn == 3 and len(pairs) == 21
    assert (pairs[0].rec_index, pairs[0].lig_index, pairs[0].cutoff) == (1763, 1196, 5.0)
    groups = parse_rxns_xml_reaction_groups(F / "thrombin_rxns.xml")
    assert len(groups) == 1 and groups[0].state_before == "before" and groups[0].name == "association"
    pairs, nn = parse_rxns_xml_criteria(F / "barnase_rxns.xml")
    assert nn == 1
    assert [(p.rec_index, p.lig_index, p.cutoff) for p in pairs] == [(1227, 632, 7.0), (847, 1211, 7.0)]


def test_rxns_atom_format(tmp_path):
    p = tmp_path / "r.xml"
    p.write_text("<top><reactions><reaction><criterion><n_needed>2</n_needed>"
                 "<pair><atom1>10 0.5 6.5</atom1><atom2>3 -0.5</atom2></pair>"
                 "<pair><atom1>11</atom1><atom2>4</atom2></pair>"
                 "</criterion></reaction></reactions></top>")
    pairs, nn = parse_rxns_xml_criteria(p)
    assert nn == 2
    assert [(q.rec_index, q.lig_index, q.cutoff) for q in pairs] == [(9, 2, 6.5), (10, 3, 5.0)]


def _result():
    return GPUBatchResult(n_trajectories=1000, n_reacted=400, n_escaped=590, n_max_steps=10,
                          reaction_counts={"rxn": 400}, r_start=30.0, r_escape=60.0, dt=0.2,
                          elapsed_sec=2.0, steps_per_sec=5e6, total_steps=10_000_000, k_b=57.47)


def test_results_json_fields(tmp_path):
    """Fields of tests:4917-4946 and output_writer.py:79-116."""
    p = write_results_json(_result(), tmp_path, D_rel=0.434, k_b=57.47)
    d = json.loads(p.read_text())
    for key in ("k_on", "k_on_low", "k_on_high", "P_rxn", "P_rxn_low", "P_rxn_high", "k_b",
                "D_rel", "n_trajectories", "n_reacted", "n_escaped", "wall_time_sec",
                "steps_per_sec", "log10_k_on", "confidence_level", "k_on_units"):
        assert key in d, key
    P = 400 / 990
    assert d["P_rxn"] == pytest.approx(P)
    assert d["k_on"] == pytest.approx(6.022e8 * 57.47 * P)
    assert d["k_on_low"] < d["k_on"] < d["k_on_high"]
    assert d["log10_k_on"] == pytest.approx(math.log10(d["k_on"]))
    assert d["n_max_steps"] == 10 and d["r_escape"] == 60.0


def test_results_dict_zero_reactions():
    r = _result()
    r.n_reacted = 0
    d = results_dict(r, 0.434, 57.47)
    assert d["k_on"] == 0.0 and "log10_k_on" not in d
