"""Reaction criteria from a BrownDye-style rxns.xml.

Restates pystarc/pipeline/geometry.py:321-389 (flattened pairs, the
batch path's criterion) and :502-620 (per-reaction groups with state
labels): indices in the file are 1-based and converted to 0-based; a pair is
either <atom1>rec charge cutoff</atom1><atom2>lig ...</atom2> (cutoff from
atom1's third field, 5.0 when absent) or <atoms>rec lig</atoms>
<distance>cutoff</distance>; n_needed is read from each criterion (-1 = all
pairs; the flattened reader keeps the last one it sees).
"""
from __future__ import annotations

import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from pathlib import Path
from typing import List, Optional, Tuple

from ..pathways.reaction_interface import PathwaySet, ReactionInterface
from ..structures.molecules import ContactPair, ReactionCriteria


@dataclass
class ReactionPair:
    rec_index: int      # 0-based
    lig_index: int      # 0-based
    cutoff: float


@dataclass
class ReactionGroup:
    name: str
    state_before: Optional[str]
    state_after: Optional[str]
    pairs: List[ReactionPair] = field(default_factory=list)
    n_needed: int = -1


def _pairs_of(crit) -> List[ReactionPair]:
    out = []
    for node in crit.findall("pair"):
        a1, a2 = node.find("atom1"), node.find("atom2")
        if a1 is not None and a2 is not None:
            try:
                p1, p2 = a1.text.split(), a2.text.split()
                out.append(ReactionPair(int(p1[0]) - 1, int(p2[0]) - 1,
                                        float(p1[2]) if len(p1) >= 3 else 5.0))
            except (ValueError, IndexError, AttributeError):
                pass
            continue
        at, dist = node.find("atoms"), node.find("distance")
        if at is not None and dist is not None:
            try:
                idx = at.text.split()
                out.append(ReactionPair(int(idx[0]) - 1, int(idx[1]) - 1, float(dist.text.strip())))
            except (ValueError, IndexError, AttributeError):
                pass
    return out


def _n_needed(crit, default=-1) -> int:
    nn = crit.find("n_needed")
    if nn is not None and nn.text is not None:
        try:
            return int(nn.text.strip())
        except ValueError:
            pass
    return default


def parse_rxns_xml_criteria(path) -> Tuple[List[ReactionPair], int]:
    """All pairs of all reactions flattened, and n_needed."""
    root = ET.parse(str(Path(path))).getroot()
    pairs: List[ReactionPair] = []
    n_needed = -1
    for rxn in root.iter("reaction"):
        crit = rxn.find("criterion")
        if crit is None:
            continue
        n_needed = _n_needed(crit, n_needed)
        pairs.extend(_pairs_of(crit))
    return pairs, n_needed


def parse_rxns_xml_reaction_groups(path) -> List[ReactionGroup]:
    root = ET.parse(str(Path(path))).getroot()
    groups = []
    for i, rxn in enumerate(root.iter("reaction")):
        crit = rxn.find("criterion")
        if crit is None:
            continue

        def txt(tag):
            node = rxn.find(tag)
            return node.text.strip() if node is not None and node.text else None
        pairs = _pairs_of(crit)
        if pairs:
            groups.append(ReactionGroup(txt("n") or f"reaction_{i}", txt("state_before"),
                                        txt("state_after"), pairs, _n_needed(crit)))
    return groups


def pathway_set_from_rxns(path, states=False) -> PathwaySet:
    """states=False: one flattened reaction (all pairs, n_needed), the batch
    step's default criterion.  states=True: one reaction per <reaction> with
    its state_before / state_after labels and first_state, which switches
    the batch step to state-machine reactions."""
    if not states:
        pairs, n_needed = parse_rxns_xml_criteria(path)
        crit = ReactionCriteria(name="rxns", pairs=[ContactPair(p.rec_index, p.lig_index,
                                                                p.cutoff) for p in pairs],
                                n_needed=n_needed)
        return PathwaySet([ReactionInterface("rxns", crit)])
    root = ET.parse(str(Path(path))).getroot()
    fs = root.find("first_state")
    first = fs.text.strip() if fs is not None and fs.text else None
    rxns = []
    for g in parse_rxns_xml_reaction_groups(path):
        crit = ReactionCriteria(name=g.name, pairs=[ContactPair(p.rec_index, p.lig_index, p.cutoff)
                                                    for p in g.pairs], n_needed=g.n_needed)
        rxns.append(ReactionInterface(g.name, crit, state_before=g.state_before,
                                      state_after=g.state_after))
    return PathwaySet(rxns, first_state=first)
