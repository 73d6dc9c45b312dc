"""Reaction pathways as the batch path reads them
(pystarc/pathways/reaction_interface.py:14-91): ``reactions[*].criteria.pairs``
and ``.n_needed``."""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

from ..structures.molecules import ReactionCriteria


@dataclass
class ReactionInterface:
    name: str
    criteria: ReactionCriteria
    probability: float = 1.0
    state_before: Optional[str] = None
    state_after: Optional[str] = None


class PathwaySet:
    def __init__(self, reactions: Optional[List[ReactionInterface]] = None,
                 first_state: Optional[str] = None):
        self.reactions: List[ReactionInterface] = reactions or []
        self.first_state = first_state

    def add(self, rxn: ReactionInterface) -> None:
        self.reactions.append(rxn)

    def __len__(self) -> int:
        return len(self.reactions)
