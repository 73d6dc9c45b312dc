"""Rigid-body data containers read by the batch BD path.

Same attribute and method names as pystarc/structures/molecules.py
(Atom :14-46, Molecule.centroid/total_charge/positions_array/charges_array
:59-93, ContactPair :184-199, ReactionCriteria :203-237), so molecules built
for the reference drive this package unchanged.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np


@dataclass
class Atom:
    index: int = 0
    name: str = ""
    residue_name: str = ""
    residue_index: int = 0
    chain: str = "A"
    x: float = 0.0
    y: float = 0.0
    z: float = 0.0
    charge: float = 0.0
    radius: float = 1.5

    @property
    def position(self) -> np.ndarray:
        return np.array([self.x, self.y, self.z], dtype=float)


@dataclass
class Molecule:
    name: str = ""
    atoms: List[Atom] = field(default_factory=list)

    @classmethod
    def from_arrays(cls, positions, charges, radii=None, name=""):
        positions = np.asarray(positions, dtype=float)
        charges = np.asarray(charges, dtype=float)
        radii = np.full(len(charges), 1.5) if radii is None else np.asarray(radii, dtype=float)
        atoms = [Atom(index=i, name=f"A{i}", x=float(p[0]), y=float(p[1]), z=float(p[2]),
                      charge=float(q), radius=float(r))
                 for i, (p, q, r) in enumerate(zip(positions, charges, radii))]
        return cls(name=name, atoms=atoms)

    def positions_array(self) -> np.ndarray:
        return np.array([[a.x, a.y, a.z] for a in self.atoms], dtype=float).reshape(-1, 3)

    def charges_array(self) -> np.ndarray:
        return np.array([a.charge for a in self.atoms], dtype=float)

    def radii_array(self) -> np.ndarray:
        return np.array([a.radius for a in self.atoms], dtype=float)

    def centroid(self) -> np.ndarray:
        """Unweighted mean of atom positions (molecules.py:59-63)."""
        if not self.atoms:
            return np.zeros(3)
        return self.positions_array().mean(axis=0)

    def total_charge(self) -> float:
        return float(sum(a.charge for a in self.atoms))

    def bounding_radius(self) -> float:
        """Largest distance from the centroid to an atom's surface (molecules.py:79-84)."""
        if not self.atoms:
            return 0.0
        d = np.linalg.norm(self.positions_array() - self.centroid(), axis=1)
        return float((d + self.radii_array()).max())


@dataclass
class ContactPair:
    mol1_atom_index: int = 0
    mol2_atom_index: int = 0
    distance_cutoff: float = 5.0


@dataclass
class ReactionCriteria:
    """Fires when at least n_needed pairs are closer than their cutoffs
    (-1: all pairs), molecules.py:203-237."""
    name: str = "reaction"
    pairs: List[ContactPair] = field(default_factory=list)
    n_needed: int = -1
    state_before: Optional[str] = None
    state_after: Optional[str] = None
