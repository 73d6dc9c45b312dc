"""BD run parameters read by the batch simulator (fields and defaults of
pystarc/simulation/nam_simulator.py:79-106, including r_escape = 1.1 b when
left at 0)."""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional


@dataclass
class NAMParameters:
    n_trajectories: int = 1_000
    dt: float = 0.2
    dt_rxn: float = 0.05
    minimum_core_dt: float = 0.0
    minimum_core_reaction_dt: float = 0.0
    max_steps: int = 1_000_000
    r_start: float = 100.0
    r_escape: float = 0.0
    seed: Optional[int] = None
    n_threads: int = 1
    use_hard_sphere: bool = True
    hydrodynamic_interactions: bool = False
    verbose: bool = False

    def __post_init__(self):
        if self.r_escape == 0.0:
            self.r_escape = self.r_start * 1.1
