"""results.json of a batch run, with the fields and units of
pystarc/pipeline/output_writer.py:79-116 (checked by the reference's
tests:4917-4946): k_on with its confidence interval, P_rxn with its Wilson
interval, k_b, D_rel, trajectory counts, r_start / r_escape, wall time,
steps per second, the confidence level and log10(k_on) when k_on > 0."""
from __future__ import annotations

import json
import math
from pathlib import Path


def results_dict(result, D_rel: float, k_b: float, confidence: float = 0.95) -> dict:
    """The results.json mapping; state-machine runs add completed_reactions."""
    k_on = result.rate_constant(D_rel, k_b)
    k_lo, k_hi = result.rate_constant_ci(D_rel, k_b, confidence)
    p_lo, p_hi = result.reaction_probability_ci(confidence)
    d = {"k_on": k_on, "k_on_low": k_lo, "k_on_high": k_hi, "k_on_units": "M-1 s-1",
         "P_rxn": result.reaction_probability, "P_rxn_low": p_lo, "P_rxn_high": p_hi,
         "k_b": k_b, "k_b_units": "A3/ps", "D_rel": D_rel, "D_rel_units": "A2/ps",
         "n_trajectories": result.n_trajectories, "n_reacted": result.n_reacted,
         "n_escaped": result.n_escaped, "n_max_steps": result.n_max_steps,
         "r_start": result.r_start, "r_escape": result.r_escape,
         "wall_time_sec": result.elapsed_sec, "steps_per_sec": result.steps_per_sec,
         "confidence_level": confidence}
    if k_on > 0:
        d["log10_k_on"] = math.log10(k_on)
    sd = getattr(result, "sim_data", None) or {}
    if "completed_reactions" in sd:
        d["completed_reactions"] = sd["completed_reactions"]
    return d


def write_results_json(result, work_dir, D_rel: float, k_b: float,
                       confidence: float = 0.95) -> Path:
    work_dir = Path(work_dir)
    work_dir.mkdir(parents=True, exist_ok=True)
    p = work_dir / "results.json"
    p.write_text(json.dumps(results_dict(result, D_rel, k_b, confidence), indent=2))
    return p
