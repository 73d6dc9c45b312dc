"""Splitting one run over several GPUs and combining the parts.

Two ways to split, both producing the reference's combined results.json:

* by trajectory range (this build's production mode): every part runs the
  same seed on a disjoint range of trajectory ids
  (``GPUBatchSimulator(..., traj_range=(first, count))``).  Random numbers are
  keyed by trajectory id, so each trajectory of a part is the one the whole
  run would have simulated; the parts together are the single run.
* by seed, as pystarc/multi_GPU/multi_GPU_runs.py:71-84 does: part i runs
  n_trajectories // n_splits trajectories with seed + i * 11111111.

``combine_results`` pools finished parts as
pystarc/multi_GPU/combine_data.py:69-140 does (summed counts, one Wilson
interval over the pooled trials, summed per-reaction firing counts, summed
wall time); ``combine_sim_data`` sums the run-wide histograms and
concatenates the per-trajectory records; ``allreduce_counts`` is the same
pooling over a torch.distributed process group (one rank per GPU), whose
collective carries only a few counters (CPU tensors, so gloo works too).
"""
from __future__ import annotations

import math
from typing import Dict, List, Sequence, Tuple

import numpy as np

CONV = 6.022e8          # A^3/ps -> M^-1 s^-1 (combine_data.py:79)
SEED_STRIDE = 11111111  # multi_GPU_runs.py:81


def split_trajectories(n_total: int, n_splits: int) -> List[Tuple[int, int]]:
    """Disjoint (first, count) ranges covering 0 .. n_total-1, sizes within one."""
    if n_splits <= 0 or n_total < n_splits:
        raise ValueError("need 1 <= n_splits <= n_total")
    edges = [i * n_total // n_splits for i in range(n_splits + 1)]
    return [(edges[i], edges[i + 1] - edges[i]) for i in range(n_splits)]


def split_seeds(n_total: int, seed: int, n_splits: int) -> List[Tuple[int, int]]:
    """(n_trajectories, seed) of each part of a seed split, as
    multi_GPU_runs.py:71-84: the trajectory count is divided down and part
    i (counted from 1) uses seed + i * 11111111."""
    if n_splits <= 0:
        raise ValueError("n_splits must be positive")
    per = n_total // n_splits
    return [(per, seed + i * SEED_STRIDE) for i in range(1, n_splits + 1)]


def _wilson(P: float, N: int, z: float = 1.96) -> Tuple[float, float]:
    if N <= 0:
        return 0.0, 0.0
    denom = 1.0 + z * z / N
    centre = (P + z * z / (2 * N)) / denom
    spread = z * math.sqrt(P * (1 - P) / N + z * z / (4 * N * N)) / denom
    return max(0.0, centre - spread), min(1.0, centre + spread)


def _as_dict(r) -> Dict:
    if isinstance(r, dict):
        return r
    return {"n_reacted": r.n_reacted, "n_escaped": r.n_escaped, "n_max_steps": r.n_max_steps,
            "k_b": r.k_b, "r_start": r.r_start, "r_escape": r.r_escape,
            "wall_time_sec": r.elapsed_sec, "steps_per_sec": r.steps_per_sec,
            "completed_reactions": (r.sim_data or {}).get("completed_reactions", [])}


def combine_results(parts: Sequence, D_rel: float = 0.0) -> Dict:
    """Pooled results of finished parts (GPUBatchResult objects or results.json
    dicts).  P = reacted / (reacted + escaped) over all parts; k_on and its
    Wilson 95 % interval from P and k_b of the first part; max-step
    trajectories are counted in n_trajectories but not in P
    (combine_data.py:69-88, :124-140, :146-175)."""
    runs = [_as_dict(r) for r in parts]
    if not runs:
        raise ValueError("no parts to combine")
    nr = sum(int(r["n_reacted"]) for r in runs)
    ne = sum(int(r["n_escaped"]) for r in runs)
    nm = sum(int(r.get("n_max_steps", 0)) for r in runs)
    N = nr + ne
    P = nr / N if N > 0 else 0.0
    k_b = float(runs[0]["k_b"])
    lo, hi = _wilson(P, N)
    wall = sum(float(r.get("wall_time_sec", 0.0)) for r in runs)
    steps = sum(float(r.get("steps_per_sec", 0.0)) * float(r.get("wall_time_sec", 0.0))
                for r in runs)
    k_on = CONV * k_b * P
    out = {"k_on": k_on, "k_on_low": CONV * k_b * lo, "k_on_high": CONV * k_b * hi,
           "k_on_units": "M-1 s-1", "P_rxn": P, "P_rxn_low": lo, "P_rxn_high": hi,
           "k_b": k_b, "k_b_units": "A3/ps", "D_rel": float(runs[0].get("D_rel", D_rel)),
           "D_rel_units": "A2/ps", "n_trajectories": N + nm, "n_reacted": nr, "n_escaped": ne,
           "n_max_steps": nm, "r_start": runs[0].get("r_start", 0.0),
           "r_escape": runs[0].get("r_escape", 0.0), "wall_time_sec": wall,
           "steps_per_sec": steps / wall if wall > 0 else 0.0, "confidence_level": 0.95,
           "log10_k_on": math.log10(k_on) if k_on > 0 else 0.0}
    per, order = {}, []
    for r in runs:
        for e in r.get("completed_reactions", []) or []:
            name = e.get("name")
            if name is None:
                continue
            if name not in per:
                per[name] = {"name": name, "n_fired": 0, "state_before": e.get("state_before"),
                             "state_after": e.get("state_after")}
                order.append(name)
            per[name]["n_fired"] += int(e.get("n_fired", 0))
    if per:
        out["completed_reactions"] = [per[n] for n in order]
    return out


# sim_data entries that are run-wide counts (summed) and per-trajectory
# records (concatenated in part order, so a trajectory's index in the
# combined arrays is its part's offset plus its index in the part)
_SUMMED = ("radial_counts", "angular_counts", "milestone_flux_out", "milestone_flux_in",
           "trans_matrix", "contact_pair_counts", "contact_total_steps",
           "n_returns_total", "n_overlap_rejected", "n_bridge", "total_steps")
_RECORDS = ("outcome", "n_steps", "start_pos", "start_q", "min_dist", "step_at_min",
            "total_time_ps", "n_returns", "bb_triggered")
# per-event records whose trajectory column is re-indexed by the part offset
_EVENTS = (("encounter_traj", ("encounter_pos", "encounter_q", "encounter_n_pairs")),
           ("near_miss_traj", ("near_miss_pos", "near_miss_q", "near_miss_dist")))
# bin edges / radii shared by every part
_SHARED = ("radial_bins", "angular_theta", "angular_phi", "milestone_radii", "trans_bins")


def combine_sim_data(parts: Sequence) -> Dict:
    """Run-wide counts summed and per-trajectory records concatenated over
    parts (combine_data.py:176-233 does the same on the written files)."""
    sds = [p.sim_data if hasattr(p, "sim_data") else p for p in parts]
    out: Dict = {}
    for k in _SUMMED:
        vals = [sd[k] for sd in sds if k in sd]
        if vals:
            out[k] = sum(np.asarray(v) for v in vals) if not np.isscalar(vals[0]) else sum(vals)
    for k in _RECORDS:
        vals = [np.asarray(sd[k]) for sd in sds if k in sd]
        if vals and len(vals) == len(sds):
            out[k] = np.concatenate(vals, axis=0)
    offsets = np.cumsum([0] + [len(sd["outcome"]) if "outcome" in sd else 0 for sd in sds])
    for tk, cols in _EVENTS:
        if all(tk in sd for sd in sds):
            out[tk] = np.concatenate([np.asarray(sd[tk]) + offsets[i] for i, sd in enumerate(sds)])
            for c in cols:
                if all(c in sd for sd in sds):
                    out[c] = np.concatenate([np.asarray(sd[c]) for sd in sds], axis=0)
    for k in _SHARED:
        if k in sds[0]:
            out[k] = sds[0][k]
    return out


def allreduce_counts(result, group=None) -> Dict:
    """Pool one part per rank over a torch.distributed group: one all-reduce
    (sum) of the counters, identical pooled results on every rank.  The
    counters travel as a CPU float64 tensor (exact for counts below 2^53)."""
    import torch
    import torch.distributed as dist
    d = _as_dict(result)
    v = torch.tensor([float(d["n_reacted"]), float(d["n_escaped"]),
                      float(d.get("n_max_steps", 0)), float(d.get("wall_time_sec", 0.0)),
                      float(d.get("steps_per_sec", 0.0)) * float(d.get("wall_time_sec", 0.0))],
                     dtype=torch.float64)
    dist.all_reduce(v, op=dist.ReduceOp.SUM, group=group)
    nr, ne, nm, wall, steps = (float(x) for x in v.tolist())
    pooled = {"n_reacted": int(round(nr)), "n_escaped": int(round(ne)),
              "n_max_steps": int(round(nm)), "k_b": d["k_b"], "r_start": d.get("r_start", 0.0),
              "r_escape": d.get("r_escape", 0.0), "wall_time_sec": wall,
              "steps_per_sec": steps / wall if wall > 0 else 0.0}
    return combine_results([pooled])
