"""Drop-in GPUBatchSimulator / GPUBatchResult on the gfx950 step kernels.

Mirrors pystarc/simulation/gpu_batch_simulator.py: constructor
``GPUBatchSimulator(mol1, mol2, mob, pathway_set, params, batch_engine)``
(:159-318), ``run() -> GPUBatchResult`` (:462-1594) and the result dataclass
with its Wilson interval and k_on (:45-141).  The parameters are read from
``params`` exactly as the reference reads them, including the attributes the
pipeline attaches with setattr (max_dt, _overlap_check, _kT_scale,
gpu_force_batch; pystarc/pipeline/pipeline.py:346-353).

Keyword-only extensions (defaults keep the reference's physics):
  rng          "philox" (default): device counter-based noise, persistent
               slots refilled from the trajectory counter, hipGraph-captured
               steps, no host synchronisation inside a launch.
               "numpy": replay mode - the host draws from
               numpy.random.default_rng(seed) in the reference's order and
               shapes and the kernels consume them, reproducing the reference's
               reaction / escape decisions.
  n_slots      concurrent trajectories in philox mode (default min(N, 2^20)).
  precision    force precision "fp64" / "fp32" (default: the engine's).
  graph_steps  steps per captured hipGraph launch in philox mode.
  record       True: keep per-trajectory outcome and step counts (sim_data);
               "observables": also the reference's per-trajectory and run-wide
               sim_data (start pose, closest approach, simulated time,
               returns, bridge reactions, encounter / near-miss poses,
               radial / angular / milestone / transition / contact counts),
               collected on the device (philox mode).
  surrogate    a trained pystarc_amd.surrogate.Surrogate: trajectories beyond
               its contact shell take the MLP force (philox mode only).
  audit_n      surrogate trajectories re-evaluated on the grid path per step.
  audit_tol    relative RMS force (or torque) error of a window of at least
               AUDIT_MIN_POSES audited poses above which the run falls back
               to the grid path for every trajectory.
  autotune     False (default) or the number of warm-up steps (True: 64)
               after which the production force launch (direct or
               brick-binned evaluation, lanes per trajectory, workgroup cap,
               XCD mapping) is tuned on live steps.
  gather       None (the engine's) or a GATHER_MODES name ("direct", "binned", "zpair",
               "split", "split_zpair"): force evaluation
               of the production step (identical results either way).
  hydro        "scalar" (default): the reference's scalar relative diffusion
               (D_t, or D_parallel(r) with hydrodynamic_interactions).
               "rpy12": both bodies translate and turn with the two-sphere
               12x12 RPY grand mobility (radii mob.radius1 / mob.radius2,
               kT/eta = 1/0.243 A^3/ps as the reference's HI constant) and its
               Cholesky noise; the pose is then re-expressed in the turned
               receptor frame (philox mode only).
  traj_range   (first, count): run trajectories first .. first+count-1 of
               the params' run (philox mode).  Each draws the random numbers
               it draws in the whole run, so shards of one run (one per GPU,
               pystarc_amd.pipeline.multi_gpu) reproduce its trajectories;
               sim_data records are indexed from 0 within the shard.
  outer        "return_prob" (default): the reference's escape test at
               r_escape with the LMZ return probability.  "lmz": a trajectory
               reaching 1.1 b is carried through the outer region (up to the
               q-sphere at 20 x the larger molecular radius) by the LMZ outer
               propagator of the reference's CPU simulator on the device,
               returning to the b-sphere or escaping (philox mode).
  snapshots    path / energetics rows of the running trajectories every
               _output_cfg.save_interval steps (sim_data "path_steps",
               "energy_steps", gpu_batch_simulator.py:1318-1340); default: on
               with record="observables", as the reference collects them
               with its sim_data.
  stall_interval  steps between the reference's stall checks (10,000).

Run control of philox mode, as the reference's loop (:1342-1420): the 99.5 %
early stop (on the device), the stall rule, live k_on at every
convergence_interval percent (sim_data "convergence", logged), a checkpoint
of the whole device state in _work_dir/checkpoint.npz every
checkpoint_interval completed trajectories, resumed from when the file
exists and removed when the run completes.  Each trajectory gets max_steps
of its own steps however many slots refill it.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import logging
import math
import os
import time
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict

import numpy as np

from .._lib import MAX_PAIRS, StepParams, ptr
from ..forces.gpu_batch_engine import GATHER_MODES
from ..constants import CONV, EPS0, MU_REF, OVERLAP_THRESHOLD, QB_FACTOR
from ..lib.numerical import romberg_integrate

log = logging.getLogger("pystarc_amd")

# the reference checks for stalled runs every 10,000 loop steps
# (gpu_batch_simulator.py:1349-1356)
STALL_EVERY = 10_000
# stated relative RMS force error of the surrogate: above it (audit per
# window of at least AUDIT_MIN_POSES audited poses) the grid path takes over
# every trajectory
AUDIT_TOL = 0.25
# fewest audited poses behind one fallback decision.  A chunk of steps in the
# run's tail, when few trajectories are left beyond the surrogate radius,
# audits only a handful of poses, whose relative RMS error scatters far
# around the surrogate's real error; such chunks are pooled until the window
# holds this many.
AUDIT_MIN_POSES = 2048


def _z_quantile(p):
    lo, hi = -10.0, 10.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if 0.5 * math.erfc(-mid / math.sqrt(2.0)) < p:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


@dataclass
class GPUBatchResult:
    n_trajectories: int
    n_reacted: int
    n_escaped: int
    n_max_steps: int
    reaction_counts: Dict[str, int]
    r_start: float
    r_escape: float
    dt: float
    elapsed_sec: float
    steps_per_sec: float
    sim_data: Dict = None
    total_steps: int = 0
    k_b: float = 0.0
    return_prob: float = 0.0

    @property
    def n_completed(self) -> int:
        return self.n_reacted + self.n_escaped

    @property
    def reaction_probability(self) -> float:
        n = self.n_completed
        return self.n_reacted / n if n > 0 else 0.0

    def reaction_probability_ci(self, confidence: float = 0.95):
        """Wilson score interval (gpu_batch_simulator.py:70-80)."""
        n = self.n_completed
        if n == 0:
            return (0.0, 1.0)
        z = _z_quantile(0.5 + confidence / 2.0)
        p = self.reaction_probability
        a = z * z / n
        centre = (p + a / 2.0) / (1.0 + a)
        half = z * math.sqrt(max(p * (1 - p) / n + a / (4.0 * n), 0.0)) / (1.0 + a)
        return (max(0.0, centre - half), min(1.0, centre + half))

    def _k_of(self, P, D_rel, k_b):
        if P <= 0.0:
            return 0.0
        if k_b > 0.0:
            return CONV * k_b * P
        k_D = 4.0 * math.pi * D_rel * self.r_start
        beta = self.r_start / self.r_escape
        return CONV * k_D * P / (1.0 - P * (1.0 - beta))

    def rate_constant(self, D_rel: float, k_b: float = 0.0) -> float:
        """k_on = CONV k_b P, Smoluchowski form when k_b = 0 (:82-96)."""
        return self._k_of(self.reaction_probability, D_rel, k_b)

    def rate_constant_ci(self, D_rel: float, k_b: float = 0.0, confidence: float = 0.95):
        lo, hi = self.reaction_probability_ci(confidence)
        return (self._k_of(lo, D_rel, k_b), self._k_of(hi, D_rel, k_b))

    def summary(self, D_rel: float, k_b: float = 0.0, confidence: float = 0.95) -> str:
        k = self.rate_constant(D_rel, k_b)
        k_lo, k_hi = self.rate_constant_ci(D_rel, k_b, confidence)
        p_lo, p_hi = self.reaction_probability_ci(confidence)
        pct = int(confidence * 100)
        lines = [f"  Trajectories : {self.n_trajectories:,}",
                 f"  Completed    : {self.n_completed:,}  ({self.n_reacted:,} reacted + "
                 f"{self.n_escaped:,} escaped)"]
        if self.n_max_steps:
            lines.append(f"  Max-steps    : {self.n_max_steps:,}  (excluded)")
        lines += [f"  P_rxn        : {self.reaction_probability:.6f}  ({pct}% CI: [{p_lo:.6f}, {p_hi:.6f}])",
                  f"  k_on         : {k:.4e} M-1 s-1",
                  f"  {pct}% CI     : [{k_lo:.4e}, {k_hi:.4e}] M-1 s-1"]
        if k > 0:
            lines.append(f"  log10(k_on)  : {math.log10(k):.3f}")
        lines += [f"  Wall time    : {self.elapsed_sec:.1f} s",
                  f"  BD steps/sec : {self.steps_per_sec:,.0f}",
                  "  Backend      : HIP gfx950 (batch GPU)"]
        return "\n".join(lines)


def state_machine(pathway_set):
    """State-machine tables of the reference (gpu_batch_simulator.py:216-306):
    reactions with a pair, their state labels; first_state (or the first
    reaction's state_before) is state 0, other labels are numbered in order
    of appearance; a state that is no reaction's state_before is terminal.
    None when no reaction carries a state_before label."""
    rxns = [r for r in pathway_set.reactions if r.criteria.pairs]
    if not any(getattr(r, "state_before", None) is not None for r in pathway_set.reactions):
        return None
    if not rxns:
        return None
    if len(rxns) > MAX_PAIRS:
        raise ValueError(f"at most {MAX_PAIRS} reactions are supported")
    first = getattr(pathway_set, "first_state", None) or rxns[0].state_before
    idx, labels = {}, []

    def get(label):
        if label is None:
            return -1
        if label not in idx:
            idx[label] = len(labels)
            labels.append(label)
        return idx[label]
    if first is not None:
        get(first)
    before = [get(r.state_before) for r in rxns]
    after = [get(r.state_after) for r in rxns]
    if len(labels) > 32:
        raise ValueError("at most 32 reaction states are supported")
    sources = set(before)
    return {"names": [r.name for r in rxns], "before": before, "after": after,
            "labels": labels, "terminal": [i not in sources for i in range(len(labels))],
            "before_labels": [r.state_before for r in rxns],
            "after_labels": [r.state_after for r in rxns]}


def sm_report(lib, ctx, h, sm):
    """Per-reaction fire counts as the reference reports them (:1558-1581)."""
    c = np.zeros(3 * MAX_PAIRS, dtype=np.uint64)
    ctx.check(lib.pst_prod_sm_counts(h, ptr(c)), "pst_prod_sm_counts")
    rows = []
    for r, name in enumerate(sm["names"]):
        d, b, u = int(c[r]), int(c[MAX_PAIRS + r]), int(c[2 * MAX_PAIRS + r])
        rows.append({"name": name, "n_fired": d + b, "state_before": sm["before_labels"][r],
                     "state_after": sm["after_labels"][r], "n_fired_discrete": d,
                     "n_fired_bridge": b, "unique_trajectories": u})
    return {"completed_reactions": rows}


def observables(lib, ctx, h, P, outcome):
    """sim_data entries of the reference (gpu_batch_simulator.py:1524-1557)
    from the device observables of a record-level-2 run."""
    td, hl = C.c_int(), C.c_int()
    lib.pst_prod_obs_sizes(C.byref(td), C.byref(hl))
    N = int(P.n_total)
    tr = np.zeros((N, td.value))
    hist = np.zeros(hl.value, dtype=np.uint64)
    ctx.check(lib.pst_prod_observables(h, ptr(tr), ptr(hist)), "pst_prod_observables")
    r_b, r_esc = float(P.r_b), float(P.r_esc)
    nrad, nth, nph, nms, ntr = 200, 36, 72, 11, 50
    o = 0
    rad = hist[o:o + nrad].astype(np.int64); o += nrad
    ang = hist[o:o + nth * nph].astype(np.int64).reshape(nth, nph); o += nth * nph
    ms_out = hist[o:o + nms].astype(np.int64); o += nms
    ms_in = hist[o:o + nms].astype(np.int64); o += nms
    trans = hist[o:o + ntr * ntr].astype(np.int64).reshape(ntr, ntr); o += ntr * ntr
    contact = hist[o:o + 16].astype(np.int64); o += 16
    contact_steps = int(hist[o])
    reacted = outcome == 1
    escaped = outcome == 2
    nm_pos, nm_q = tr[:, 9:12], tr[:, 12:16]
    return {
        "start_pos": tr[:, 0:3].copy(), "start_q": tr[:, 3:7].copy(),
        "min_dist": tr[:, 7].copy(), "step_at_min": tr[:, 8].astype(np.int64),
        "total_time_ps": tr[:, 16].copy(), "n_returns": tr[:, 17].astype(np.int64),
        "bb_triggered": tr[:, 18].astype(np.int64),
        "encounter_pos": tr[reacted, 19:22].copy(), "encounter_q": tr[reacted, 22:26].copy(),
        "encounter_traj": np.where(reacted)[0].astype(np.int64),
        "encounter_n_pairs": tr[reacted, 26].astype(np.int64),
        "near_miss_pos": nm_pos[escaped].copy(), "near_miss_q": nm_q[escaped].copy(),
        "near_miss_traj": np.where(escaped)[0].astype(np.int64),
        "near_miss_dist": tr[escaped, 7].copy(),
        "radial_bins": np.linspace(0, 1.2 * r_esc, nrad + 1), "radial_counts": rad,
        "angular_theta": np.linspace(0, np.pi, nth), "angular_phi": np.linspace(0, 2 * np.pi, nph),
        "angular_counts": ang,
        "milestone_radii": np.linspace(r_b, r_esc, nms), "milestone_flux_out": ms_out,
        "milestone_flux_in": ms_in,
        "contact_pair_counts": contact[:int(P.n_pairs)] if P.n_pairs else contact[:1],
        "contact_total_steps": contact_steps,
        "trans_bins": np.linspace(0, 1.2 * r_esc, ntr + 1), "trans_matrix": trans,
    }


def autotune_sim(lib, ctx, h, rounds=3, steps_per_cand=1):
    """Runs pst_prod_autotune on a production sim handle; returns the table of
    candidates {G, cap, xcd, ms} and the chosen launch."""
    table = np.zeros((96, 5))
    n = C.c_int()
    best = (C.c_int * 5)()
    ctx.check(lib.pst_prod_autotune(h, int(rounds), int(steps_per_cand), ptr(table), 96,
                                    C.byref(n), best), "pst_prod_autotune")
    names = GATHER_MODES
    rows = [{"mode": names[int(r[0])], "G": int(r[1]), "cap": int(r[2]),
             "xcd": int(r[3]), "ms": float(r[4])} for r in table[:n.value]]
    return {"best": {"mode": names[int(best[0])], "G": best[1], "cap": best[2],
                     "xcd": best[3]},
            "table": rows, "steps": int(best[4])}


def _shoemake_host(u):
    return np.stack([np.sqrt(1 - u[:, 0]) * np.sin(2 * np.pi * u[:, 1]),
                     np.sqrt(1 - u[:, 0]) * np.cos(2 * np.pi * u[:, 1]),
                     np.sqrt(u[:, 0]) * np.sin(2 * np.pi * u[:, 2]),
                     np.sqrt(u[:, 0]) * np.cos(2 * np.pi * u[:, 2])], axis=1)


def stall_stop(n_total, done, done_before):
    """The reference's stall rule (gpu_batch_simulator.py:1349-1356), checked
    every STALL_EVERY steps: stop when more than 95 % of the trajectories are
    done and fewer than max(2, 0.1 % of them) finished since the last check."""
    return done > int(n_total * 0.95) and done - done_before < max(2, int(n_total * 0.001))


class AuditWindow:
    """Pools the surrogate audit's sums (squared force error, squared force,
    audited poses, squared torque error, squared torque) over chunks of steps
    until at least `min_poses` poses are in, then yields the window's relative
    RMS force and torque errors for one fallback decision and starts a new
    window."""

    def __init__(self, min_poses=AUDIT_MIN_POSES, sums=None):
        self.min_poses = int(min_poses)
        self.sums = np.zeros(5) if sums is None else np.asarray(sums, dtype=np.float64).copy()

    def add(self, err2, ref2, n, terr2=0.0, tref2=0.0):
        """Adds one chunk; returns (force error, torque error) when the window
        is complete (torque error 0 without audited torques), else None."""
        self.sums += (err2, ref2, n, terr2, tref2)
        e2, r2, k, t2, tr2 = self.sums
        if k < self.min_poses or r2 <= 0:
            return None
        self.sums[:] = 0.0
        return math.sqrt(e2 / r2), (math.sqrt(t2 / tr2) if tr2 > 0 else 0.0)


def is_stall_check(step, every=STALL_EVERY):
    """The reference checks after loop index i > 0 with i % 10000 == 0, i.e.
    after i + 1 moves; `step` here counts moves."""
    return step > 1 and (step - 1) % every == 0


def next_stall_check(step, every=STALL_EVERY):
    """The first move count after `step` at which is_stall_check holds."""
    return (max(step - 1, 0) // every + 1) * every + 1


def next_snapshot(step, every):
    """The first move count after `step` at which the reference records a
    snapshot (loop index i with i % save_interval == 0, i.e. after i + 1 moves)."""
    return (max(step - 1, -1) // every + 1) * every + 1


def convergence_milestones(n_total, interval):
    """Done counts at which the reference reports live k_on
    (gpu_batch_simulator.py:753-761): every `interval` percent below 100."""
    interval = int(interval or 0)
    if interval <= 0:
        return []
    return sorted({int(n_total * q / 100) for q in range(interval, 100, interval)})


def device_step_bound(max_steps, n_total, n_slots):
    """Device steps that suffice for every trajectory to get its own
    max_steps: while the pool still holds trajectories every slot is busy, so
    all of them start within ceil(N / S) * max_steps steps, and the last one
    ends max_steps later.  (The reference runs every trajectory from step 0,
    so its loop bound max_steps is also each trajectory's.)"""
    waves = -(-int(n_total) // max(1, int(n_slots)))
    return int(max_steps) * (waves + 1)


def snapshot_config(params):
    """(save_interval, paths, energetics) of the reference's _output_cfg
    (gpu_batch_simulator.py:578-581, :1318-1340): every 10 steps, both on,
    when no output configuration is attached."""
    cfg = getattr(params, "_output_cfg", None)
    if cfg is None:
        return 10, True, True
    return (max(int(getattr(cfg, "save_interval", 10)), 1), bool(getattr(cfg, "full_paths", True)),
            bool(getattr(cfg, "energetics", True)))


_TUNE_CACHE = {}


def _tune_cache_file():
    return os.environ.get("PYSTARC_AMD_TUNE_CACHE")


def tune_cache_get(key):
    """Launch chosen earlier for this key (this process, or the JSON file
    named by PYSTARC_AMD_TUNE_CACHE)."""
    if key in _TUNE_CACHE:
        return _TUNE_CACHE[key]
    f = _tune_cache_file()
    if f and os.path.exists(f):
        import json
        try:
            with open(f) as fh:
                d = json.load(fh)
            v = d.get(key)
            if v is not None:
                _TUNE_CACHE[key] = v
            return v
        except (OSError, ValueError):
            return None
    return None


def tune_cache_put(key, best):
    _TUNE_CACHE[key] = best
    f = _tune_cache_file()
    if f:
        import json
        d = {}
        try:
            if os.path.exists(f):
                with open(f) as fh:
                    d = json.load(fh)
        except (OSError, ValueError):
            d = {}
        d[key] = best
        tmp = f + ".tmp"
        with open(tmp, "w") as fh:
            json.dump(d, fh)
        os.replace(tmp, f)


_CONV_KEYS = ("milestone", "done", "pct", "P_rxn", "k_on", "n_reacted", "n_escaped", "step",
              "elapsed_s")


def _split_rows(rows, counts):
    out, o = [], 0
    for n in np.asarray(counts).tolist():
        out.append(np.asarray(rows[o:o + n]))
        o += n
    return out


class Checkpoint:
    """checkpoint.npz of a production run in params._work_dir (default
    "bd_sims"), with the reference's file contract
    (gpu_batch_simulator.py:548-575 resume when the file exists, :764-768 and
    :1400-1420 written whenever checkpoint_interval more trajectories have
    completed, :1505-1510 removed when the run completes).  It holds the
    library's saved device state (pst_prod_save_state: slots, counters,
    outcomes, observables), the host loop's counters and the sim_data lists
    collected so far, all as plain arrays (loaded with allow_pickle=False),
    and a fingerprint of the run's parameters; a file of another run (or of
    the reference) is ignored, as the reference ignores one it cannot load."""

    def __init__(self, work_dir, fingerprint):
        self.dir = Path(work_dir)
        self.path = self.dir / "checkpoint.npz"
        self.fingerprint = fingerprint

    def exists(self):
        return self.path.exists()

    def save(self, lib, ctx, h, host):
        nb = C.c_longlong()
        ctx.check(lib.pst_prod_state_bytes(h, C.byref(nb)), "pst_prod_state_bytes")
        blob = np.zeros(nb.value, dtype=np.uint8)
        ctx.check(lib.pst_prod_save_state(h, ptr(blob), nb.value), "pst_prod_save_state")
        self.dir.mkdir(parents=True, exist_ok=True)
        arrays = {"state": blob, "fingerprint": np.array(self.fingerprint)}
        for k, v in host.items():
            arrays[k] = np.asarray(v)
        tmp = self.dir / "checkpoint.tmp.npz"
        np.savez_compressed(str(tmp), **arrays)
        os.replace(tmp, self.path)

    def load(self, lib, ctx, h):
        """Restores the device state; returns the host entries, or None when
        the file is not this run's (it is then left alone)."""
        try:
            with np.load(str(self.path), allow_pickle=False) as z:
                if str(z["fingerprint"]) != self.fingerprint:
                    log.warning("checkpoint %s belongs to another run; starting fresh", self.path)
                    return None
                host = {k: z[k] for k in z.files if k not in ("state", "fingerprint")}
                blob = np.ascontiguousarray(z["state"])
        except (OSError, ValueError, KeyError) as e:
            log.warning("checkpoint load failed: %s. Starting fresh.", e)
            return None
        ctx.check(lib.pst_prod_load_state(h, ptr(blob), blob.size), "pst_prod_load_state")
        return host

    def remove(self):
        if self.path.exists():
            self.path.unlink()


class GPUBatchSimulator:
    def __init__(self, mol1, mol2, mob, pathway_set, params, batch_engine, *, rng="philox",
                 n_slots=None, precision=None, graph_steps=16, record=False, surrogate=None,
                 audit_n=64, audit_tol=AUDIT_TOL, hydro="scalar", autotune=False, gather=None,
                 traj_range=None, outer="return_prob", snapshots=None,
                 stall_interval=STALL_EVERY):
        if rng not in ("philox", "numpy"):
            raise ValueError("rng must be 'philox' or 'numpy'")
        if traj_range is not None:
            first, count = (int(x) for x in traj_range)
            if first < 0 or count <= 0:
                raise ValueError("traj_range needs first >= 0 and count > 0")
            if rng != "philox":
                raise ValueError("traj_range needs rng='philox' (replay shards split by seed)")
            traj_range = (first, count)
        self.traj_range = traj_range
        if outer not in ("return_prob", "lmz"):
            raise ValueError("outer must be 'return_prob' or 'lmz'")
        if outer == "lmz" and rng != "philox":
            raise ValueError("outer='lmz' needs rng='philox'")
        self.outer = outer
        self.outer_propagator = None
        if hydro not in ("scalar", "rpy12"):
            raise ValueError("hydro must be 'scalar' or 'rpy12'")
        if hydro == "rpy12" and rng != "philox":
            raise ValueError("replay follows the reference's scalar diffusion; rpy12 needs rng='philox'")
        if hydro == "rpy12" and not (getattr(mob, "radius1", 0) > 0 and getattr(mob, "radius2", 0) > 0):
            raise ValueError("rpy12 needs both hydrodynamic radii (mob.radius1, mob.radius2)")
        self.hydro = hydro
        self.autotune = int(autotune) if autotune is not True else 64
        if gather is not None and gather not in GATHER_MODES:
            raise ValueError(f"gather must be None or one of {GATHER_MODES}")
        self.gather = gather
        self.mol1, self.mol2, self.mob = mol1, mol2, mob
        self.pathway_set, self.params, self.engine = pathway_set, params, batch_engine
        self.rng_mode = rng
        if surrogate is not None and rng != "philox":
            raise ValueError("the surrogate runs in philox mode only")
        self.surrogate = surrogate
        self.audit_n = int(audit_n)
        self.audit_tol = float(audit_tol)
        self.graph_steps = int(graph_steps)
        if record not in (False, True, "observables"):
            raise ValueError("record must be False, True or 'observables'")
        self.record = 2 if record == "observables" else int(bool(record))
        if snapshots not in (None, True, False):
            raise ValueError("snapshots must be None, True or False")
        if snapshots and rng != "philox":
            raise ValueError("snapshots need rng='philox'")
        self.snapshots = (self.record == 2) if snapshots is None else bool(snapshots)
        self.stall_interval = int(stall_interval)
        if self.stall_interval < 1:
            raise ValueError("stall_interval must be positive")
        prec = precision if precision is not None else getattr(batch_engine, "precision", "fp64")
        if prec not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.precision = prec
        c0 = mol2.centroid()
        self._mol2_pos0 = np.asarray(mol2.positions_array() - c0, dtype=np.float64)
        self._charges = np.asarray(mol2.charges_array(), dtype=np.float64)
        self._N_lig = len(self._charges)
        self._use_hi = bool(getattr(params, "hydrodynamic_interactions", False))
        self._qb_factor = QB_FACTOR
        # state-machine reactions (:216-318): one pair (pair0) per reaction
        self._sm = state_machine(pathway_set)
        if self._sm is not None and rng != "philox":
            raise ValueError("state-machine reactions run in philox mode")
        # flattened reaction pairs (:186-205); in state-machine mode pair r is
        # reaction r's pair0
        rec_idx, lig_idx, cut = [], [], []
        for rxn in pathway_set.reactions:
            pairs = rxn.criteria.pairs
            if self._sm is not None:
                pairs = pairs[:1]
            for pair in pairs:
                rec_idx.append(pair.mol1_atom_index)
                lig_idx.append(pair.mol2_atom_index)
                cut.append(pair.distance_cutoff)
        self._rec_gho_indices, self._lig_gho_indices = rec_idx, lig_idx
        self._rxn_cutoffs = np.array(cut, dtype=np.float64)
        if len(cut) > MAX_PAIRS:
            raise ValueError(f"at most {MAX_PAIRS} reaction pairs are supported")
        if pathway_set.reactions:
            nn = pathway_set.reactions[0].criteria.n_needed
            self._rxn_n_needed = len(cut) if (nn < 0 or nn > len(cut)) else nn
        else:
            self._rxn_n_needed = 1
        if self._sm is not None:
            self._rxn_n_needed = 1
        rec_pos = np.asarray(mol1.positions_array(), dtype=np.float64)
        self._rec_gho_pos = rec_pos[rec_idx] if rec_idx else None
        N = int(params.n_trajectories) if traj_range is None else traj_range[1]
        self.n_slots = int(n_slots) if n_slots else min(N, 1 << 20)
        if rng == "numpy":
            self.n_slots = N
        self._return_prob = 0.0
        self._k_b = 0.0

    # -- outer propagator: k_b and return probability (:320-404) ----------------
    def _make_outer(self):
        from .outer_propagator import OuterPropagator
        q_rec, q_lig = float(self.mol1.total_charge()), float(self.mol2.total_charge())
        kT_scale = float(getattr(self.params, "_kT_scale", 1.0))
        V = q_rec * q_lig / (4.0 * math.pi * 78.0 * EPS0) / kT_scale
        return OuterPropagator(
            float(self.params.r_start),
            max(self.mol1.bounding_radius(), self.mol2.bounding_radius()), V,
            float(getattr(self.engine, "_debye", 7.858)),
            float(self.mob.relative_translational_diffusion()), self._use_hi,
            float(getattr(self.mob, "radius1", 0.0)), float(getattr(self.mob, "radius2", 0.0)),
            float(self.mob.D_rot1), float(self.mob.D_rot2))

    def _compute_return_prob(self) -> float:
        b = float(self.params.r_start)
        q_out = float(self.params.r_escape)
        q_rec = float(self.mol1.total_charge())
        q_lig = float(self.mol2.total_charge())
        debye = float(getattr(self.engine, "_debye", 7.858))
        eps = 78.0 * EPS0
        D_t = float(self.mob.relative_translational_diffusion())
        kT_scale = float(getattr(self.params, "_kT_scale", 1.0))
        if abs(q_rec * q_lig) < 1e-9:
            self._k_b = 4.0 * math.pi * D_t * b
            return b / q_out
        if self._use_hi:
            a0, a1 = float(self.mob.radius1), float(self.mob.radius2)
            a2 = 0.5 * (a0 ** 2 + a1 ** 2)
            ainv = 1.0 / a0 + 1.0 / a1
            dpre = D_t / ainv

        def integrand(s):
            if s == 0.0:
                return 1.0 / D_t
            r = 1.0 / s
            w = math.exp(q_rec * q_lig / (4.0 * math.pi * eps * r) * math.exp(-r / debye) / kT_scale)
            if self._use_hi:
                return w / (dpre * (ainv - 3.0 / r + 2.0 * a2 / r ** 3))
            return w / D_t

        try:
            self._k_b = 4.0 * math.pi / romberg_integrate(integrand, 0.0, 1.0 / b, tol=1e-8)
            k_q = 4.0 * math.pi / romberg_integrate(integrand, 0.0, 1.0 / q_out, tol=1e-8)
            rp = self._k_b / k_q if k_q > 0 else 0.5
            return float(max(0.0, min(1.0, rp)))
        except (OverflowError, ZeroDivisionError, ValueError):
            self._k_b = 4.0 * math.pi * D_t * b
            return b / q_out

    def _step_params(self) -> StepParams:
        p = self.params
        P = StepParams()
        N = int(p.n_trajectories)
        if self.traj_range is not None:
            P.traj_base, N = self.traj_range
        if self.outer == "lmz":
            self.outer_propagator = self._make_outer()
            self.outer_propagator.fill(P)
        P.n_total = N
        P.n_slots = self.n_slots
        P.seed = int(p.seed if p.seed is not None else 0) & ((1 << 64) - 1)
        P.r_b = float(p.r_start)
        P.r_esc = float(p.r_escape)
        P.return_prob = self._return_prob
        D_t = float(self.mob.relative_translational_diffusion())
        P.D_t = D_t
        P.D_r = float(self.mob.relative_rotational_diffusion())
        P.use_hi = int(self._use_hi)
        if self._use_hi:
            Df = 1.0 / MU_REF
            a0 = float(getattr(self.mob, "radius1", 0.0))
            a1 = float(getattr(self.mob, "radius2", 0.0))
            P.hi_Df = Df
            P.hi_a2 = 0.5 * (a0 ** 2 + a1 ** 2)
            P.hi_ainv = (Df / (6.0 * math.pi * a0) + Df / (6.0 * math.pi * a1)) / Df
        # adaptive-dt constants (:622-636)
        P.adt_debye = float(getattr(self.engine, "_debye", 7.828))
        V0 = float(self.mol1.total_charge()) * float(self.mol2.total_charge()) / (4.0 * math.pi * 78.0 * EPS0)
        P.adt_V0 = V0
        P.adt_has_force = int(abs(V0) > 1e-9)
        P.min_core_dt = float(getattr(p, "minimum_core_dt", 0.0))
        P.max_dt = float(getattr(p, "max_dt", 0.0) or 0.0)
        P.use_overlap = int(bool(getattr(p, "_overlap_check", True)))
        P.overlap_thresh = OVERLAP_THRESHOLD
        n_pairs = len(self._rxn_cutoffs)
        P.n_pairs = n_pairs
        P.n_needed = self._rxn_n_needed
        if n_pairs:
            # Brownian-bridge D_eff per pair (:643-660)
            a_r = float(getattr(self.mob, "radius1", 1.0))
            a_l = float(getattr(self.mob, "radius2", 1.0))
            a_sum = a_r + a_l if (a_r + a_l) > 0 else 1.0
            Dtr, Dtl = D_t * a_l / a_sum, D_t * a_r / a_sum
            Drr = Dtr * 3.0 / (4.0 * a_r ** 2) if a_r > 0 else 0.0
            Drl = Dtl * 3.0 / (4.0 * a_l ** 2) if a_l > 0 else 0.0
            lig_body = self._mol2_pos0[self._lig_gho_indices]
            L_rec = np.linalg.norm(self._rec_gho_pos, axis=1)
            L_lig = np.linalg.norm(lig_body, axis=1)
            deff = D_t + Drr * L_rec ** 2 + Drl * L_lig ** 2
            P.has_bb_deff = 1
            for k in range(n_pairs):
                for d in range(3):
                    P.rec_gho[3 * k + d] = float(self._rec_gho_pos[k, d])
                    P.lig_gho[3 * k + d] = float(lig_body[k, d])
                P.cutoff[k] = float(self._rxn_cutoffs[k])
                P.bb_deff[k] = float(deff[k])
        P.max_steps = int(p.max_steps)
        P.hydro_mode = 1 if self.hydro == "rpy12" else 0
        P.a_rec = float(getattr(self.mob, "radius1", 0.0))
        P.a_lig = float(getattr(self.mob, "radius2", 0.0))
        P.kT_over_eta = 1.0 / MU_REF
        P.surrogate_r = -1.0
        P.stop_done = int(N * 0.995)
        if self._sm is not None:
            sm = self._sm
            P.sm_active = 1
            for r, (b, a) in enumerate(zip(sm["before"], sm["after"])):
                P.sm_before[r] = b
                P.sm_after[r] = a
            P.sm_terminal = sum(1 << i for i, t in enumerate(sm["terminal"]) if t)
            P.min_core_rxn_dt = float(getattr(p, "minimum_core_reaction_dt", 0.0) or 0.0)
        return P

    def _make_sim(self, P):
        eng = self.engine
        eng.set_ligand(self._mol2_pos0, self._charges)
        h = C.c_void_p()
        prec = 0 if self.precision == "fp64" else 1
        eng.ctx.check(eng.ctx.lib.pst_sim_create(eng.ctx.h, C.byref(P), prec, C.byref(h)),
                      "pst_sim_create")
        if self.gather is not None:
            eng.ctx.check(eng.ctx.lib.pst_prod_set_launch(
                h, GATHER_MODES.index(self.gather), 0, 2048, 1),
                "pst_prod_set_launch")
        return h

    # -- run -----------------------------------------------------------------------
    def run(self, max_device_steps=None) -> GPUBatchResult:
        """Run to completion (the reference's stop rules).  max_device_steps
        (philox mode) ends the run earlier after that many device steps, as an
        interrupted run: its checkpoint, if one is being written, is kept."""
        self._return_prob = self._compute_return_prob()
        P = self._step_params()
        if self.rng_mode == "numpy":
            return self._run_replay(P)
        return self._run_philox(P, max_device_steps)

    def _result(self, N, n_r, n_e, n_m, total, elapsed, sim_data):
        return GPUBatchResult(
            n_trajectories=N, n_reacted=int(n_r), n_escaped=int(n_e), n_max_steps=int(n_m),
            reaction_counts={"stage_0": int(n_r)}, r_start=float(self.params.r_start),
            r_escape=float(self.params.r_escape), dt=float(getattr(self.params, "dt", 0.2)),
            elapsed_sec=elapsed, steps_per_sec=total / elapsed if elapsed > 0 else 0.0,
            sim_data=sim_data, total_steps=int(total), k_b=self._k_b, return_prob=self._return_prob)

    def tune_key(self):
        """Autotuner cache key: ligand and receptor atoms, grid dimensions,
        precision, slots, hydrodynamics, torque mode and surrogate on/off."""
        eng = self.engine
        dims = [(g["nx"], g["ny"], g["nz"]) for g in getattr(eng, "_elec_grids_gpu", [])
                + getattr(eng, "_born_grids_gpu", []) + getattr(eng, "_lig_born_grids_gpu", [])]
        n_rec = 0 if getattr(eng, "_rec_pos", None) is None else len(eng._rec_pos)
        return (f"M{self._N_lig}-R{n_rec}-G{dims}-{self.precision}-S{self.n_slots}-{self.hydro}-"
                f"{getattr(eng, 'torque_mode', 'reference')}-surr{int(self.surrogate is not None)}")

    def fingerprint(self, P):
        """Identity of a run for its checkpoint: the step parameters (seed,
        trajectory count, physics, pairs) and the options that change results."""
        hsh = hashlib.sha256(bytes(P))
        extra = (self.precision, self.record, self.n_slots, self.hydro, self.outer,
                 None if self.surrogate is None else
                 tuple(np.round(self.surrogate.sim_params(self.audit_n), 12).tolist()))
        hsh.update(repr(extra).encode())
        return hsh.hexdigest()

    def _run_philox(self, P, max_device_steps=None) -> GPUBatchResult:
        eng = self.engine
        lib, ctx = eng.ctx.lib, eng.ctx
        p = self.params
        h = self._make_sim(P)
        N = int(P.n_total)
        try:
            ctx.check(lib.pst_prod_init(h, int(self.record)), "pst_prod_init")
            surr_on = False
            if self.surrogate is not None:
                p9 = np.ascontiguousarray(self.surrogate.sim_params(self.audit_n), dtype=np.float64)
                ctx.check(lib.pst_sim_set_surrogate(h, self.surrogate.h, ptr(p9)),
                          "pst_sim_set_surrogate")
                surr_on = True
            audit = {"err2": 0.0, "ref2": 0.0, "n": 0, "worst": 0.0, "fallback_step": None,
                     "terr2": 0.0, "tref2": 0.0, "worst_torque": 0.0}
            window = AuditWindow()
            c16 = np.zeros(16)
            steps = stall_done = last_ckpt = 0
            conv_done = 0
            paths, energies, conv = [], [], []
            save_every, want_paths, want_energy = snapshot_config(p)
            snap = self.snapshots and (want_paths or want_energy)
            ckpt_every = int(getattr(p, "checkpoint_interval", 0) or 0)
            ckpt = Checkpoint(getattr(p, "_work_dir", "bd_sims"), self.fingerprint(P))
            resumed = False
            if ckpt.exists():
                host = ckpt.load(lib, ctx, h)
                if host is not None:
                    resumed = True
                    steps, stall_done, last_ckpt, conv_done = (int(x) for x in host["loop"])
                    a = host["audit"]
                    audit.update(err2=float(a[0]), ref2=float(a[1]), n=int(a[2]),
                                 worst=float(a[3]), terr2=float(a[4]), tref2=float(a[5]),
                                 worst_torque=float(a[6]),
                                 fallback_step=None if a[7] < 0 else int(a[7]))
                    if len(a) >= 13:
                        window = AuditWindow(sums=a[8:13])
                    if surr_on and audit["fallback_step"] is not None:
                        ctx.check(lib.pst_sim_set_surrogate(h, None, None), "pst_sim_set_surrogate")
                        surr_on = False
                    if "paths" in host:
                        paths = _split_rows(host["paths"], host["paths_n"])
                    if "energies" in host:
                        energies = _split_rows(host["energies"], host["energies_n"])
                    if "conv" in host:
                        conv = [dict(zip(_CONV_KEYS, r.tolist())) for r in host["conv"]]
                    ctx.check(lib.pst_prod_counters(h, ptr(c16)), "pst_prod_counters")
                    log.info("resumed from checkpoint: step=%d, %d/%d done", steps,
                             int(c16[0] + c16[1] + c16[2]), N)
            milestones = convergence_milestones(N, getattr(p, "convergence_interval", 10))
            t0 = time.time()
            last_report = t0
            tuned = None
            if self.autotune and not resumed:
                key = self.tune_key()
                best = tune_cache_get(key)
                if best is not None:
                    ctx.check(lib.pst_prod_set_launch(h, GATHER_MODES.index(best["mode"]),
                        int(best["G"]), int(best["cap"]), int(best["xcd"])),
                        "pst_prod_set_launch")
                    tuned = {"best": best, "table": [], "steps": 0, "cached": True}
                else:
                    ctx.check(lib.pst_prod_run(h, self.autotune, self.graph_steps, ptr(c16)),
                              "pst_prod_run")
                    steps += int(c16[8])
                    tuned = autotune_sim(lib, ctx, h)
                    steps += tuned["steps"]
                    tune_cache_put(key, tuned["best"])
            limit = device_step_bound(P.max_steps, N, self.n_slots)
            if max_device_steps is not None:
                limit = min(limit, int(max_device_steps))
            chunk = max(self.graph_steps, 1) * 64
            stop_reason = "step_limit"
            while steps < limit:
                todo = min(chunk, limit - steps)
                # end chunks where the reference's loop checks for stalls and
                # records snapshots
                todo = min(todo, next_stall_check(steps, self.stall_interval) - steps)
                if snap:
                    todo = min(todo, next_snapshot(steps, save_every) - steps)
                todo = max(int(todo), 1)
                ctx.check(lib.pst_prod_run(h, todo, self.graph_steps, ptr(c16)), "pst_prod_run")
                ran = int(c16[8])
                steps += ran
                if surr_on and c16[13] > 0:
                    a6 = np.zeros(6)
                    ctx.check(lib.pst_prod_audit(h, ptr(a6)), "pst_prod_audit")
                    audit["err2"] += a6[0]
                    audit["ref2"] += a6[1]
                    audit["n"] += int(a6[2])
                    audit["terr2"] += a6[3]
                    audit["tref2"] += a6[4]
                    ctx.check(lib.pst_sim_reset_audit(h), "pst_sim_reset_audit")
                    res = window.add(a6[0], a6[1], a6[2], a6[3], a6[4])
                    err, terr = res if res is not None else (0.0, 0.0)
                    audit["worst"] = max(audit["worst"], err)
                    audit["worst_torque"] = max(audit["worst_torque"], terr)
                    if err > self.audit_tol or terr > self.audit_tol:
                        # the grid path takes over every trajectory from here on
                        ctx.check(lib.pst_sim_set_surrogate(h, None, None), "pst_sim_set_surrogate")
                        surr_on = False
                        audit["fallback_step"] = steps
                n_r, n_e, n_m = int(c16[0]), int(c16[1]), int(c16[2])
                done = n_r + n_e + n_m
                if snap and ran and (steps - 1) % save_every == 0:
                    self._snapshot(lib, ctx, h, steps - 1, want_paths, want_energy, paths, energies)
                for m in milestones:
                    if conv_done < m <= done:
                        conv_done = m
                        conv.append(self._conv_report(m, n_r, n_e, done, N, steps,
                                                      time.time() - t0))
                now = time.time()
                if now - last_report >= 10.0:
                    last_report = now
                    log.info("[T] %6.0fs done=%d/%d reacted=%d escaped=%d", now - t0, done, N,
                             n_r, n_e)
                if ckpt_every > 0 and done - last_ckpt >= ckpt_every:
                    last_ckpt = done
                    ckpt.save(lib, ctx, h, self._ckpt_host(steps, stall_done, last_ckpt, conv_done,
                                                           audit, paths, energies, conv, window))
                    log.info("checkpoint saved (%d/%d done) -> %s", done, N, ckpt.path)
                if c16[9]:
                    stop_reason = "done_99.5"
                    break
                if (c16[10] + c16[11]) == 0:
                    stop_reason = "complete"
                    break
                if ran < todo:
                    stop_reason = "complete"
                    break
                if is_stall_check(steps, self.stall_interval):
                    if stall_stop(N, done, stall_done):
                        log.info("stall detected: %d trajectories trapped (+%d in the last %d "
                                 "steps); stopping", N - done, done - stall_done,
                                 self.stall_interval)
                        stop_reason = "stall"
                        break
                    stall_done = done
            elapsed = time.time() - t0
            interrupted = (max_device_steps is not None and stop_reason == "step_limit"
                           and steps >= int(max_device_steps))
            if ckpt_every > 0 and interrupted:
                ctx.check(lib.pst_prod_counters(h, ptr(c16)), "pst_prod_counters")
                ckpt.save(lib, ctx, h, self._ckpt_host(steps, stall_done, last_ckpt, conv_done,
                                                       audit, paths, energies, conv, window))
            elif not interrupted and ckpt_every > 0:
                ckpt.remove()
            n_r, n_e = int(c16[0]), int(c16[1])
            n_m = N - n_r - n_e
            sim_data = {"n_returns_total": int(c16[5]), "n_overlap_rejected": int(c16[6]),
                        "n_bridge": int(c16[7]), "bd_steps": steps, "stop_reason": stop_reason,
                        "resumed": resumed}
            if milestones:
                sim_data["convergence"] = conv
            if snap:
                sim_data["path_steps"] = paths if want_paths else []
                sim_data["energy_steps"] = energies if want_energy else []
            if tuned is not None:
                sim_data["autotune"] = tuned
            if self.surrogate is not None:
                audit["rel_rms"] = math.sqrt(audit["err2"] / audit["ref2"]) if audit["ref2"] > 0 else 0.0
                audit["torque_rel_rms"] = (math.sqrt(audit["terr2"] / audit["tref2"])
                                           if audit["tref2"] > 0 else None)
                sim_data["surrogate_audit"] = audit
            if self.record:
                st = np.zeros(N, dtype=np.int8)
                ns = np.zeros(N, dtype=np.int32)
                ctx.check(lib.pst_prod_outcomes(h, ptr(st), ptr(ns)), "pst_prod_outcomes")
                sim_data["outcome"] = st
                sim_data["n_steps"] = ns
            if self.record == 2:
                sim_data.update(observables(lib, ctx, h, P, sim_data["outcome"]))
            if self._sm is not None:
                sim_data.update(sm_report(lib, ctx, h, self._sm))
            return self._result(N, n_r, n_e, n_m, c16[3], elapsed, sim_data)
        finally:
            lib.pst_sim_destroy(h)

    def _snapshot(self, lib, ctx, h, step, want_paths, want_energy, paths, energies):
        S = self.n_slots
        p8 = np.zeros((S, 8)) if want_paths else None
        e6 = np.zeros((S, 6)) if want_energy else None
        n = C.c_int()
        ctx.check(lib.pst_prod_snapshot(h, int(step), ptr(p8), ptr(e6), C.byref(n)),
                  "pst_prod_snapshot")
        if want_paths:
            paths.append(p8[:n.value].copy())
        if want_energy:
            energies.append(e6[:n.value].copy())

    def _conv_report(self, milestone, n_r, n_e, done, N, steps, elapsed):
        """Live k_on when `milestone` trajectories are done
        (gpu_batch_simulator.py:1358-1372); the counts are those of the end of
        the chunk of steps in which the milestone was passed."""
        n_c = n_r + n_e
        P_live = n_r / n_c if n_c else 0.0
        k_live = CONV * self._k_b * P_live if self._k_b > 0 else 0.0
        pct = milestone / N * 100
        log.info("  -- %.0f%% converged --  P_rxn=%.5f  k_on=%.3e  (%d react / %d esc)  [%.0fs]",
                 pct, P_live, k_live, n_r, n_e, elapsed)
        return dict(zip(_CONV_KEYS, [float(milestone), float(done), pct, P_live, k_live,
                                     float(n_r), float(n_e), float(steps), elapsed]))

    def _ckpt_host(self, steps, stall_done, last_ckpt, conv_done, audit, paths, energies, conv,
                   window=None):
        a = [audit["err2"], audit["ref2"], audit["n"], audit["worst"], audit["terr2"],
             audit["tref2"], audit["worst_torque"],
             -1 if audit["fallback_step"] is None else audit["fallback_step"]]
        # the open audit window, so a resumed run decides its fallback where
        # the uninterrupted run does
        a += list(window.sums) if window is not None else [0.0] * 5
        host = {"loop": np.array([steps, stall_done, last_ckpt, conv_done], dtype=np.int64),
                "audit": np.array(a, dtype=np.float64)}
        if paths:
            host["paths"] = np.concatenate(paths)
            host["paths_n"] = np.array([len(x) for x in paths], dtype=np.int64)
        if energies:
            host["energies"] = np.concatenate(energies)
            host["energies_n"] = np.array([len(x) for x in energies], dtype=np.int64)
        if conv:
            host["conv"] = np.array([[r[k] for k in _CONV_KEYS] for r in conv], dtype=np.float64)
        return host

    def replay_driver(self) -> "ReplayDriver":
        """Step-by-step replay handle (run() drives it to completion)."""
        self._return_prob = self._compute_return_prob()
        return ReplayDriver(self, self._step_params())

    def production_driver(self, record=True) -> "ProductionDriver":
        """Step-by-step handle on the production kernels (philox mode, no
        surrogate): advance(k) runs k device steps, state() reads the slots."""
        if self.rng_mode != "philox":
            raise ValueError("production_driver needs rng='philox'")
        self._return_prob = self._compute_return_prob()
        return ProductionDriver(self, self._step_params(), record)


    def _run_replay(self, P) -> GPUBatchResult:
        drv = ReplayDriver(self, P)
        try:
            t0 = time.time()
            while drv.step():
                pass
            elapsed = time.time() - t0
            st, ns = drv.status_steps()
            N = int(P.n_total)
            n_r, n_e = int((st == 1).sum()), int((st == 2).sum())
            return self._result(N, n_r, n_e, int((st == 0).sum()), drv.total_steps, elapsed,
                                {"outcome": st, "n_steps": ns})
        finally:
            drv.close()


class ReplayDriver:
    """One reference loop iteration per step() with host NumPy draws in the
    reference's order (gpu_batch_simulator.py:472-513, :783-1356)."""

    def __init__(self, sim: GPUBatchSimulator, P: StepParams):
        self.sim = sim
        self.P = P
        eng = sim.engine
        self.lib, self.ctx = eng.ctx.lib, eng.ctx
        self.N = int(P.n_total)
        self.rng = np.random.default_rng(sim.params.seed)
        v = self.rng.standard_normal((self.N, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        pos = np.ascontiguousarray(v * float(P.r_b))
        q = np.ascontiguousarray(_shoemake_host(self.rng.random((self.N, 3)).astype(np.float64)))
        self.h = sim._make_sim(P)
        self.ctx.check(self.lib.pst_sim_set_state(self.h, self.N, ptr(pos), ptr(q)),
                       "pst_sim_set_state")
        self.total_steps = 0
        self.n_done = 0
        self.stall_done = 0
        self.step_index = 0
        self.last = {}

    def step(self) -> bool:
        """Returns False when the reference loop would stop."""
        lib, ctx, rng, h = self.lib, self.ctx, self.rng, self.h
        P = self.P
        if self.step_index >= int(P.max_steps):
            return False
        a, b_, c_ = C.c_int(), C.c_int(), C.c_int()
        ctx.check(lib.pst_replay_begin(h, C.byref(a), C.byref(b_), C.byref(c_)), "pst_replay_begin")
        n_run, n_esc, n_react = a.value, b_.value, c_.value
        if n_run == 0:
            return False
        self.n_done += n_react
        u_outer = u_quat = None
        n_ret = 0
        if n_esc:
            u_outer = rng.random(n_esc)
            n_ret = int((u_outer < self.sim._return_prob).sum())
            if n_ret:
                u_quat = np.ascontiguousarray(rng.random((n_ret, 3)).astype(np.float64))
            self.n_done += n_esc - n_ret
        ctx.check(lib.pst_replay_escape(h, ptr(u_outer), n_esc, ptr(u_quat), n_ret, C.byref(a)),
                  "pst_replay_escape")
        n_sr = a.value
        self.last = {"n_run": n_run, "n_reacted_start": n_react, "n_return": n_ret,
                     "n_escape": n_esc - n_ret, "n_sr": n_sr, "bridge": False, "n_bridge": 0}
        step = self.step_index
        self.step_index += 1
        if n_sr == 0:
            return self.step_index < int(P.max_steps)
        nt = rng.standard_normal((n_sr, 3))
        nr = rng.standard_normal((n_sr, 3))
        ctx.check(lib.pst_replay_move(h, ptr(nt), ptr(nr), n_sr, C.byref(a), C.byref(b_)),
                  "pst_replay_move")
        self.last["n_overlap"] = b_.value
        u_bb = None
        if P.n_pairs and a.value:
            u_bb = np.ascontiguousarray(rng.random((n_sr, int(P.n_pairs))))
            self.last["bridge"] = True
        ctx.check(lib.pst_replay_bridge(h, ptr(u_bb), n_sr, C.byref(a)), "pst_replay_bridge")
        self.last["n_bridge"] = a.value
        self.n_done += a.value
        self.total_steps += n_sr
        if self.n_done >= int(self.N * 0.995):
            return False
        if step > 0 and step % 10000 == 0:
            if self.n_done > int(self.N * 0.95) and self.n_done - self.stall_done < max(2, int(self.N * 0.001)):
                return False
            self.stall_done = self.n_done
        return self.step_index < int(P.max_steps)

    def forces(self, n_sr):
        F = np.zeros((n_sr, 3))
        self.ctx.check(self.lib.pst_replay_forces(self.h, ptr(F), n_sr), "pst_replay_forces")
        return F

    def state(self):
        N = self.N
        pos = np.zeros((N, 3))
        q = np.zeros((N, 4))
        st = np.zeros(N, dtype=np.int32)
        ns = np.zeros(N, dtype=np.int64)
        self.ctx.check(self.lib.pst_sim_get_state(self.h, ptr(pos), ptr(q), ptr(st), ptr(ns)),
                       "pst_sim_get_state")
        return pos, q, st, ns

    def status_steps(self):
        _, _, st, ns = self.state()
        return st, ns

    def close(self):
        if self.h:
            self.lib.pst_sim_destroy(self.h)
            self.h = None


class ProductionDriver:
    """Production step in host-chosen increments.  With n_slots >= N slot s
    holds trajectory s for the whole run (no refill), so the slot state is
    the trajectory state."""

    def __init__(self, sim: GPUBatchSimulator, P: StepParams, record: bool):
        self.sim, self.P = sim, P
        self.lib, self.ctx = sim.engine.ctx.lib, sim.engine.ctx
        self.h = sim._make_sim(P)
        self.record = 2 if record == "observables" else int(bool(record))
        self.ctx.check(self.lib.pst_prod_init(self.h, self.record), "pst_prod_init")
        self.c16 = np.zeros(16)
        self.steps = 0

    def advance(self, k: int, graph_steps: int = 0) -> int:
        """Runs up to k steps; returns the number run (fewer once the run stops)."""
        self.ctx.check(self.lib.pst_prod_run(self.h, int(k), int(graph_steps), ptr(self.c16)),
                       "pst_prod_run")
        n = int(self.c16[8])
        self.steps += n
        return n

    def autotune(self, rounds=2, steps_per_cand=1) -> Dict:
        """Tunes the force launch on live steps (they count as run steps)."""
        out = autotune_sim(self.lib, self.ctx, self.h, rounds, steps_per_cand)
        self.steps += out["steps"]
        return out

    def counters(self) -> Dict:
        c = self.c16
        return {"n_reacted": int(c[0]), "n_escaped": int(c[1]), "n_max_steps": int(c[2]),
                "total_steps": int(c[3]), "n_returns": int(c[5]), "n_overlap": int(c[6]),
                "n_bridge": int(c[7]), "stopped": bool(c[9])}

    def state(self):
        """(pos (S,3), quat (S,4), status (S,), n_steps (S,)) of the slots."""
        S = int(self.P.n_slots)
        pos = np.zeros((S, 3))
        q = np.zeros((S, 4))
        st = np.zeros(S, dtype=np.int32)
        ns = np.zeros(S, dtype=np.int64)
        self.ctx.check(self.lib.pst_sim_get_state(self.h, ptr(pos), ptr(q), ptr(st), ptr(ns)),
                       "pst_sim_get_state")
        return pos, q, st, ns

    def completed_reactions(self):
        """State-machine fire counts so far (state-machine pathway sets)."""
        if self.sim._sm is None:
            return None
        return sm_report(self.lib, self.ctx, self.h, self.sim._sm)["completed_reactions"]

    def observables(self):
        """The reference's sim_data entries collected so far (record='observables')."""
        st, _ = self.outcomes()
        return observables(self.lib, self.ctx, self.h, self.P, st)

    def outcomes(self):
        """Per-trajectory (outcome, steps): 0 running, 1 reacted, 2 escaped, 3 max steps."""
        N = int(self.P.n_total)
        st = np.zeros(N, dtype=np.int8)
        ns = np.zeros(N, dtype=np.int32)
        self.ctx.check(self.lib.pst_prod_outcomes(self.h, ptr(st), ptr(ns)), "pst_prod_outcomes")
        return st, ns

    def close(self):
        if self.h is not None:
            self.lib.pst_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GPUBatchGroup:
    """Several ligands against one receptor in one run: each ligand is a
    GPUBatchSimulator (its own engine, built with ``receptor_from`` so the
    receptor grids are resident once), and every step evaluates the
    grid-path forces of all ligands in ONE force launch (pst_group_run).
    ``run()`` returns one GPUBatchResult per ligand.  Philox mode; every
    simulator must use the same force precision.  Each ligand may carry its
    own surrogate: its audit is read after every chunk of steps and, when the
    relative RMS force error of a window of at least AUDIT_MIN_POSES audited
    poses exceeds that simulator's audit_tol, the
    grid path takes over that ligand's trajectories from then on (the other
    ligands keep theirs), as in a single run."""

    def __init__(self, simulators, graph_steps=16, stall_interval=STALL_EVERY):
        if not simulators:
            raise ValueError("no simulators")
        for s in simulators:
            if s.rng_mode != "philox":
                raise ValueError("group runs need rng='philox'")
        if len({s.precision for s in simulators}) != 1:
            raise ValueError("all simulators of a group need the same precision")
        self.sims = list(simulators)
        self.graph_steps = int(graph_steps)
        self.stall_interval = int(stall_interval)
        self.last_ms_per_step = 0.0

    def _open(self, handles, Ps, record=None):
        for s in self.sims:
            s._return_prob = s._compute_return_prob()
            P = s._step_params()
            h = s._make_sim(P)
            handles.append(h)
            Ps.append(P)
            rec = int(s.record) if record is None else int(record)
            s.engine.ctx.check(s.engine.ctx.lib.pst_prod_init(h, rec), "pst_prod_init")
            if s.surrogate is not None:
                p9 = np.ascontiguousarray(s.surrogate.sim_params(s.audit_n), dtype=np.float64)
                s.engine.ctx.check(s.engine.ctx.lib.pst_sim_set_surrogate(h, s.surrogate.h, ptr(p9)),
                                   "pst_sim_set_surrogate")

    def timed_window(self, warmup, steps):
        """Throughput of the group on live steps: `warmup` untimed steps, then
        exactly `steps` steps timed on the host around device synchronisation
        (slots refill from each ligand's pool, so every step is a full batch
        while trajectories remain).  Returns trajectory-steps, wall time and
        the device time per step; surrogates stay on (no audit fallback)."""
        import torch
        sims = self.sims
        lib, ctx = sims[0].engine.ctx.lib, sims[0].engine.ctx
        handles, Ps = [], []
        try:
            self._open(handles, Ps, 0)
            L = len(sims)
            arr = (C.c_void_p * L)(*[h.value for h in handles])
            c16 = np.zeros((L, 16))
            ms = C.c_double()
            if warmup > 0:
                ctx.check(lib.pst_group_run(arr, L, int(warmup), self.graph_steps, ptr(c16),
                                            C.byref(ms)), "pst_group_run")
            for s, h in zip(sims, handles):
                if s.surrogate is not None:
                    ctx.check(lib.pst_sim_reset_audit(h), "pst_sim_reset_audit")
            base = c16[:, 3].copy()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.check(lib.pst_group_run(arr, L, int(steps), self.graph_steps, ptr(c16), C.byref(ms)),
                      "pst_group_run")
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            per = c16[:, 3] - base
            return {"steps": int(c16[0, 8]), "wall_s": wall, "traj_steps": float(per.sum()),
                    "traj_steps_per_ligand": per.tolist(), "device_ms_per_step": ms.value,
                    "value": float(per.sum()) / wall if wall > 0 else 0.0,
                    "audit_rel_rms": [math.sqrt(c[12] / c[13]) if c[13] > 0 else None for c in c16]}
        finally:
            for s, h in zip(sims, handles):
                s.engine.ctx.lib.pst_sim_destroy(h)

    def run(self, progress=None):
        """Run every ligand's trajectories to completion.  ``progress``, if
        given, is called after each chunk of steps with (steps so far, the
        (L, 16) counter array)."""
        sims = self.sims
        lib, ctx = sims[0].engine.ctx.lib, sims[0].engine.ctx
        handles, Ps = [], []
        try:
            self._open(handles, Ps)
            L = len(sims)
            c16 = np.zeros((L, 16))
            ms = C.c_double()
            limit = max(device_step_bound(P.max_steps, P.n_total, s.n_slots)
                        for s, P in zip(sims, Ps))
            surr_on = [s.surrogate is not None for s in sims]
            audits = [{"err2": 0.0, "ref2": 0.0, "n": 0, "worst": 0.0, "fallback_step": None}
                      for _ in sims]
            windows = [AuditWindow() for _ in sims]
            stall_done = [0] * L
            stalled = [False] * L
            t0 = time.time()
            steps = 0
            active = list(range(L))
            while steps < limit and active:
                todo = int(min(max(self.graph_steps, 1) * 64, limit - steps))
                todo = max(1, min(todo, next_stall_check(steps, self.stall_interval) - steps))
                La = len(active)
                arr = (C.c_void_p * La)(*[handles[l].value for l in active])
                ca = np.zeros((La, 16))
                ctx.check(lib.pst_group_run(arr, La, todo, self.graph_steps, ptr(ca), C.byref(ms)),
                          "pst_group_run")
                for k, l in enumerate(active):
                    c16[l] = ca[k]
                ran = int(ca[0, 8])
                steps += ran
                for l in active:
                    s = sims[l]
                    if not surr_on[l] or c16[l, 13] <= 0:
                        continue
                    a = audits[l]
                    a["err2"] += c16[l, 12]
                    a["ref2"] += c16[l, 13]
                    a["n"] += int(c16[l, 14])
                    ctx.check(lib.pst_sim_reset_audit(handles[l]), "pst_sim_reset_audit")
                    res = windows[l].add(c16[l, 12], c16[l, 13], c16[l, 14])
                    err = res[0] if res is not None else 0.0
                    a["worst"] = max(a["worst"], err)
                    if err > s.audit_tol:
                        ctx.check(lib.pst_sim_set_surrogate(handles[l], None, None),
                                  "pst_sim_set_surrogate")
                        surr_on[l] = False
                        a["fallback_step"] = steps
                if progress is not None:
                    progress(steps, c16)
                if is_stall_check(steps, self.stall_interval):
                    # the reference's stall rule, per ligand: a stalled ligand
                    # leaves the group (its remaining trajectories end as
                    # max-steps) while the others run on
                    for l in list(active):
                        P = Ps[l]
                        d = int(c16[l, 0] + c16[l, 1] + c16[l, 2])
                        if stall_stop(int(P.n_total), d, stall_done[l]):
                            stalled[l] = True
                            active.remove(l)
                            log.info("ligand %d: stall detected, %d trajectories trapped", l,
                                     int(P.n_total) - d)
                        stall_done[l] = d
                active = [l for l in active if not (c16[l, 9] or (c16[l, 10] + c16[l, 11]) == 0)]
                if ran < todo:
                    break
            elapsed = time.time() - t0
            self.last_ms_per_step = ms.value
            out = []
            for l, (s, h, P) in enumerate(zip(sims, handles, Ps)):
                c = c16[l]
                N = int(P.n_total)
                n_r, n_e = int(c[0]), int(c[1])
                sd = {"n_returns_total": int(c[5]), "n_overlap_rejected": int(c[6]),
                      "n_bridge": int(c[7]), "bd_steps": steps, "group_size": L,
                      "stop_reason": "stall" if stalled[l] else
                      ("done_99.5" if c[9] else "complete")}
                if s.surrogate is not None:
                    a = audits[l]
                    a["rel_rms"] = math.sqrt(a["err2"] / a["ref2"]) if a["ref2"] > 0 else 0.0
                    sd["surrogate_audit"] = a
                if s.record:
                    st = np.zeros(N, dtype=np.int8)
                    ns = np.zeros(N, dtype=np.int32)
                    ctx.check(lib.pst_prod_outcomes(h, ptr(st), ptr(ns)), "pst_prod_outcomes")
                    sd["outcome"] = st
                    sd["n_steps"] = ns
                out.append(s._result(N, n_r, n_e, N - n_r - n_e, c[3], elapsed, sd))
            return out
        finally:
            for s, h in zip(sims, handles):
                s.engine.ctx.lib.pst_sim_destroy(h)
