"""Benchmark: trajectory-steps/s of the batched BD step on one MI355X,
surrogate vs plain, with k_on of both paths from the same run.

Workloads (--workload):
  protein_protein  (default, the headline) the 1,700 / 1,400-atom pair:
                   193^3 FP32 receptor potential + Born grids and ligand Born
                   grid at 0.5 A, bidirectional Born, two-body RPY 12x12
                   mobility with hydrodynamic radii 17 / 15 A and FP64
                   Cholesky noise (--hydro scalar: the reference's scalar
                   D_parallel), overlap rejection, 2-of-5 contacts < 5.5 A,
                   b = 60 A, r_esc = 120 A.
  protein_ligand   the 3,220 / 18-atom protein-small-ligand system (161^3 /
                   65^3 grids, 2-of-3 contacts < 4 A, b = 40 A) and, as the
                   surrogate path, the same system with the learned surrogate
                   trained on plain-path trajectory poses plus shell samples.
  multi_ligand     16 ligands of U{12..64} atoms against a 4,700-atom receptor
                   on 257^3 grids, one grouped force launch per step.

A "step" is one BD step of every live trajectory slot; value =
trajectory-steps per second over the timed steps of the surrogate + FP32
path (inputs resident in HBM; the step graph is captured before the timed
window).  Paths timed on identical seeds:

  fp64       plain grid path, FP64 forces (the reference's arithmetic)
  fp32       plain grid path, FP32 forces
  surrogate  MLP beyond contact + 8 A, FP32 grid path inside, per-step audit

k_on: fp64 and surrogate each run --kon-trajectories complete trajectories;
P_rxn with Wilson 95 % intervals go in the JSON line, with whether the
surrogate's P lies in the plain path's interval and the two-sample z of the
difference (two independent estimates differ by more than the plain
interval's half-width about one time in six).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "trajectory-steps/sec on 1 MI355X, surrogate vs plain at equal k_on (Wilson 95%)"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
MLP_FLOP_PER_TRAJ = 2 * (9 * 256 + 3 * 256 * 256 + 256 * 13)   # 404,480
TUNE_STEPS = 150  # at most 3 rounds x 42 launch candidates + a 4 x 4 runoff; the warm-up is topped up
FORCE_TIMING_STEPS = 20


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_system(args, n_traj=None):
    from pystarc_amd import systems as S
    from pystarc_amd._lib import Context
    if args.workload == "protein_ligand":
        sysm = S.protein_ligand(n_traj=n_traj or args.trajectories, seed=2012)
    else:
        sysm = S.protein_protein(n_traj=n_traj or args.trajectories, seed=3013)
    t0 = time.time()
    ctx = Context(0)
    grids = S.build_grids(sysm, ctx)
    ctx.close()
    log(f"grids built in {time.time() - t0:.1f}s")
    return sysm, grids


def make_sim(sysm, eng, args, surrogate=None, n_slots=None, autotune=False):
    from pystarc_amd.simulation.gpu_batch_simulator import GPUBatchSimulator
    return GPUBatchSimulator(sysm.mol1, sysm.mol2, sysm.mob, sysm.pathway_set, sysm.params, eng,
                             n_slots=n_slots or args.slots, graph_steps=args.graph_steps,
                             surrogate=surrogate, audit_n=args.audit_n, audit_tol=args.audit_tol,
                             hydro=args.hydro, autotune=autotune, outer=args.outer)


def run_window(sysm, eng, args, surrogate=None):
    """Production run: warm-up steps, then exactly `steps` timed steps."""
    import torch
    from pystarc_amd._lib import ptr
    from pystarc_amd.simulation.gpu_batch_simulator import autotune_sim
    sim = make_sim(sysm, eng, args, surrogate)
    sim._return_prob = sim._compute_return_prob()
    P = sim._step_params()
    lib, ctx = eng.ctx.lib, eng.ctx
    h = sim._make_sim(P)
    c16 = np.zeros(16)
    ctx.check(lib.pst_prod_init(h, 0), "pst_prod_init")
    if surrogate is not None:
        p9 = np.ascontiguousarray(surrogate.sim_params(args.audit_n), dtype=np.float64)
        ctx.check(lib.pst_sim_set_surrogate(h, surrogate.h, ptr(p9)), "pst_sim_set_surrogate")
    # warm-up: plain steps, then (with --autotune) the launch tuner's steps,
    # which are real steps under each candidate launch
    tune = None
    n_tune = TUNE_STEPS if args.autotune else 0
    done = 0
    if args.warmup > n_tune:
        ctx.check(lib.pst_prod_run(h, args.warmup - n_tune, args.graph_steps, ptr(c16)), "warmup")
        done = int(c16[8])
    if n_tune:
        tune = autotune_sim(lib, ctx, h)
        done += tune["steps"]
        log(f"autotune: {tune['best']} after {tune['steps']} tuning steps")
    if done < args.warmup:
        # the tuner ran fewer steps than budgeted: the warm-up is exactly W steps
        ctx.check(lib.pst_prod_run(h, args.warmup - done, args.graph_steps, ptr(c16)), "warmup")
        done = args.warmup
    # the step graph of the chosen launch is captured and instantiated here,
    # outside the timed window (pst_prod_run re-launches it while nothing it
    # captured changes); these steps count as warm-up
    ctx.check(lib.pst_prod_run(h, args.graph_steps, args.graph_steps, ptr(c16)), "capture")
    done += int(c16[8])
    if surrogate is not None:
        ctx.check(lib.pst_sim_reset_audit(h), "pst_sim_reset_audit")
    # trajectory-steps so far (warm-up and tuning steps), read right before timing
    ctx.check(lib.pst_prod_counters(h, ptr(c16)), "pst_prod_counters")
    base_steps = c16[3]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.check(lib.pst_prod_run(h, args.steps, args.graph_steps, ptr(c16)), "timed")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = C.c_double()
    lib.pst_prod_last_ms_per_step(h, C.byref(ms))
    S_ = sim.n_slots
    pos = np.zeros((S_, 3)); q = np.zeros((S_, 4)); st = np.zeros(S_, dtype=np.int32)
    ctx.check(lib.pst_sim_get_state(h, ptr(pos), ptr(q), ptr(st), None), "get_state")
    # the force kernel as the step launches it: FORCE_TIMING_STEPS further steps,
    # each force launch timed by HIP events on the library's stream
    m6 = np.zeros(6)
    ctx.check(lib.pst_prod_time_forces(h, FORCE_TIMING_STEPS, ptr(m6)), "pst_prod_time_forces")
    lib.pst_sim_destroy(h)
    alg = m6[2] * (56 + 48) + (m6[3] + m6[4] + m6[5]) * 20 * 4
    roof_force = {"ms": m6[0], "bytes": alg, "launches": int(m6[1]),
                  "census": [m6[2], m6[3], m6[4], m6[5]],
                  "achieved_gbs": alg / (m6[0] * 1e-3) / 1e9 if m6[0] > 0 else 0.0}
    live = st == 0
    traj_steps = c16[3] - base_steps
    audit = math.sqrt(c16[12] / c16[13]) if c16[13] > 0 else None
    return {"wall_s": wall, "device_ms_per_step": ms.value, "steps": int(c16[8]),
            "warmup_steps_run": done, "tune_steps": tune["steps"] if tune else 0,
            "traj_steps": traj_steps, "value": traj_steps / wall if wall > 0 else 0.0,
            "pos": pos[live], "quat": q[live], "counters": c16.tolist(), "audit_rel_rms": audit,
            "n_surrogate_last_step": int(c16[11]),
            "launch": tune["best"] if tune else None, "roof_force": roof_force,
            "tune_fastest": sorted((r for r in tune["table"] if r["ms"] > 0),
                                   key=lambda r: r["ms"])[:8] if tune else None}


def mlp_roofline(surr, pos, quat, reps=20):
    """Average duration of the fused inference kernel on the features of n
    live surrogate poses (HIP events around the kernel's launches on the
    library's stream, pst_mlp_time_infer) and its FLOP rate; `call_ms` is
    the library's whole forward call (feature packing, kernel, unpacking)
    timed the same way from torch's stream."""
    import torch
    n = min(len(pos), surr.max_rows)
    X = torch.as_tensor(surr.features(pos[:n], quat[:n]), device=surr.device)
    surr.forward(X)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        surr.forward(X)
    b.record()
    torch.cuda.synchronize()
    call_ms = a.elapsed_time(b) / reps
    ms = C.c_double()
    # warm-up launches first: after the host-side feature work the GPU's
    # clocks ramp back up over ~200 launches (0.18 -> 0.15 ms per launch)
    surr.ctx.check(surr.lib.pst_mlp_time_infer(surr.h, n, 200, C.byref(ms)), "pst_mlp_time_infer")
    surr.ctx.check(surr.lib.pst_mlp_time_infer(surr.h, n, reps, C.byref(ms)), "pst_mlp_time_infer")
    ms = ms.value
    return {"ms": ms, "n": n, "tflops": MLP_FLOP_PER_TRAJ * n / (ms * 1e-3) / 1e12,
            "call_ms": call_ms}


def full_run(sysm, eng, args, surrogate=None):
    """Complete trajectories for k_on (99.5 % rule), returns P and Wilson CI."""
    sysm.params.n_trajectories = args.kon_trajectories
    res = make_sim(sysm, eng, args, surrogate, n_slots=min(args.kon_trajectories, args.slots),
                   autotune=64 if args.autotune else False).run()
    lo, hi = res.reaction_probability_ci(0.95)
    out = {"n": res.n_trajectories, "reacted": res.n_reacted, "escaped": res.n_escaped,
           "P": res.reaction_probability, "P_ci95": [lo, hi],
           "k_on": res.rate_constant(0.0, res.k_b),
           "k_on_ci95": list(res.rate_constant_ci(0.0, res.k_b)),
           "traj_steps_per_s": res.steps_per_sec}
    if surrogate is not None:
        out["audit"] = {k: (float(v) if v is not None else None)
                        for k, v in res.sim_data["surrogate_audit"].items()}
    sysm.params.n_trajectories = args.trajectories
    return out


def two_sample_z(a, b):
    """z of P_a - P_b for two reaction probabilities from independent runs
    (pooled binomial variance over reacted / (reacted + escaped))."""
    na, nb = a["reacted"] + a["escaped"], b["reacted"] + b["escaped"]
    p = (a["reacted"] + b["reacted"]) / max(na + nb, 1)
    se = math.sqrt(max(p * (1.0 - p) * (1.0 / max(na, 1) + 1.0 / max(nb, 1)), 1e-300))
    return (a["P"] - b["P"]) / se


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def _cpu_worker(task):
    """One host core: the NumPy FP64 production oracle on its own poses for
    about `seconds` of wall time; returns (trajectory-steps, seconds, loops)."""
    import os as _os
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        _os.environ[k] = "1"
    root, npz, system_kind, hydro, seconds, wid = task
    import sys as _sys
    if root not in _sys.path:
        _sys.path.insert(0, root)
    import numpy as _np
    import time as _time
    from oracle.production import ProductionOracle
    from pystarc_amd import systems as S
    from tests._systems import oracle_params, oracle_system
    z = _np.load(npz, allow_pickle=False)
    sysm = (S.protein_ligand(n_traj=100) if system_kind == "protein_ligand"
            else S.protein_protein(n_traj=100))
    grids = {}
    for role in ("elec", "born", "lig_born"):
        grids[role] = [(z[f"{role}{i}_data"], z[f"{role}{i}_origin"], float(z[f"{role}{i}_h"]))
                       for i in range(int(z[f"{role}_n"]))]
    pos, quat = z["pos"][wid::int(z["n_workers"])], z["quat"][wid::int(z["n_workers"])]
    ora = ProductionOracle(oracle_system(sysm, grids, chunk=0),
                           oracle_params(sysm, n_trajectories=len(pos)), hydro=hydro)
    ora.pos[:] = pos
    ora.q[:] = quat
    w0 = _time.perf_counter()
    steps = loops = 0
    while _time.perf_counter() - w0 < seconds:
        before = ora.total_steps
        if not ora.step():
            break
        loops += 1
        steps += ora.total_steps - before
    return steps, _time.perf_counter() - w0, loops


def cpu_baseline(sysm, grids, pos, quat, hydro, args, seconds=15.0, per_worker=32):
    """The NumPy FP64 oracle of the same production step (same
    hydrodynamics) on a bounded sample of live poses from the GPU run: one
    host core alone, then every core of the box's CPU share at once (one
    process per core, each on its own poses), each for about `seconds`."""
    import multiprocessing as mp
    import tempfile
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = os.cpu_count() or 1
    n_workers = max(1, min(share, 16))
    sel = np.random.default_rng(0).choice(len(pos), size=min(per_worker * n_workers, len(pos)),
                                          replace=False)
    arrays = {"pos": pos[sel], "quat": quat[sel], "n_workers": np.array(n_workers)}
    for role, lst in grids.items():
        arrays[f"{role}_n"] = np.array(len(lst))
        for i, g in enumerate(lst):
            arrays[f"{role}{i}_data"] = np.asarray(g.data)
            arrays[f"{role}{i}_origin"] = np.asarray(g.origin)
            arrays[f"{role}{i}_h"] = np.array(float(np.asarray(g.delta)[0, 0]))
    tmp = tempfile.NamedTemporaryFile(suffix=".npz", delete=False)
    tmp.close()
    try:
        np.savez(tmp.name, **arrays)
        ctx = mp.get_context("spawn")
        base = (ROOT, tmp.name, args.workload, hydro, seconds)
        with ctx.Pool(1) as pool:
            s1, w1, l1 = pool.map(_cpu_worker, [base + (0,)])[0]
        log(f"cpu baseline, 1 core: {s1 / w1:.4g} traj-steps/s")
        with ctx.Pool(n_workers) as pool:
            res = pool.map(_cpu_worker, [base + (w,) for w in range(n_workers)])
    finally:
        os.unlink(tmp.name)
    total = sum(r[0] / r[1] for r in res)
    return {"value": total, "unit": "trajectory-steps/s", "cores": n_workers, "kind": "port",
            "cpu_model": _cpu_model(), "one_core_value": s1 / w1,
            "sample": f"{len(sel)} live poses from the GPU run, {per_worker} per process, stepped "
                      f"for ~{seconds:.0f}s by the NumPy FP64 production oracle (hydro {hydro}): "
                      f"{n_workers} single-threaded processes at once, summed "
                      f"({sum(r[0] for r in res)} trajectory-steps); one core alone "
                      f"{s1} trajectory-steps in {w1:.1f}s"}


def gather_ceiling(eng, n_points=1 << 24, reps=10):
    """Gather ceiling of the receptor grids (pst_gather_ceiling): uniform
    and clustered points, algorithmic GB/s of the 160-byte stencil, and the
    locality sweep (each workgroup's points inside a cube of edge 12/mode A)."""
    def probe(mode):
        o4 = np.zeros(4)
        eng.ctx.check(eng.ctx.lib.pst_gather_ceiling(eng.ctx.h, mode, n_points, reps, ptr_np(o4)),
                      "pst_gather_ceiling")
        return {"ms": float(o4[0]), "gbs": float(o4[1]), "points": int(o4[2])}
    out = {"uniform": probe(0), "clustered": probe(1)}
    out["locality_sweep"] = [dict(cube_edge_A=12.0 / m, **(out["clustered"] if m == 1 else probe(m)))
                             for m in (1, 2, 3, 4, 6, 12)]
    return out


def ptr_np(a):
    from pystarc_amd._lib import ptr
    return ptr(a)


def harvest_poses(sysm, eng, args, n_want, r_on, every=4, n_slots=None):
    """Poses of the plain FP64 path's own trajectories beyond r_on: a fresh
    production run whose live slots are read every `every` steps."""
    from pystarc_amd._lib import ptr
    sim = make_sim(sysm, eng, args, n_slots=n_slots or args.slots)
    sim._return_prob = sim._compute_return_prob()
    P = sim._step_params()
    lib, ctx = eng.ctx.lib, eng.ctx
    h = sim._make_sim(P)
    S_ = sim.n_slots
    c16 = np.zeros(16)
    pos = np.zeros((S_, 3)); q = np.zeros((S_, 4)); st = np.zeros(S_, dtype=np.int32)
    got_p, got_q, n = [], [], 0
    try:
        ctx.check(lib.pst_prod_init(h, 0), "pst_prod_init")
        while n < n_want:
            ctx.check(lib.pst_prod_run(h, every, 0, ptr(c16)), "harvest")
            ctx.check(lib.pst_sim_get_state(h, ptr(pos), ptr(q), ptr(st), None), "get_state")
            keep = (st == 0) & (np.linalg.norm(pos, axis=1) > r_on)
            got_p.append(pos[keep].copy())
            got_q.append(q[keep].copy())
            n += int(keep.sum())
            if c16[9] or int(c16[8]) < every:
                break
    finally:
        lib.pst_sim_destroy(h)
    return np.concatenate(got_p)[:n_want], np.concatenate(got_q)[:n_want]


def measured_traffic(workload, path, slots):
    """HBM bytes per force evaluation of `path` from the committed rocprofv3
    --pmc summary of the workload at the same slot count (FETCH_SIZE and
    WRITE_SIZE, calibrated), or None."""
    f = os.path.join(ROOT, "profiles", "k_forces_traffic.json")
    try:
        with open(f) as fh:
            d = json.load(fh)
        if d.get("workload_key", "protein_protein") != workload or d.get("slots") != slots:
            return None
        return d["paths"][path]["hbm_bytes_per_eval"]
    except (OSError, KeyError, ValueError):
        return None


def measured_limiter(workload, path, slots):
    """The committed cache-counter summary of the force kernel for `path`
    (L1 tag lookups per stencil, L1 / L2 hit rates, address / data unit busy
    fractions), or None."""
    f = os.path.join(ROOT, "profiles", "k_forces_traffic.json")
    try:
        with open(f) as fh:
            d = json.load(fh)
        cc = d.get("cache_counters") or {}
        if d.get("workload_key") != workload or d.get("slots") != slots or cc.get("path") != path:
            return None
        return dict(cc["derived"], source="profiles/k_forces_traffic.json")
    except (OSError, KeyError, ValueError):
        return None


def workload_text(args):
    if args.workload == "protein_ligand":
        return ("protein-small-ligand 3220/18 atoms, 161^3 FP32 receptor potential+Born grids "
                "(0.5 A), 65^3 ligand Born grid (0.4 A), bidirectional Born, "
                + ("two-body RPY 12x12 mobility" if args.hydro == "rpy12" else "scalar diffusion")
                + ", b=40 A, 2-of-3 contacts < 4 A; surrogate path = the same system with the "
                "MLP surrogate")
    return ("protein-protein 1700/1400 atoms, 193^3 FP32 grids (receptor potential+Born, ligand "
            "Born), bidirectional Born, "
            + ("two-body RPY 12x12 mobility (radii 17/15 A) with FP64 Cholesky noise"
               if args.hydro == "rpy12" else "scalar HI")
            + ", b=60 A, 2-of-5 contacts < 5.5 A")


def train_surrogate(sysm, eng64, args, extra_poses=None):
    from pystarc_amd.surrogate import build_surrogate
    t0 = time.time()
    s = build_surrogate(eng64, sysm, n_train=args.train_poses, epochs=args.train_epochs,
                        batch=args.train_batch, max_rows=max(args.slots, args.train_batch),
                        log=log, extra_poses=extra_poses, lr=args.train_lr)
    log(f"surrogate trained in {time.time() - t0:.1f}s, validation force error {s.val_error:.3f}")
    return s


def pmc_mode(args):
    """For rocprofv3 --pmc runs: the production window of one path (its
    force-timing steps sit between two k_timing_window_mark dispatches), then
    one plain device copy of 256 MiB whose FETCH_SIZE / WRITE_SIZE calibrate
    the counters' byte scale.  The surrogate path trains on the plain path's
    FP64 forces first (--train-poses / --train-epochs)."""
    import torch
    from pystarc_amd import systems as S
    sysm, grids = build_system(args)
    p = args.paths.split(",")[0]
    eng = S.make_engine(sysm, grids, precision="fp64" if p == "fp64" else "fp32")
    surr = None
    if p == "surrogate":
        surr = train_surrogate(sysm, S.make_engine(sysm, grids, precision="fp64"), args)
    r = run_window(sysm, eng, args, surr)
    log(f"pmc mode: path {p}, force kernel in the step {r['roof_force']['ms']:.3f} ms, "
        f"census {r['roof_force']['census']}, bytes {r['roof_force']['bytes']:.4g}")
    src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"pmc mode done: {len(r['pos'])} live poses, copy of {src.numel() * 4} bytes")


ML_WORKLOAD = ("multi-ligand: 4,700-atom receptor on 257^3 FP32 potential+Born grids (0.5 A), "
               "16 ligands of U{12..64} atoms each with its own 65^3 Born grid (0.4 A), one force "
               "launch per step for all 16, reverse Born over all receptor atoms, b=40 A, "
               "2-of-3 contacts < 4 A")


def multi_ligand_bench(args):
    """The multi-ligand workload: 16 ligands against one receptor whose grids
    are resident once, every step one grouped force launch.  Paths fp64 (plain)
    and surrogate (FP32 grid path + one MLP per ligand, audited, with the grid
    path taking over a ligand whose audit fails).  Prints one JSON line with
    the group's trajectory-steps/s and per-ligand k_on of both paths."""
    import torch
    from pystarc_amd import systems as S
    from pystarc_amd._lib import Context
    from pystarc_amd.simulation.gpu_batch_simulator import GPUBatchGroup, GPUBatchSimulator
    from pystarc_amd.surrogate import build_surrogate
    ls = S.multi_ligand(n_traj=args.kon_trajectories)
    t0 = time.time()
    ctx = Context(0)
    grids = [S.build_grids(s, ctx) for s in ls]
    ctx.close()
    log(f"multi-ligand: {len(ls)} ligands, atoms {[s.meta['atoms'] for s in ls]}, "
        f"grids built in {time.time() - t0:.1f}s")
    engines = {}
    for prec in ("fp64", "fp32"):
        e0 = S.make_engine(ls[0], grids[0], precision=prec)
        engines[prec] = [e0] + [S.make_engine(s, g, precision=prec, receptor_from=e0)
                                for s, g in zip(ls[1:], grids[1:])]
    paths = [p for p in args.paths.split(",") if p in ("fp64", "fp32", "surrogate")]
    surr = []
    t0 = time.time()
    for k, (s, e) in enumerate(zip(ls, engines["fp64"]) if "surrogate" in paths else ()):
        # training poses: shell samples plus poses of the ligand's own plain
        # FP64 trajectories
        r_on = float(s.mob.radius1 + s.mob.radius2) + 8.0
        extra = harvest_poses(s, e, args, args.ml_train_poses, r_on, n_slots=args.ligand_slots)
        surr.append(build_surrogate(e, s, n_train=args.ml_train_poses, epochs=args.ml_train_epochs,
                                    batch=args.train_batch, max_rows=max(args.ligand_slots, args.train_batch),
                                    seed=k, extra_poses=extra, lr=args.train_lr))
        log(f"ligand {k}: surrogate validation force error {surr[-1].val_error:.3f} "
            f"({len(extra[0])} trajectory poses)")
    t_train = time.time() - t0
    if surr:
        log(f"{len(surr)} surrogates trained in {t_train:.1f}s")

    def group(path, n_slots):
        prec = "fp64" if path == "fp64" else "fp32"
        sims = [GPUBatchSimulator(s.mol1, s.mol2, s.mob, s.pathway_set, s.params, e,
                                  n_slots=n_slots, graph_steps=args.graph_steps,
                                  surrogate=surr[k] if path == "surrogate" else None,
                                  audit_n=args.audit_n, audit_tol=args.audit_tol,
                                  hydro=args.hydro)
                for k, (s, e) in enumerate(zip(ls, engines[prec]))]
        return GPUBatchGroup(sims, graph_steps=args.graph_steps)

    results, kon = {}, {}
    for p in paths:
        for s in ls:
            s.params.n_trajectories = 10_000_000
        r = group(p, args.ligand_slots).timed_window(args.warmup, args.steps)
        log(f"path {p}: {r['value']:.4g} traj-steps/s, {r['device_ms_per_step']:.3f} ms/step device")
        results[p] = r
    if not args.no_kon:
        for p in ("fp64", "surrogate"):
            if p not in paths:
                continue
            for s in ls:
                s.params.n_trajectories = args.kon_trajectories
            t0 = time.time()
            last = [t0]

            def progress(steps, c16, p=p, t0=t0, last=last):
                if time.time() - last[0] > 20.0:
                    last[0] = time.time()
                    done = int((c16[:, 0] + c16[:, 1]).sum())
                    running = {k: int(c[0] + c[1]) for k, c in enumerate(c16) if not c[9]}
                    log(f"k_on {p}: step {steps}, {done} of {len(ls) * args.kon_trajectories} "
                        f"trajectories ended, {time.time() - t0:.0f}s; ligands still running "
                        f"(ended so far): {running}")

            res = group(p, min(args.kon_trajectories, args.ligand_slots)).run(progress)
            wall = time.time() - t0
            per = []
            for k, x in enumerate(res):
                lo, hi = x.reaction_probability_ci(0.95)
                d = {"ligand": k, "atoms": ls[k].meta["atoms"], "net": ls[k].meta["net"],
                     "n": x.n_trajectories, "reacted": x.n_reacted, "escaped": x.n_escaped,
                     "P": x.reaction_probability, "P_ci95": [lo, hi],
                     "k_on": x.rate_constant(0.0, x.k_b)}
                if "surrogate_audit" in x.sim_data:
                    a = x.sim_data["surrogate_audit"]
                    d["audit_rel_rms"] = a["rel_rms"]
                    d["fallback_step"] = a["fallback_step"]
                per.append(d)
            tot = sum(x.total_steps for x in res)
            kon[p] = {"wall_s": wall, "traj_steps": tot, "traj_steps_per_s": tot / wall,
                      "ligands": per}
            log(f"k_on {p}: {len(res)} ligands x {args.kon_trajectories} trajectories in "
                f"{wall:.1f}s ({tot / wall:.4g} traj-steps/s)")
            log(f"k_on result {p}: " + json.dumps(kon[p]))
        if "fp64" in kon and "surrogate" in kon:
            kon["surrogate_in_plain_ci95"] = [
                bool(a["P_ci95"][0] <= b["P"] <= a["P_ci95"][1])
                for a, b in zip(kon["fp64"]["ligands"], kon["surrogate"]["ligands"])]
            kon["two_sample_z"] = [two_sample_z(b, a) for a, b in
                                   zip(kon["fp64"]["ligands"], kon["surrogate"]["ligands"])]
    head = "surrogate" if "surrogate" in results else paths[0]
    best = results[head]
    plain = results.get("fp64")
    out = {
        "metric": METRIC, "value": best["value"], "unit": "trajectory-steps/s", "n_gpus": 1,
        "steps": best["steps"], "warmup": args.warmup,
        "ms_per_step": best["wall_s"] * 1e3 / max(best["steps"], 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": {"surrogate": "f64 state, f32 grid forces, bf16 MLP (f32 accumulate)",
                  "fp32": "f64 state, f32 forces", "fp64": "f64"}[head],
        "data": "synthetic",
        "config": {"workload": ML_WORKLOAD, "path": head, "slots_per_ligand": args.ligand_slots,
                   "graph_steps": args.graph_steps, "hydro": args.hydro},
        "speedup_vs_plain_fp64": (best["value"] / plain["value"]) if plain else None,
        "roofline": None, "cpu_baseline": None,
        "surrogates": {"train_poses": args.ml_train_poses, "epochs": args.ml_train_epochs,
                       "train_s": t_train, "validation_force_rel_rms": [x.val_error for x in surr]},
        "k_on": kon,
        "paths": {k: {kk: v[kk] for kk in ("value", "device_ms_per_step", "traj_steps_per_ligand",
                                           "audit_rel_rms")} for k, v in results.items()},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--slots", type=int, default=1 << 18,
                    help="trajectories in flight (throughput saturates from ~2^18 on the "
                         "protein pair)")
    ap.add_argument("--trajectories", type=int, default=10_000_000)
    ap.add_argument("--graph-steps", type=int, default=10)
    ap.add_argument("--paths", default="fp64,fp32,surrogate")
    ap.add_argument("--kon-trajectories", type=int, default=500_000)
    # the surrogate workload's 2e7 (pose, force) training pairs
    ap.add_argument("--train-poses", type=int, default=20_000_000)
    ap.add_argument("--train-epochs", type=int, default=64)
    ap.add_argument("--train-batch", type=int, default=16384)
    ap.add_argument("--train-lr", type=float, default=4e-3)
    ap.add_argument("--audit-n", type=int, default=64)
    ap.add_argument("--audit-tol", type=float, default=0.25)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-kon", action="store_true")
    ap.add_argument("--pmc", action="store_true", help="counter-collection mode (no JSON line)")
    ap.add_argument("--hydro", default=None, choices=("rpy12", "scalar"),
                    help="two-body RPY 12x12 mobility (the protein pair's default) or scalar "
                         "D_parallel (the protein-ligand system's default)")
    ap.add_argument("--train-from-trajectories", type=int, default=None,
                    help="training poses taken from the plain FP64 path's own trajectories "
                         "(default: 10,000,000 for protein_ligand, 0 for protein_protein)")
    ap.add_argument("--outer", default="return_prob", choices=("return_prob", "lmz"),
                    help="escape handling: the reference's return probability at r_esc, or "
                         "the LMZ outer propagator from 1.1 b")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the heuristic force launch instead of tuning it in the warm-up")
    ap.add_argument("--workload", default="protein_protein",
                    choices=("protein_protein", "protein_ligand", "multi_ligand"),
                    help="protein_protein (the headline), protein_ligand (with the surrogate as "
                         "its surrogate path) or multi_ligand (16 ligands, one launch)")
    ap.add_argument("--ligand-slots", type=int, default=8192, help="multi_ligand: slots per ligand")
    ap.add_argument("--ml-train-poses", type=int, default=1_000_000)
    ap.add_argument("--ml-train-epochs", type=int, default=24)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU in this build (asked for {args.gpus}); "
              "the multi-GPU design is trajectory sharding, documented but not launched.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(2)
    if args.hydro is None:
        args.hydro = "rpy12" if args.workload == "protein_protein" else "scalar"
    if args.train_from_trajectories is None:
        args.train_from_trajectories = 10_000_000 if args.workload == "protein_ligand" else 0
    if args.pmc:
        pmc_mode(args)
        return
    if args.workload == "multi_ligand":
        multi_ligand_bench(args)
        return
    from pystarc_amd import systems as S
    sysm, grids = build_system(args)
    paths = args.paths.split(",")
    engines = {p: S.make_engine(sysm, grids, precision="fp64" if p == "fp64" else "fp32")
               for p in ("fp64", "fp32")}
    surr = None
    traj_poses = 0
    results = {}
    # plain paths first: their live poses join the surrogate's training set
    order = [p for p in paths if p != "surrogate"] + [p for p in paths if p == "surrogate"]
    for p in order:
        if p == "surrogate":
            extra = None
            if results:
                extra = (np.concatenate([r["pos"] for r in results.values()]),
                         np.concatenate([r["quat"] for r in results.values()]))
            if args.train_from_trajectories > 0:
                mob = sysm.mob
                r_on = float(mob.radius1 + mob.radius2) + 8.0
                t0 = time.time()
                hp, hq = harvest_poses(sysm, engines["fp64"], args, args.train_from_trajectories,
                                       r_on)
                log(f"harvested {len(hp)} poses beyond {r_on:.1f} A from plain FP64 trajectories "
                    f"in {time.time() - t0:.1f}s")
                extra = (hp, hq) if extra is None else (np.concatenate([extra[0], hp]),
                                                        np.concatenate([extra[1], hq]))
                traj_poses = len(hp)
            surr = train_surrogate(sysm, engines["fp64"], args, extra)
        eng = engines["fp32" if p == "surrogate" else p]
        log(f"path {p}: warmup {args.warmup}, timed {args.steps} steps, {args.slots} slots")
        r = run_window(sysm, eng, args, surr if p == "surrogate" else None)
        log(f"path {p}: {r['value']:.4g} traj-steps/s, {r['device_ms_per_step']:.3f} ms/step device"
            + (f", audit rel rms {r['audit_rel_rms']:.3f}" if r["audit_rel_rms"] is not None else ""))
        log(f"path {p}: force kernel in the step {r['roof_force']['ms']:.3f} ms over "
            f"{r['roof_force']['census'][0]:.0f} trajectories, "
            f"{r['roof_force']['achieved_gbs']:.1f} GB/s algorithmic")
        if p == "surrogate":
            rr = np.linalg.norm(r["pos"], axis=1)
            r["roof_mlp"] = mlp_roofline(surr, r["pos"][rr > surr.r_on], r["quat"][rr > surr.r_on])
            log(f"path {p}: mlp {r['roof_mlp']['ms']:.3f} ms for {r['roof_mlp']['n']} poses "
                f"({r['roof_mlp']['tflops']:.1f} TFLOP/s)")
        results[p] = r
    kon = {}
    if not args.no_kon:
        for p in ("fp64", "surrogate"):
            if p in paths:
                kon[p] = full_run(sysm, engines["fp32" if p == "surrogate" else p], args,
                                  surr if p == "surrogate" else None)
                log(f"k_on {p}: P={kon[p]['P']:.5f} CI95 {kon[p]['P_ci95']}")
        if "fp64" in kon and "surrogate" in kon:
            lo, hi = kon["fp64"]["P_ci95"]
            kon["surrogate_in_plain_ci95"] = bool(lo <= kon["surrogate"]["P"] <= hi)
            kon["two_sample_z"] = two_sample_z(kon["surrogate"], kon["fp64"])
    head = "surrogate" if "surrogate" in results else max(results, key=lambda k: results[k]["value"])
    best = results[head]
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle) ...")
        cpu = cpu_baseline(sysm, grids, best["pos"], best["quat"], args.hydro, args)
        log(f"cpu baseline {cpu['value']:.4g} traj-steps/s on {cpu['cores']} cores "
            f"({cpu['cpu_model']}), {cpu['one_core_value']:.4g} on one")
    # gather ceiling of the receptor grids (the force kernel's bound below HBM)
    ceiling = gather_ceiling(engines["fp32"])
    log(f"gather ceiling: uniform {ceiling['uniform']['gbs']:.0f} GB/s, clustered "
        f"{ceiling['clustered']['gbs']:.0f} GB/s (algorithmic, 160 B per stencil)")
    # dominant kernel of the headline path: MFMA roofline of the MLP when it
    # takes longer than the force kernel on the near trajectories, else HBM
    if head == "surrogate" and (best["roof_force"] is None
                                or best["roof_mlp"]["ms"] >= best["roof_force"]["ms"]):
        rm = best["roof_mlp"]
        roof = {"bound": "mfma", "achieved": rm["tflops"], "peak": BF16_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": rm["tflops"] / BF16_PEAK_TFLOPS, "traffic": None,
                "kernel": "k_mlp_infer (fused 5-layer MLP forward)", "ms_per_launch": rm["ms"],
                "rows": rm["n"]}
    else:
        rf = best["roof_force"]
        lc = best.get("launch")
        tc = "double" if head == "fp64" else "float"
        zp = "true" if lc and lc.get("mode") in ("zpair", "split_zpair") else "false"
        if lc and lc.get("mode", "").startswith("split"):
            kname = f"k_forces_split<{tc}, {lc['G']}, {zp}> + k_rev_apply"
        else:
            kname = f"k_forces<{tc}, {lc['G']}, true, {zp}>" if lc else "k_forces"
        roof = {"bound": "hbm", "achieved": rf["achieved_gbs"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": rf["achieved_gbs"] / HBM_PEAK_GBS,
                "traffic": measured_traffic(args.workload, head, args.slots),
                "kernel": kname, "ms_per_launch": rf["ms"], "launches": rf["launches"],
                "census": rf["census"], "algorithmic_bytes": rf["bytes"],
                "gather_ceiling_gbs": ceiling["clustered"]["gbs"],
                "frac_of_gather_ceiling": rf["achieved_gbs"] / ceiling["clustered"]["gbs"],
                "limiter": measured_limiter(args.workload, head, args.slots)}
    plain = results.get("fp64")
    out = {
        "metric": METRIC,
        "value": best["value"],
        "unit": "trajectory-steps/s",
        "n_gpus": 1,
        "steps": best["steps"],
        "warmup": args.warmup,
        "ms_per_step": best["wall_s"] * 1e3 / max(best["steps"], 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"surrogate": "f64 state, f32 grid forces, bf16 MLP (f32 accumulate)",
                  "fp32": "f64 state, f32 forces", "fp64": "f64"}[head],
        "data": "synthetic",
        "config": {"workload": workload_text(args),
                   "path": head, "slots": args.slots, "graph_steps": args.graph_steps,
                   "hydro": args.hydro, "outer": args.outer, "launch": best.get("launch")},
        "speedup_vs_plain_fp64": (best["value"] / plain["value"]) if plain else None,
        "roofline": roof,
        "cpu_baseline": cpu,
        "k_on": kon,
        "surrogate": None if surr is None else {
            "validation_force_rel_rms": surr.val_error,
            "validation_torque_rel_rms": getattr(surr, "val_torque_error", None),
            "audit_tol": args.audit_tol, "r_on": surr.r_on, "train_poses": args.train_poses,
            "trajectory_poses": traj_poses, "epochs": args.train_epochs,
            "k_on_shift": (kon["surrogate"]["P"] / kon["fp64"]["P"] - 1.0)
            if "surrogate" in kon and "fp64" in kon and kon["fp64"]["P"] > 0 else None},
        "gather_ceiling": ceiling,
        "paths": {k: {"value": v["value"], "device_ms_per_step": v["device_ms_per_step"],
                      "wall_ms_per_step": v["wall_s"] * 1e3 / max(v["steps"], 1),
                      "tune_steps": v["tune_steps"], "warmup_steps_run": v["warmup_steps_run"],
                      "audit_rel_rms": v["audit_rel_rms"],
                      "force_ms": v["roof_force"]["ms"] if v.get("roof_force") else None,
                      "force_gbs": v["roof_force"]["achieved_gbs"] if v.get("roof_force") else None,
                      "mlp": v.get("roof_mlp"), "launch": v.get("launch"),
                      "tune_fastest": v.get("tune_fastest")}
                  for k, v in results.items()},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
