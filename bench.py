"""Benchmark: trajectory-steps/s of the batched BD step on one MI355X.

Workload: the 1,700 / 1,400-atom protein-protein pair (193^3 FP32 receptor
potential + Born grids and ligand Born grid at 0.5 A, bidirectional Born,
scalar HI, overlap rejection, 2-of-5 contact reaction, b = 60 A).  A "step" is
one BD step of every live trajectory slot; value = trajectory-steps per second
over the timed steps (inputs resident in HBM).  Synthetic data from seeded
generators; random-free physics identical to the reference's step.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--slots S] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "trajectory-steps/sec on 1 MI355X, surrogate vs plain at equal k_on (Wilson 95%)"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_system(args):
    from pystarc_amd import systems as S
    from pystarc_amd._lib import Context
    sysm = S.protein_protein(n_traj=args.trajectories, seed=3013)
    t0 = time.time()
    ctx = Context(0)
    grids = S.build_grids(sysm, ctx)
    ctx.close()
    log(f"grids built in {time.time() - t0:.1f}s")
    return sysm, grids


def run_path(sysm, grids, precision, args):
    """Production run: warm-up steps, then exactly `steps` timed steps."""
    import ctypes as C
    import torch
    from pystarc_amd import systems as S
    from pystarc_amd._lib import ptr
    from pystarc_amd.simulation.gpu_batch_simulator import GPUBatchSimulator
    eng = S.make_engine(sysm, grids, precision=precision)
    sim = GPUBatchSimulator(sysm.mol1, sysm.mol2, sysm.mob, sysm.pathway_set, sysm.params, eng,
                            n_slots=args.slots, graph_steps=args.graph_steps)
    sim._return_prob = sim._compute_return_prob()
    P = sim._step_params()
    lib, ctx = eng.ctx.lib, eng.ctx
    h = sim._make_sim(P)
    c16 = np.zeros(16)
    ctx.check(lib.pst_prod_init(h, 0), "pst_prod_init")
    if args.warmup:
        ctx.check(lib.pst_prod_run(h, args.warmup, args.graph_steps, ptr(c16)), "warmup")
    base_steps = c16[3]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.check(lib.pst_prod_run(h, args.steps, args.graph_steps, ptr(c16)), "timed")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = C.c_double()
    lib.pst_prod_last_ms_per_step(h, C.byref(ms))
    steps_run = int(c16[8])
    traj_steps = c16[3] - base_steps
    # snapshot of the live state for the kernel roofline and the CPU baseline
    S_ = sim.n_slots
    pos = np.zeros((S_, 3)); q = np.zeros((S_, 4)); st = np.zeros(S_, dtype=np.int32)
    ctx.check(lib.pst_sim_get_state(h, ptr(pos), ptr(q), ptr(st), None), "get_state")
    lib.pst_sim_destroy(h)
    live = st == 0
    return {"eng": eng, "wall_s": wall, "device_ms_per_step": ms.value, "steps": steps_run,
            "traj_steps": traj_steps, "value": traj_steps / wall if wall > 0 else 0.0,
            "pos": pos[live], "quat": q[live], "counters": c16.tolist()}


def force_roofline(eng, pos, quat, precision, reps=20):
    """Average duration of the fused force kernel on the live poses (HIP events
    on the context stream) and its algorithmic bytes from an atom census."""
    import ctypes as C
    import torch
    from pystarc_amd._lib import ptr
    body = eng._lig_key is not None
    assert body
    n = len(pos)
    dp = torch.as_tensor(np.ascontiguousarray(pos), device=eng.device)
    dq = torch.as_tensor(np.ascontiguousarray(quat), device=eng.device)
    F = torch.zeros((n, 3), dtype=torch.float64, device=eng.device)
    T = torch.zeros_like(F)
    ms = C.c_double()
    census = np.zeros(4)
    eng.ctx.check(eng.ctx.lib.pst_time_forces_pose(
        eng.ctx.h, n, ptr(dp), ptr(dq), ptr(F), ptr(T), None, 0 if precision == "fp64" else 1,
        reps, C.byref(ms), ptr(census)), "pst_time_forces_pose")
    grid_bytes = 4  # FP32 grids
    # per trajectory: pose in (24 + 32 B) and force + torque out (48 B);
    # per in-grid atom and grid: 20 distinct stencil nodes
    alg = census[0] * (56 + 48) + (census[1] + census[2] + census[3]) * 20 * grid_bytes
    sec = ms.value * 1e-3
    return {"ms": ms.value, "bytes": alg, "census": census.tolist(),
            "achieved_gbs": alg / sec / 1e9 if sec > 0 else 0.0}


def cpu_baseline(sysm, grids, pos, quat, seconds=15.0, n_traj=128):
    """The NumPy FP64 oracle of the same step on a bounded sample: n_traj live
    poses taken from the GPU run, stepped until ~`seconds` of CPU time."""
    from oracle.simulator import OracleSim
    from tests._systems import oracle_params, oracle_system
    np_grids = {k: [(np.asarray(g.data), np.asarray(g.origin), float(g.delta[0, 0])) for g in v]
                for k, v in grids.items()}
    sel = np.random.default_rng(0).choice(len(pos), size=min(n_traj, len(pos)), replace=False)
    ora = OracleSim(oracle_system(sysm, np_grids), oracle_params(sysm, n_trajectories=len(sel)))
    ora.pos[:] = pos[sel]
    ora.q[:] = quat[sel]
    t0 = time.process_time()
    w0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - w0 < seconds:
        before = ora.total_steps
        if not ora.step():
            break
        steps += ora.total_steps - before
    wall = time.perf_counter() - w0
    return {"value": steps / wall if wall > 0 else 0.0, "unit": "trajectory-steps/s", "cores": 1,
            "kind": "port",
            "sample": f"{len(sel)} live poses from the GPU run stepped {ora.step_index} times "
                      f"({steps} trajectory-steps, {wall:.1f}s, single-threaded NumPy FP64 oracle)",
            "cpu_seconds": time.process_time() - t0}


def pmc_mode(args):
    """For rocprofv3 --pmc runs: a few production steps, `--pmc-launches` fused
    force launches on the live poses, then one plain device copy of 256 MiB
    whose FETCH_SIZE / WRITE_SIZE calibrate the counters' byte scale."""
    import torch
    sysm, grids = build_system(args)
    prec = args.paths.split(",")[0]
    r = run_path(sysm, grids, prec, args)
    force_roofline(r["eng"], r["pos"], r["quat"], prec, reps=args.pmc_launches)
    src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"pmc mode done: {len(r['pos'])} live poses, copy of {src.numel() * 4} bytes")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pmc", action="store_true", help="counter-collection mode (no JSON line)")
    ap.add_argument("--pmc-launches", type=int, default=3)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--slots", type=int, default=1 << 17)
    ap.add_argument("--trajectories", type=int, default=10_000_000)
    ap.add_argument("--graph-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--paths", default="fp64,fp32")
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
    if args.pmc:
        pmc_mode(args)
        return
    sysm, grids = build_system(args)
    results = {}
    for prec in args.paths.split(","):
        log(f"path {prec}: warmup {args.warmup}, timed {args.steps} steps, {args.slots} slots")
        r = run_path(sysm, grids, prec, args)
        log(f"path {prec}: {r['value']:.4g} traj-steps/s, {r['device_ms_per_step']:.3f} ms/step device")
        rf = force_roofline(r["eng"], r["pos"], r["quat"], prec)
        log(f"path {prec}: force kernel {rf['ms']:.3f} ms, {rf['achieved_gbs']:.1f} GB/s algorithmic")
        r["roof"] = rf
        results[prec] = r
    best_key = max(results, key=lambda k: results[k]["value"])
    best = results[best_key]
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle) ...")
        cpu = cpu_baseline(sysm, grids, best["pos"], best["quat"])
        log(f"cpu baseline {cpu['value']:.4g} traj-steps/s")
    rf = best["roof"]
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
        "dtype": "f64 state, " + ("f64" if best_key == "fp64" else "f32") + " forces",
        "data": "synthetic",
        "config": {"workload": "protein-protein 1700/1400 atoms, 193^3 FP32 grids (receptor "
                               "potential+Born, ligand Born), bidirectional Born, scalar HI, "
                               "b=60 A, 2-of-5 contacts < 5.5 A",
                   "path": best_key, "slots": args.slots, "graph_steps": args.graph_steps},
        "roofline": {"bound": "hbm", "achieved": rf["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": rf["achieved_gbs"] / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_forces", "ms_per_launch": rf["ms"],
                     "census": rf["census"]},
        "cpu_baseline": cpu,
        "paths": {k: {"value": v["value"], "device_ms_per_step": v["device_ms_per_step"],
                      "force_ms": v["roof"]["ms"], "force_gbs": v["roof"]["achieved_gbs"],
                      "counters": v["counters"]} for k, v in results.items()},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
