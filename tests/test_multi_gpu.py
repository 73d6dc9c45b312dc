"""Splitting a run into parts and pooling them (pystarc/multi_GPU/
multi_GPU_runs.py, combine_data.py): ranges, seeds, pooled counts and
Wilson interval, summed per-reaction fires, observables, and the same
pooling as one all-reduce over a two-rank gloo group on the CPU."""
import math
import os
import socket

import numpy as np
import pytest

from pystarc_amd.pipeline.multi_gpu import (allreduce_counts, combine_results, combine_sim_data,
                                            split_seeds, split_trajectories)
from pystarc_amd.pipeline.output_writer import results_dict
from pystarc_amd.simulation.gpu_batch_simulator import GPUBatchResult


def test_split_trajectories_cover_the_run():
    for n, k in ((10, 3), (1_000_000, 8), (7, 7)):
        r = split_trajectories(n, k)
        assert len(r) == k and r[0][0] == 0
        assert all(a[0] + a[1] == b[0] for a, b in zip(r, r[1:]))
        assert r[-1][0] + r[-1][1] == n
        assert max(c for _, c in r) - min(c for _, c in r) <= 1
    with pytest.raises(ValueError):
        split_trajectories(3, 4)


def test_split_seeds_follow_the_reference():
    # multi_GPU_runs.py:71-84: per_split = total // n, seed = base + i * 11111111, i = 1..n
    assert split_seeds(100_000, 1, 4) == [(25_000, 11111112), (25_000, 22222223),
                                          (25_000, 33333334), (25_000, 44444445)]


def _res(nr, ne, nm, wall=2.0, sps=1e6, k_b=57.47, rx=None):
    return GPUBatchResult(n_trajectories=nr + ne + nm, n_reacted=nr, n_escaped=ne, n_max_steps=nm,
                          reaction_counts={"rxn": nr}, r_start=30.0, r_escape=60.0, dt=0.2,
                          elapsed_sec=wall, steps_per_sec=sps, total_steps=int(sps * wall),
                          k_b=k_b, sim_data={"completed_reactions": rx or []})


def test_combined_equals_the_pooled_single_run():
    parts = [_res(40, 460, 3), _res(55, 440, 1, wall=3.0), _res(47, 452, 0)]
    c = combine_results(parts)
    one = results_dict(_res(142, 1352, 4), 0.434, 57.47)
    for k in ("P_rxn", "k_on"):
        assert c[k] == pytest.approx(one[k], rel=1e-12), k
    # the combine step uses z = 1.96 (combine_data.py:84), a single run's
    # interval the exact normal quantile 1.95996...: bounds agree to ~2e-5
    for k in ("k_on_low", "k_on_high", "P_rxn_low", "P_rxn_high"):
        assert c[k] == pytest.approx(one[k], rel=1e-4), k
    assert (c["n_reacted"], c["n_escaped"], c["n_max_steps"], c["n_trajectories"]) == \
        (142, 1352, 4, 1498)
    # Wilson 95 % interval written out (combine_data.py:85-92)
    N, P, z = 1494, 142 / 1494, 1.96
    d = 1 + z * z / N
    ctr = (P + z * z / (2 * N)) / d
    sp = z * math.sqrt(P * (1 - P) / N + z * z / (4 * N * N)) / d
    assert c["P_rxn_low"] == pytest.approx(ctr - sp, rel=1e-14)
    assert c["P_rxn_high"] == pytest.approx(ctr + sp, rel=1e-14)
    assert c["wall_time_sec"] == pytest.approx(7.0)
    assert c["steps_per_sec"] == pytest.approx(1e6)


def test_combined_reaction_fires_and_observables():
    rx1 = [{"name": "a", "n_fired": 3, "state_before": "s0", "state_after": "s1"},
           {"name": "b", "n_fired": 1, "state_before": "s1", "state_after": "done"}]
    rx2 = [{"name": "a", "n_fired": 4, "state_before": "s0", "state_after": "s1"}]
    c = combine_results([_res(1, 9, 0, rx=rx1), _res(2, 8, 0, rx=rx2)])
    assert c["completed_reactions"] == [
        {"name": "a", "n_fired": 7, "state_before": "s0", "state_after": "s1"},
        {"name": "b", "n_fired": 1, "state_before": "s1", "state_after": "done"}]
    a = {"outcome": np.array([1, 2, 2]), "radial_counts": np.array([1, 2, 3]),
         "encounter_traj": np.array([0, 2]), "encounter_pos": np.zeros((2, 3)),
         "n_returns_total": 4, "radial_bins": np.arange(4)}
    b = {"outcome": np.array([2, 1]), "radial_counts": np.array([5, 0, 1]),
         "encounter_traj": np.array([1]), "encounter_pos": np.ones((1, 3)),
         "n_returns_total": 1, "radial_bins": np.arange(4)}
    s = combine_sim_data([a, b])
    assert s["radial_counts"].tolist() == [6, 2, 4] and s["n_returns_total"] == 5
    assert s["outcome"].tolist() == [1, 2, 2, 2, 1]
    assert s["encounter_traj"].tolist() == [0, 2, 4]
    assert s["encounter_pos"].shape == (3, 3)


def _rank(rank, world, port, out):
    import torch.distributed as dist
    dist.init_process_group("gloo", init_method=f"tcp://127.0.0.1:{port}", rank=rank,
                            world_size=world)
    try:
        part = [_res(40, 460, 3), _res(55, 440, 1, wall=3.0)][rank]
        c = allreduce_counts(part)
        out.put((rank, c["n_reacted"], c["n_escaped"], c["n_max_steps"], c["P_rxn"],
                 c["k_on_low"], c["k_on_high"]))
    finally:
        dist.destroy_process_group()


def test_allreduce_over_two_gloo_ranks():
    import torch.multiprocessing as mp
    with socket.socket() as s:
        s.bind(("127.0.0.1", 0))
        port = s.getsockname()[1]
    ctx = mp.get_context("spawn")
    q = ctx.Queue()
    procs = [ctx.Process(target=_rank, args=(r, 2, port, q)) for r in range(2)]
    for p in procs:
        p.start()
    got = sorted(q.get(timeout=120) for _ in procs)
    for p in procs:
        p.join(timeout=60)
        assert p.exitcode == 0
    ref = combine_results([_res(40, 460, 3), _res(55, 440, 1, wall=3.0)])
    for r in got:
        assert r[1:4] == (95, 900, 4)
        assert r[4] == pytest.approx(ref["P_rxn"], rel=1e-15)
        assert (r[5], r[6]) == pytest.approx((ref["k_on_low"], ref["k_on_high"]), rel=1e-15)
