"""Host-side run control of the production loop (no device): the reference's
stall rule and its check steps, snapshot steps, convergence milestones, the
per-trajectory step bound, the output configuration, the launch-tuner cache
and the checkpoint file contract.

Reference: pystarc/simulation/gpu_batch_simulator.py:1342-1356 (99.5 % stop
and stall rule), :753-761 (milestones), :578-581 and :1318-1340 (snapshots),
:548-575 / :1400-1420 / :1505-1510 (checkpoint).
"""
import json
from types import SimpleNamespace

import numpy as np
import pytest

from pystarc_amd.simulation import gpu_batch_simulator as G


def reference_stall(N, done, stall_done):
    # the reference's expression, spelled out on its own terms
    return done > int(N * 0.95) and done - stall_done < max(2, int(N * 0.001))


@pytest.mark.parametrize("N,done,before,expect", [
    (1000, 951, 950, True),      # 1 finished since the last check, fewer than 2
    (1000, 951, 949, False),     # 2 finished: not a stall
    (1000, 950, 950, False),     # not more than 95 % done
    (100000, 96000, 95950, True),   # 50 < 0.1 % of N = 100
    (100000, 96000, 95900, False),  # 100 is not fewer than 100
])
def test_stall_rule(N, done, before, expect):
    assert G.stall_stop(N, done, before) is expect
    assert G.stall_stop(N, done, before) == reference_stall(N, done, before)


def test_stall_check_steps_follow_the_reference_loop_index():
    # the reference checks after loop index 10000, 20000, ... i.e. after
    # 10001, 20001, ... moves; never after the first
    checks = [s for s in range(1, 40003) if G.is_stall_check(s)]
    assert checks == [10001, 20001, 30001, 40001]
    assert G.next_stall_check(0) == 10001
    assert G.next_stall_check(10001) == 20001
    assert G.next_stall_check(12345) == 20001
    assert [s for s in range(1, 12) if G.is_stall_check(s, 3)] == [4, 7, 10]


def test_snapshot_steps_follow_the_reference_loop_index():
    # loop index 0, 10, 20 (after moves 1, 11, 21) with save_interval 10
    got, s = [], 0
    while s < 40:
        s = G.next_snapshot(s, 10)
        got.append(s)
    assert got == [1, 11, 21, 31, 41]
    assert G.next_snapshot(0, 1) == 1 and G.next_snapshot(1, 1) == 2


def test_convergence_milestones():
    assert G.convergence_milestones(1000, 10) == [100, 200, 300, 400, 500, 600, 700, 800, 900]
    assert G.convergence_milestones(1000, 25) == [250, 500, 750]
    assert G.convergence_milestones(1000, 0) == []
    assert G.convergence_milestones(7, 50) == [3]


def test_device_step_bound_gives_every_trajectory_its_budget():
    # 2,000 trajectories in 256 slots: 8 waves; the last trajectory can start
    # no later than 8 * max_steps and needs max_steps more
    assert G.device_step_bound(50, 2000, 256) == 50 * 9
    assert G.device_step_bound(50, 256, 256) == 100
    # a greedy refill simulation with every trajectory taking max_steps
    # finishes within the bound
    N, S, m = 1000, 64, 7
    free_at = np.zeros(S, dtype=int)
    for _ in range(N):
        k = int(np.argmin(free_at))
        free_at[k] += m
    assert free_at.max() <= G.device_step_bound(m, N, S)


def test_snapshot_config():
    assert G.snapshot_config(SimpleNamespace()) == (10, True, True)
    cfg = SimpleNamespace(save_interval=5, full_paths=False, energetics=True)
    assert G.snapshot_config(SimpleNamespace(_output_cfg=cfg)) == (5, False, True)
    cfg0 = SimpleNamespace(save_interval=0)
    assert G.snapshot_config(SimpleNamespace(_output_cfg=cfg0))[0] == 1


def test_tune_cache_roundtrip(tmp_path, monkeypatch):
    f = tmp_path / "tune.json"
    monkeypatch.setenv("PYSTARC_AMD_TUNE_CACHE", str(f))
    monkeypatch.setattr(G, "_TUNE_CACHE", {})
    best = {"mode": "zpair", "G": 64, "cap": 2048, "xcd": 1}
    assert G.tune_cache_get("k1") is None
    G.tune_cache_put("k1", best)
    assert json.loads(f.read_text())["k1"] == best
    monkeypatch.setattr(G, "_TUNE_CACHE", {})   # a new process reads the file
    assert G.tune_cache_get("k1") == best


def test_checkpoint_of_another_run_is_ignored(tmp_path):
    ck = G.Checkpoint(tmp_path, "abc")
    np.savez_compressed(str(ck.path), state=np.zeros(8, np.uint8), fingerprint=np.array("other"))
    assert ck.exists()
    # the fingerprint is checked before the library is touched
    assert ck.load(None, None, None) is None
    assert ck.exists()
    (tmp_path / "checkpoint.npz").write_bytes(b"not an npz")
    assert ck.load(None, None, None) is None
    ck.remove()
    assert not ck.exists()


def test_checkpoint_host_entries_roundtrip():
    sim = G.GPUBatchSimulator.__new__(G.GPUBatchSimulator)
    audit = {"err2": 1.0, "ref2": 2.0, "n": 3, "worst": 0.1, "terr2": 0.5, "tref2": 4.0,
             "worst_torque": 0.2, "fallback_step": None}
    paths = [np.arange(16.0).reshape(2, 8), np.arange(8.0).reshape(1, 8)]
    conv = [dict(zip(G._CONV_KEYS, [1.0, 1.0, 10.0, 0.5, 1e9, 1.0, 1.0, 5.0, 0.1]))]
    host = sim._ckpt_host(11, 3, 4, 5, audit, paths, [], conv)
    assert host["loop"].tolist() == [11, 3, 4, 5]
    assert host["audit"][7] == -1
    back = G._split_rows(host["paths"], host["paths_n"])
    assert len(back) == 2 and np.array_equal(back[0], paths[0]) and np.array_equal(back[1], paths[1])
    assert dict(zip(G._CONV_KEYS, host["conv"][0].tolist())) == conv[0]


def test_audit_window_pools_small_chunks_before_deciding():
    w = G.AuditWindow(min_poses=100)
    # tail chunks: 10 poses each with a large error, no decision until 100 poses
    for _ in range(9):
        assert w.add(err2=4.0, ref2=10.0, n=10) is None
    err, terr = w.add(err2=4.0, ref2=10.0, n=10)
    # pooled error sqrt(40 / 100), not any single chunk's
    assert err == pytest.approx(np.sqrt(0.4)) and terr == 0.0
    # the window starts over after a decision
    assert w.add(err2=1.0, ref2=100.0, n=50) is None
    err, terr = w.add(err2=1.0, ref2=100.0, n=50, terr2=9.0, tref2=100.0)
    assert err == pytest.approx(np.sqrt(2.0 / 200.0)) and terr == pytest.approx(0.3)


def test_audit_window_one_large_chunk_decides_at_once():
    # a steady-state chunk (640 steps x 64 audited poses) is a window by itself
    err, _ = G.AuditWindow().add(err2=1.0, ref2=100.0, n=640 * 64)
    assert err == pytest.approx(0.1)


def test_audit_window_resumes_from_its_sums():
    a = G.AuditWindow(min_poses=100)
    a.add(2.0, 10.0, 60)
    b = G.AuditWindow(min_poses=100, sums=a.sums)
    assert a.add(2.0, 10.0, 60) == b.add(2.0, 10.0, 60)
