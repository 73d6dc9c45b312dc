"""Host-side parts of bench.py's contract (no GPU): the multi-GPU refusal,
the committed HBM-traffic summary it reads for `roofline.traffic`, and the
arithmetic of that summary (FETCH_SIZE scaled by the copy calibration plus
WRITE_SIZE, in KiB per dispatch)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TRAFFIC = os.path.join(ROOT, "profiles", "k_forces_traffic.json")


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""      # no JSON line


def test_traffic_summary_covers_every_path():
    sys.path.insert(0, ROOT)
    import bench
    slots = json.load(open(TRAFFIC))["slots"]
    for path in ("surrogate", "fp32", "fp64"):
        v = bench.measured_traffic("protein_protein", path, slots)
        assert isinstance(v, int) and v > 0
    assert bench.measured_traffic("protein_protein", "no_such_path", slots) is None
    # a summary measured on another workload or slot count is not reported
    assert bench.measured_traffic("protein_ligand", "surrogate", slots) is None
    assert bench.measured_traffic("protein_protein", "surrogate", slots // 4) is None


@pytest.mark.parametrize("path", ["surrogate", "fp32", "fp64"])
def test_traffic_bytes_follow_from_the_counters(path):
    d = json.load(open(TRAFFIC))
    e = d["paths"][path]
    # the copy calibration: FETCH_SIZE reads half of a 262,144 KiB copy's
    # reads, WRITE_SIZE all of its writes
    expect = (e["fetch_kib_per_eval"] * 2 + e["write_kib_per_eval"]) * 1024
    assert abs(e["hbm_bytes_per_eval"] - expect) <= 1024
    assert e["launches"] == 20
    assert all("rocprofv3 --pmc" in c and "bench.py --pmc" in c for c in e["commands"])


def test_cache_counter_summary_follows_from_its_counters():
    # the derived figures of the cache-counter summary (and the limiter the
    # bench line carries) follow from the per-dispatch counters
    sys.path.insert(0, ROOT)
    import bench
    d = json.load(open(TRAFFIC))
    cc = d["cache_counters"]
    c, dv = cc["per_dispatch"], cc["derived"]
    stencils = c["census"][1] + c["census"][3]
    assert dv["l1_tag_lookups_per_stencil"] == pytest.approx(
        c["TCP_TOTAL_CACHE_ACCESSES_sum"] / stencils, abs=0.05)
    assert dv["l1_hit_rate"] == pytest.approx(
        1 - c["TCP_TCC_READ_REQ_sum"] / c["TCP_TOTAL_CACHE_ACCESSES_sum"], abs=0.002)
    assert dv["l2_hit_rate"] == pytest.approx(
        c["TCC_HIT_sum"] / (c["TCC_HIT_sum"] + c["TCC_MISS_sum"]), abs=0.002)
    cycles = c["GRBM_GUI_ACTIVE"] / 8            # per XCD
    assert dv["ta_busy_frac"] == pytest.approx(c["TA_TA_BUSY_sum"] / 256 / cycles, abs=0.02)
    assert dv["td_busy_frac"] == pytest.approx(c["TD_TD_BUSY_sum"] / 256 / cycles, abs=0.02)
    assert dv["l2_requests_per_clock_per_xcd"] == pytest.approx(
        c["TCP_TCC_READ_REQ_sum"] / 8 / cycles, abs=0.1)
    lim = bench.measured_limiter("protein_protein", cc["path"], d["slots"])
    assert lim["l1_hit_rate"] == dv["l1_hit_rate"]
    assert bench.measured_limiter("protein_protein", "fp64", d["slots"]) is None


def test_cpu_model_and_workload_text():
    sys.path.insert(0, ROOT)
    import argparse
    import bench
    assert isinstance(bench._cpu_model(), str) and bench._cpu_model()
    a = argparse.Namespace(workload="protein_ligand", hydro="scalar")
    assert "3220/18" in bench.workload_text(a)
    a = argparse.Namespace(workload="protein_protein", hydro="rpy12")
    assert "1700/1400" in bench.workload_text(a) and "RPY" in bench.workload_text(a)


def test_two_sample_z():
    sys.path.insert(0, ROOT)
    import bench
    a = {"reacted": 20, "escaped": 80, "P": 0.2}
    b = {"reacted": 10, "escaped": 90, "P": 0.1}
    # pooled p = 0.15, se = sqrt(0.15 * 0.85 * (1/100 + 1/100)) = 0.050498
    assert abs(bench.two_sample_z(a, b) - 0.1 / 0.0504975) < 1e-4
    assert abs(bench.two_sample_z(b, a) + bench.two_sample_z(a, b)) < 1e-12
    assert bench.two_sample_z(a, a) == 0.0
