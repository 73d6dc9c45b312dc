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
    for path in ("surrogate", "fp32", "fp64"):
        v = bench.measured_traffic("protein_protein", path)
        assert isinstance(v, int) and v > 0
    assert bench.measured_traffic("protein_protein", "no_such_path") is None
    # a summary measured on another workload is not reported for this one
    assert bench.measured_traffic("protein_ligand", "surrogate") is None


@pytest.mark.parametrize("path", ["surrogate", "fp32", "fp64"])
def test_traffic_bytes_follow_from_the_counters(path):
    d = json.load(open(TRAFFIC))
    e = d["paths"][path]
    # the copy calibration: FETCH_SIZE reads half of a 262,144 KiB copy's
    # reads, WRITE_SIZE all of its writes
    expect = (e["fetch_kib_per_dispatch"] * 2 + e["write_kib_per_dispatch"]) * 1024
    assert abs(e["hbm_bytes_per_launch"] - expect) <= 1024
    assert e["launches"] == 20
    assert all("rocprofv3 --pmc" in c and "bench.py --pmc" in c for c in e["commands"])


def test_surrogate_traffic_exceeds_its_algorithmic_bytes():
    # algorithmic bytes of one surrogate-window launch: ~4,100 trajectories x
    # 104 B plus 80 B per atom-in-grid (5.6 M + 5.6 M + 6.8 M); line traffic
    # cannot be below it
    e = json.load(open(TRAFFIC))["paths"]["surrogate"]
    alg = 4100 * 104 + (5.6e6 + 5.6e6 + 6.8e6) * 80
    assert e["hbm_bytes_per_launch"] > alg


def test_cpu_model_and_workload_text():
    sys.path.insert(0, ROOT)
    import argparse
    import bench
    assert isinstance(bench._cpu_model(), str) and bench._cpu_model()
    a = argparse.Namespace(workload="protein_ligand", hydro="scalar")
    assert "3220/18" in bench.workload_text(a)
    a = argparse.Namespace(workload="protein_protein", hydro="rpy12")
    assert "1700/1400" in bench.workload_text(a) and "RPY" in bench.workload_text(a)
