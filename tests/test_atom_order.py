"""Host-side atom orders of the device arrays (csrc/pst_order.hpp), built
with the host C++ compiler and checked on the CPU: every order is a
permutation; points on a line come out in runs of consecutive positions
(what bisection along the only extended axis must give); runs of the
synthetic protein clouds are compact; receptor run spheres bound their
atoms; small inputs keep their order; ties do not break determinism."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "pystarc_amd", "csrc")

DRIVER = r"""
#include <cstdio>
#include <vector>
#include "pst_order.hpp"
int main() {
  char mode[8];
  int n;
  if (scanf("%7s %d", mode, &n) != 2) return 1;
  std::vector<double> p(3 * (size_t)n);
  for (auto& x : p) if (scanf("%lf", &x) != 1) return 1;
  std::vector<int> order;
  std::vector<double> spheres;
  if (mode[0] == 'l') pst::ligand_order(n, p.data(), order);
  else pst::rec_chunk_order(n, p.data(), order, spheres);
  printf("%zu %zu\n", order.size(), spheres.size());
  for (int v : order) printf("%d\n", v);
  for (double v : spheres) printf("%.17g\n", v);
  return 0;
}
"""


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    d = tmp_path_factory.mktemp("order")
    src = d / "drv.cpp"
    src.write_text(DRIVER)
    exe = d / "drv"
    subprocess.run([cxx, "-std=c++17", "-O2", "-I", CSRC, str(src), "-o", str(exe)], check=True)

    def run(mode, pos):
        pos = np.asarray(pos, dtype=np.float64)
        inp = f"{mode} {len(pos)}\n" + "\n".join(" ".join(repr(float(x)) for x in r) for r in pos)
        out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True)
        vals = out.stdout.split()
        no, ns = int(vals[0]), int(vals[1])
        order = np.array([int(v) for v in vals[2:2 + no]], dtype=np.int64)
        spheres = np.array([float(v) for v in vals[2 + no:2 + no + ns]]).reshape(-1, 4)
        return order, spheres
    return run


def is_perm(order, n):
    return len(order) == n and np.array_equal(np.sort(order), np.arange(n))


def run_radius(pos, order, run):
    r = []
    for k in range(0, len(order) - run + 1, run):
        p = pos[order[k:k + run]]
        r.append(np.linalg.norm(p - p.mean(0), axis=1).max())
    return float(np.mean(r))


@pytest.mark.parametrize("n", [65, 200, 1000, 1400])
def test_points_on_a_line_form_runs_of_consecutive_positions(driver, n):
    rng = np.random.default_rng(n)
    x = rng.permutation(n).astype(np.float64)        # distinct positions 0..n-1, shuffled
    pos = np.stack([x, np.zeros(n), np.zeros(n)], axis=1)
    order, _ = driver("lig", pos)
    assert is_perm(order, n)
    xs = x[order]
    # splits fall at multiples of 64 (then 32, 16) from the start, so every
    # full aligned run is one span of the bisection: a set of consecutive x
    for run in (64, 32, 16):
        for k in range(0, (n // run) * run, run):
            seg = xs[k:k + run]
            assert seg.max() - seg.min() == run - 1, (run, k)


def test_protein_cloud_runs_are_compact(driver):
    from pystarc_amd import systems as S
    sysm = S.protein_protein(n_traj=10)
    lig = sysm.mol2.positions_array() - sysm.mol2.centroid()
    order, _ = driver("lig", lig)
    assert is_perm(order, len(lig))
    ident = np.arange(len(lig))
    for run in (64, 32):
        assert run_radius(lig, order, run) < 0.7 * run_radius(lig, ident, run)
    # each aligned 32-run is at most as spread as the 64-run that holds it
    assert run_radius(lig, order, 32) < run_radius(lig, order, 64)


def test_receptor_runs_are_bounded_by_their_spheres(driver):
    from pystarc_amd import systems as S
    sysm = S.protein_protein(n_traj=10)
    rec = sysm.mol1.positions_array()
    order, sph = driver("rec", rec)
    n = len(rec)
    assert is_perm(order, n)
    assert len(sph) == (n + 63) // 64
    for k, (cx, cy, cz, r) in enumerate(sph):
        p = rec[order[64 * k:min(n, 64 * (k + 1))]]
        assert np.all(np.linalg.norm(p - [cx, cy, cz], axis=1) <= r)
    assert run_radius(rec, order, 32) < 0.7 * run_radius(rec, np.arange(n), 32)


@pytest.mark.parametrize("mode,n", [("lig", 1), ("lig", 18), ("lig", 64), ("rec", 40), ("rec", 64)])
def test_small_inputs_keep_their_order(driver, mode, n):
    pos = np.random.default_rng(n).normal(size=(n, 3))
    order, sph = driver(mode, pos)
    assert np.array_equal(order, np.arange(n))
    if mode == "rec":
        assert len(sph) == 1


def test_ties_are_deterministic(driver):
    pos = np.zeros((300, 3))
    pos[::3, 0] = 1.0                                 # two coincident clusters
    a, _ = driver("lig", pos)
    b, _ = driver("lig", pos)
    assert is_perm(a, 300) and np.array_equal(a, b)
    xs = pos[a, 0]
    # splits along x order the aligned 16-atom blocks (the bisection's leaves
    # keep no order inside): no block holds an x above the next block's
    blocks = [xs[k:k + 16] for k in range(0, 300, 16)]
    assert all(b0.max() <= b1.min() for b0, b1 in zip(blocks, blocks[1:]))
