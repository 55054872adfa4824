"""CPU: bench.py's command-line contract and its seeded workload generators
(no GPU work is started)."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_workload_table_matches_the_configurations():
    import bench

    w = bench.WORKLOADS
    assert w["ch3d"]["shape"] == (256, 256, 256) and w["ch3d"]["prec"] == "f64" and w["ch3d"]["steps"] == 2000
    assert w["ch3d"]["params"] == [1.0, 1.0, 0.005] and w["ch3d"]["bytes_per_cell"] == 16
    assert w["ch3d_f32"]["bytes_per_cell"] == 8
    assert w["ac2d"]["shape"] == (512, 512) and w["ac2d"]["steps"] == 1000
    assert w["ac2d"]["params"] == [1.0, 1.0, 1.0, 1.0, 0.01] and (w["ac2d"]["lo"], w["ac2d"]["hi"]) == (-0.01, 0.01)
    assert w["gg2d"]["shape"] == (2048, 2048) and w["gg2d"]["fields"] == 32 and w["gg2d"]["params"] == [0.05]
    assert w["gg2d"]["bytes_per_cell"] == 2 * 32 * 4


def test_batched_members_are_seeded_and_uneven():
    pytest.importorskip("mmsp_amd")
    import bench

    sides, kap = bench.c4_members()
    assert len(sides) == 64 and set(sides) <= {128, 256, 512, 1024, 2048} and len(set(sides)) > 1
    assert np.all((kap >= 0.5) & (kap < 2.0))
    assert bench.c4_members()[0] == sides  # the same seed gives the same members


def test_served_trace_follows_the_configuration():
    pytest.importorskip("mmsp_amd")
    import bench

    tr = bench.serve_trace(400)
    assert tr == bench.serve_trace(400)
    t = np.array([r["t"] for r in tr])
    assert t[0] == 0.0 and np.all(np.diff(t) >= 0)
    rate = (len(t) - 1) / t[-1]
    assert 150 < rate < 260  # Poisson arrivals at 200 per second
    d2 = [r for r in tr if len(r["shape"]) == 2]
    d3 = [r for r in tr if len(r["shape"]) == 3]
    assert 0.4 < len(d2) / len(tr) < 0.6
    assert all(64 <= r["shape"][0] <= 1024 and r["shape"][0] == r["shape"][1] for r in d2)
    assert all(64 <= r["shape"][0] <= 256 and len(set(r["shape"])) == 1 for r in d3)
    steps = np.array([r["steps"] for r in tr])
    assert steps.min() >= 100 and steps.max() <= 10000
    # log-uniform: about half below the geometric middle, 1000
    assert 0.4 < np.mean(steps < 1000) < 0.6
