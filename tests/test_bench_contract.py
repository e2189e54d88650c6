"""bench.py's contract, checked without a GPU: the metric is the one the
measurement plan names, more than one GPU is refused with a message (this
build measures one), and the roofline's traffic figure comes from the
committed rocprof summary of the step's persistent launch."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_metric_is_the_planned_one():
    with open(os.path.join(ROOT, "BASELINE.json")) as f:
        assert bench.METRIC == json.load(f)["metric"]
    assert bench.HBM_PEAK_GBS == 8000.0


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_traffic_comes_from_the_committed_profile():
    t = bench.load_traffic("batch_train_step_traffic")
    assert t is not None and t > 0
    with open(os.path.join(ROOT, "profiles", "batch_train_step_traffic.txt")) as f:
        text = f.read()
    assert f"traffic_bytes_per_launch: {int(t)}" in text
    # the same summary states the algorithmic bytes the bench divides by
    assert "227.49 MB algorithmic" in text
    assert bench.load_traffic("no_such_profile") is None
