"""bench.py's contract, checked without a GPU: the roofline's peaks, more
than one GPU refused with a message (this build measures one), and the
roofline's traffic figure read from the committed rocprof summary of the
step's persistent launch."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_roofline_constants():
    # MI355X's published peaks: 8 TB/s of HBM, 2.5 PFLOP/s of dense FP16 MFMA
    assert bench.HBM_PEAK_GBS == 8000.0
    assert bench.MFMA_F16_PEAK_TFLOPS == 2500.0


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
