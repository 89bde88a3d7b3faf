"""bench.py's contract on the host: one GPU only (a request for more exits
with status 2 and says so in one JSON line), and the op census and the
committed traffic summary it reports from."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=ROOT)
    assert p.returncode == 2
    line = json.loads(p.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_traffic_summary_is_read():
    sys.path.insert(0, ROOT)
    import bench
    t = bench.measured_traffic("cdm_forward")
    assert t is None or t > 0.3e9    # at least the residual matrix (16,384 x 5,120 FP32) is written
    assert bench.CDM_FLOPS_PER_PAIR == 12 * 11 + 12 * 194 + 10
