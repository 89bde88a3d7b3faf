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
    assert bench.CDM_FLOPS_PER_PAIR == 12 * 10 + 12 * 177 + 10


def test_autocorrelation_times_on_ar1_chains():
    """bench.autocorrelation_times on AR(1) chains x_t = phi x_{t-1} + e_t,
    whose integrated autocorrelation time is (1 + phi) / (1 - phi): 1 for
    independent draws, 19 at phi = 0.9, also with a drift common to all
    chains added (the per-step centring removes it).  A record too short for
    the sequence to end is flagged as a lower bound."""
    import numpy as np
    import torch
    import bench
    rng = np.random.default_rng(5)
    L, C = 2000, 512
    phis = np.array([0.0, 0.5, 0.9])
    x = np.zeros((L, C, len(phis)))
    x[0] = rng.standard_normal((C, len(phis))) / np.sqrt(1 - phis ** 2)
    for t in range(1, L):
        x[t] = phis * x[t - 1] + rng.standard_normal((C, len(phis)))
    tau, cut, _ = bench.autocorrelation_times(torch.from_numpy(x).float())
    want = (1 + phis) / (1 - phis)
    assert np.all(np.abs(tau / want - 1) < 0.1), (tau, want)
    assert not cut.any()
    drift = np.linspace(0.0, 3.0, L)[:, None, None]
    tau_d, cut_d, _ = bench.autocorrelation_times(torch.from_numpy(x + drift).float())
    assert np.all(np.abs(tau_d / want - 1) < 0.1) and not cut_d.any()
    tau_short, cut_short, _ = bench.autocorrelation_times(torch.from_numpy(x[:40]).float(), max_lag=5)
    assert cut_short[2] and tau_short[2] < want[2]
