"""bench.py's contract pieces that run without a GPU: the one-GPU guard, the
committed traffic figure it reports, and the MLP FLOP count it divides by."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_traffic_figure_is_read_from_the_committed_profile():
    path = os.path.join(ROOT, "profiles", "fill_kernel_workload2_fp64_traffic.txt")
    traffic = bench.read_traffic(path)
    # the 5M-event fill moves about 1.07x its 5.76 GB of algorithmic bytes
    assert traffic is not None and 5.5e9 < traffic < 7.5e9
    assert bench.read_traffic(os.path.join(ROOT, "profiles", "no_such_file.txt")) is None


def test_mlp_flop_count_per_row():
    """Forward 2(n_in h + h^2 + h n_out); backward: the weight gradients of
    the three layers (the same again) plus the input gradients of the two
    upper layers (2(h^2 + h n_out)); the input layer's input gradient is not
    formed.  Counted here layer by layer, as the GEMMs of one training step."""
    n_in, h, n_out = 300, 512, 6000
    gemms = [(n_in, h), (h, h), (h, n_out)]           # (fan-in, fan-out) per layer
    fwd = sum(2 * a * b for a, b in gemms)
    dw = sum(2 * a * b for a, b in gemms)
    dx = sum(2 * a * b for a, b in gemms[1:])
    per_row = fwd + dw + dx
    assert per_row == 2.0 * (3 * (n_in * h + h * h + h * n_out) - n_in * h)
    # 5.3 GFLOP per 256-row step, as the profile summary states
    assert abs(per_row * 256 / 1e9 - 5.3) < 0.05
