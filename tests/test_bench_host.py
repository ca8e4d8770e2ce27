"""Host-side checks of bench.py's measurement arithmetic (no GPU needed):
the algorithmic-bytes convention on a hand-computed case, the HBM traffic
figure against the committed counter summary it is read from, the BP
iteration percentiles, and the one-GPU contract of --gpus."""
import os
import re
import subprocess
import sys

import numpy as np

from conftest import ROOT

sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_algorithmic_bytes_hand_computed():
    # graph a: 10 x 6-wide and 20 x 4-wide elements, 100 edges; the rotamer
    # graph 30 x 6 both sides, 50 edges; BP classes 3x3: 5, 3x6: 7, 6x6: 11;
    # 4 three-state and 2 six-state nodes; 90 atoms; 1000 bytes of small nodes
    graphs = {"a": {"shape": (10, 6, 20, 4), "n_edge": 100},
              "rotamer": {"shape": (30, 6, 30, 6), "n_edge": 50}}
    per_step = (8 * (60 + 80) + 8 * 100) + (8 * (180 + 180) + 8 * 50) + 8 * 50 \
        + 2 * 4 * (9 * 5 + 18 * 7 + 36 * 11) + 60 * 90 + 1000
    per_pass = 4 * (5 * (9 + 18) + 7 * (18 + 27) + 11 * (36 + 36)) + 8 * (3 * 4 + 6 * 2)
    total, s, p = bench.algorithmic_bytes(graphs, [0, 0, 0, 5, 7, 11], (0, 4, 2), 90, 1000, 3, 40)
    assert (s, p) == (per_step, per_pass) == (16536, 5160)
    assert total == 3 * per_step + 40 * per_pass


def test_traffic_matches_the_committed_counter_summary():
    """traffic_bytes_per_step in the profile equals its own counter readings
    (FETCH and WRITE KiB of the timed dispatch, scaled by the copy-kernel
    calibration, over the 300 steps of the launch)."""
    text = open(bench.TRAFFIC_FILE).read()
    num = r"([0-9.]+)"
    f_kib = float(re.search(r"FETCH_SIZE = " + num + r" KiB +-> [0-9.e+]+ bytes", text).group(1))
    w_kib = float(re.search(r"WRITE_SIZE = " + num + r" KiB +-> [0-9.e+]+ bytes", text).group(1))
    f_scale = float(re.search(r"FETCH_SIZE = [0-9.]+ KiB +-> scale " + num, text).group(1))
    w_scale = float(re.search(r"WRITE_SIZE = [0-9.]+ KiB +-> scale " + num, text).group(1))
    steps = int(re.search(r"--steps (\d+)", text).group(1))
    want = (f_kib * f_scale + w_kib * w_scale) * 1024 / steps
    got = bench.traffic_per_step()
    assert got is not None and abs(got - want) <= 2e-4 * want, (got, want)


def test_percentile_from_histogram():
    h = np.zeros(64, np.int64)
    h[[10, 12, 30]] = [50, 49, 1]
    assert bench.percentile_from_hist(h, 0.5) == 10
    assert bench.percentile_from_hist(h, 0.99) == 12
    assert bench.percentile_from_hist(h, 1.0) == 30
    assert bench.percentile_from_hist(np.zeros(8, np.int64), 0.5) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "one GPU" in (r.stdout + r.stderr)
