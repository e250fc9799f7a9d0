"""bench.py's command-line contract that does not need a GPU."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_traffic_figure_is_read_from_the_committed_profile():
    sys.path.insert(0, ROOT)
    import bench
    for name in ("amg_prolong_smooth_fine_channel200k_pmc.txt", "bsr_spmv_dot_channel200k_pmc.txt"):
        t = bench.read_traffic(name)
        assert t is not None and 1.0e8 < t < 5.0e8
