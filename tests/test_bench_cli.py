"""bench.py's command line on the host: more than one GPU is refused with a
message and exit status 2 before any device is touched, and the all-cores
CPU baseline uses the cores the process may run on, capped by
OMP_NUM_THREADS."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    res = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=60)
    assert res.returncode == 2
    assert "one GPU" in res.stderr and res.stdout == ""


def test_host_threads(monkeypatch):
    assert bench.host_threads(5) == 5
    monkeypatch.setenv("OMP_NUM_THREADS", "1")
    assert bench.host_threads() == 1
    monkeypatch.delenv("OMP_NUM_THREADS")
    n = len(os.sched_getaffinity(0))
    assert bench.host_threads() == n
