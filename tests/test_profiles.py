"""The committed HBM traffic profile is the one bench.py's roofline reads: it
names bench.py's default workload (512^3 volume fused from 300 frames), so
the bench line carries the measured traffic, and its total is the sum of its
extraction kernels' scaled FETCH_SIZE and WRITE_SIZE bytes."""
import json
import os

import bench
from marching_cube_amd import rocprof_summary

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TRAFFIC = os.path.join(ROOT, "profiles", "extract_kinfu512_traffic.json")


def test_traffic_profile_matches_bench_workload(monkeypatch):
    tj = json.load(open(TRAFFIC))
    monkeypatch.setattr("sys.argv", ["bench.py"])
    args = bench.parse()
    assert tj["dim"] == args.dim and tj["frames"] == args.frames
    assert os.path.abspath(args.traffic) == TRAFFIC
    ks = tj["kernels"]
    assert ks and all(any(k.startswith(e) for e in rocprof_summary.EXTRACT_KERNELS) for k in ks)
    total = sum(v["fetch_bytes"] + v["write_bytes"] for v in ks.values())
    assert tj["hbm_bytes_per_launch"] == total
    assert tj["fetch_scale"] > 0 and tj["write_scale"] > 0
