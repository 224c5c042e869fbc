#!/usr/bin/env python3
"""Headline benchmark: marching-cubes extraction of the 512^3 short2
KinectFusion-scale TSDF volume on one MI355X.

Workload (built on the device before timing): a 512^3 short2 volume over a
3 m cube fused from 300 synthetic 640x480 uint16 depth frames (fx=fy=525,
cx=319.5, cy=239.5; 40 seeded spheres/boxes; sigma 1.5 mm noise, 2 % dropout;
trunc 30 mm, max weight 128).  A step is one fused extraction pass
(classify + scan + triangle emission, vertex_voxel output) over the whole
volume; two copies of the volume alternate between steps so the 512 MiB input
cannot be served from the 256 MB MALL.

Prints one JSON line: value = Gvoxel/s (voxels of the volume per second of
step time), roofline of the extraction kernel (algorithmic bytes
4*X*Y*Z + 28*V per launch over its HIP-event duration on the launch stream,
against 8 TB/s), and the CPU baseline (the oracle's C restatement of the
same extraction, OpenMP over z planes, timed on the host cores).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--frames", type=int, default=300)
    ap.add_argument("--dim", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "extract_kinfu512_traffic.json"),
                    help="JSON with measured HBM bytes per launch (from rocprofv3 --pmc)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding of the path is described in DESIGN.md)", file=sys.stderr)
        return 2
    import torch
    from marching_cube_amd import fusion, mesh

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    cfg = dict(fusion.KINFU)
    dim = args.dim

    # ---- build the fused volume (integration timed separately)
    t0 = time.perf_counter()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    from marching_cube_amd import synth
    h = cfg["extent"] / dim
    intr = [cfg["fx"], cfg["fy"], cfg["cx"], cfg["cy"]]
    objs = synth.kinfu_scene(cfg["nobj"], cfg["extent"], cfg["seed"])
    poses = synth.orbit_poses(args.frames, cfg["extent"], seed=cfg["seed"])
    depth = fusion.synth_depth(intr, poses, objs, cfg["W"], cfg["H"], sigma=cfg["sigma"],
                               dropout=cfg["dropout"], seed=cfg["seed"], device=dev)
    vox = fusion.voxel_new((dim, dim, dim), cfg["trunc"], h, device=dev)
    torch.cuda.synchronize()
    e0.record()
    fusion.integrate(vox, intr, poses, depth, trunc=cfg["trunc"], max_w=cfg["max_w"])
    e1.record()
    torch.cuda.synchronize()
    integrate_ms = e0.elapsed_time(e1)
    del depth
    vols = [vox.data, vox.data.clone()]
    print(f"[bench] volume {dim}^3 fused from {args.frames} frames in {integrate_ms:.1f} ms "
          f"(setup {time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)

    # ---- size the output and check the two copies agree
    ws = mesh.Workspace(dev)
    nvox = dim ** 3
    count = mesh.extract_into(vols[0], h, cap=0, workspace=ws)
    cap = count
    vv = torch.empty((cap, mesh.VV_WORDS), dtype=torch.int32, device=dev)
    d_count = torch.zeros(1, dtype=torch.int64, device=dev)
    assert mesh.extract_into(vols[1], h, vv=vv, cap=cap, workspace=ws) == count

    def step(i):
        mesh.extract_into(vols[i & 1], h, vv=vv, cap=cap, d_count=d_count, workspace=ws,
                          sync=False)

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        evs[i][0].record()
        step(i)
        evs[i][1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    assert int(d_count.item()) == count
    kernel_ms = sorted(a.elapsed_time(b) for a, b in evs)
    avg_kernel_ms = sum(kernel_ms) / len(kernel_ms)
    med_kernel_ms = kernel_ms[len(kernel_ms) // 2]
    ms_per_step = wall * 1e3 / args.steps

    alg_bytes = 4 * nvox + 28 * count
    achieved = alg_bytes / (avg_kernel_ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(args.traffic):
        try:
            tj = json.load(open(args.traffic))
            if tj.get("dim") == dim and tj.get("frames") == args.frames:
                traffic = tj.get("hbm_bytes_per_launch")
        except (ValueError, OSError):
            traffic = None

    # ---- full reference API variant (vertex_voxel + int3 map + ver_voxel)
    v2 = torch.empty((cap, 3), dtype=torch.int32, device=dev)
    verv = torch.empty((dim, dim, dim), dtype=torch.int16, device=dev)
    mesh.extract_into(vols[0], h, vv=vv, v2vox=v2, ver_voxel=verv, cap=cap, workspace=ws)
    torch.cuda.synchronize()
    tf = time.perf_counter()
    nfull = max(3, args.steps // 4)
    for i in range(nfull):
        mesh.extract_into(vols[i & 1], h, vv=vv, v2vox=v2, ver_voxel=verv, cap=cap,
                          d_count=d_count, workspace=ws, sync=False)
    torch.cuda.synchronize()
    full_ms = (time.perf_counter() - tf) * 1e3 / nfull
    del v2, verv

    # ---- CPU baseline: the oracle's extraction on the host cores
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        import numpy as np
        host = vols[0].cpu().numpy().view(np.uint32)
        nt = max(1, min(args.cpu_threads, os.cpu_count() or 1))
        O.extract(host, h, want=("vv",), nthreads=nt)  # warm-up
        times = []
        n_cpu = None
        for _ in range(3):
            tc = time.perf_counter()
            r = O.extract(host, h, want=("vv",), nthreads=nt)
            times.append(time.perf_counter() - tc)
            n_cpu = r["n"]
        tmed = sorted(times)[1]
        assert n_cpu == count, (n_cpu, count)
        cpu = {"value": round(nvox / tmed / 1e9, 4), "unit": "Gvoxel/s", "cores": nt,
               "kind": "port",
               "sample": f"full {dim}^3 volume (same input), vertex_voxel output, median of 3 "
                         f"after 1 warm-up, {tmed:.3f} s per pass, {nt} OpenMP threads"}

    out = {
        "metric": "Gvoxel/s MC extract at 512³ (% HBM peak); fwd+bwd vs fwd-only time",
        "value": round(nvox / (ms_per_step * 1e-3) / 1e9, 3),
        "unit": "Gvoxel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16",
        "data": "synthetic",
        "config": {"workload": f"kinfu{dim}_extract", "volume": [dim, dim, dim],
                   "voxel": "short2", "frames": args.frames, "frame_size": [cfg["W"], cfg["H"]],
                   "extent_m": cfg["extent"], "trunc_m": cfg["trunc"], "max_weight": cfg["max_w"],
                   "outputs": "vertex_voxel"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic},
        "cpu_baseline": cpu,
        "vertices": count,
        "triangles": count // 3,
        "mtri_per_s": round(count / 3 / (ms_per_step * 1e-3) / 1e6, 1),
        "kernel_ms_avg": round(avg_kernel_ms, 4),
        "kernel_ms_median": round(med_kernel_ms, 4),
        "algorithmic_bytes": alg_bytes,
        "full_api_ms": round(full_ms, 4),
        "integrate_ms": round(integrate_ms, 2),
        "integrate_gvoxel_frames_per_s": round(nvox * args.frames / (integrate_ms * 1e-3) / 1e9, 2),
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
