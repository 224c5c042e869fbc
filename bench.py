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
4*X*Y*Z + 28*V per launch over the average launch duration: the span of one
HIP event pair on the launch stream around the K timed calls, over K; against
8 TB/s), and the CPU baseline (the oracle's C restatement of the
same extraction, OpenMP over z planes, timed on the host cores).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

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
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the all-cores CPU baseline (0 = every core this process "
                         "may use: its CPU affinity, capped by OMP_NUM_THREADS when set)")
    ap.add_argument("--diff-dim", type=int, default=256,
                    help="side of the fp32 bumps field of the forward/backward leg (0 = skip)")
    ap.add_argument("--batch-count", type=int, default=3000,
                    help="volumes of the batched leg (0 = skip)")
    ap.add_argument("--graph-count", type=int, default=256,
                    help="volumes of the graph-captured differentiable fusion leg (0 = skip)")
    ap.add_argument("--no-mesh-io", dest="mesh_io", action="store_false",
                    help="skip the indexed-mesh and .vol codec leg")
    ap.add_argument("--raycast-views", type=int, default=8,
                    help="camera poses of the raycast leg on the fused volume (0 = skip)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the launch-shape autotuner (default shapes)")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="after the legs, copy a 512 MiB buffer this many times (a plain copy "
                         "kernel of known size: the scale for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "extract_kinfu512_traffic.json"),
                    help="JSON with measured HBM bytes per launch (from rocprofv3 --pmc)")
    return ap.parse_args()


def nvox_int(dim):
    return dim * dim * dim


def host_threads(requested=0):
    """Threads for the all-cores CPU baseline: `requested`, or every core this
    process may run on (its affinity set), capped by OMP_NUM_THREADS."""
    if requested > 0:
        return requested
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_extract_times(vols, hs, nthreads, reps):
    """Seconds per pass of the oracle's extraction (vertex_voxel output) over
    every volume of `vols`: median of `reps` passes after one warm-up pass."""
    from oracle import oracle as O
    for v, h in zip(vols, hs):
        O.extract(v, float(h), want=("vv",), nthreads=nthreads)
    times, n = [], 0
    for _ in range(reps):
        t0 = time.perf_counter()
        n = 0
        for v, h in zip(vols, hs):
            n += O.extract(v, float(h), want=("vv",), nthreads=nthreads)["n"]
        times.append(time.perf_counter() - t0)
    return sorted(times)[len(times) // 2], n


def cpu_baseline_entry(vols, hs, threads_all, reps_single, reps_all, sample):
    """CPU baseline of an extraction workload: single thread and all cores
    (OpenMP over each volume's planes), Gvoxel/s and Mtri/s of the sample."""
    nvox = int(sum(int(np.prod(v.shape[:3])) for v in vols))
    t1, n1 = cpu_extract_times(vols, hs, 1, reps_single)
    ta, na = cpu_extract_times(vols, hs, threads_all, reps_all)
    assert n1 == na
    return {"value": round(nvox / ta / 1e9, 4), "unit": "Gvoxel/s", "cores": threads_all,
            "kind": "port", "mtri_per_s": round(na / 3 / ta / 1e6, 2),
            "seconds_per_pass": round(ta, 4),
            "single_thread": {"value": round(nvox / t1 / 1e9, 4), "unit": "Gvoxel/s",
                              "cores": 1, "mtri_per_s": round(n1 / 3 / t1 / 1e6, 2),
                              "seconds_per_pass": round(t1, 4)},
            "nproc": os.cpu_count(), "sample": sample}


def _event_ms(torch, fn, steps, warmup, spread=False):
    """Average ms of fn() over `steps` calls issued back to back: the span of
    one HIP event pair on the current stream around them, over `steps` (an
    event pair around every call would add idle time of its own).  With
    spread, also the median of `steps` further calls, each between its own
    event pair; else None."""
    for i in range(warmup):
        fn(i)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for i in range(steps):
        fn(i)
    e1.record()
    torch.cuda.synchronize()
    avg = e0.elapsed_time(e1) / steps
    if not spread:
        return avg, None
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    for i in range(steps):
        evs[i][0].record()
        fn(i)
        evs[i][1].record()
    torch.cuda.synchronize()
    ms = sorted(a.elapsed_time(b) for a, b in evs)
    return avg, ms[len(ms) // 2]


def bench_diff(torch, dev, dim, steps, warmup):
    """Forward vs forward+backward on the 256^3 fp32 sum-of-64-bumps field:
    the plain extraction (float3 positions, mcx_extract) and the
    differentiable forward (mcx_extract_diff, which also keeps the per-cube
    records the backward reads), both called directly through the C ABI, and
    the backward (mcx_extract_backward) for the gradient of the sum of
    squared distances of the vertices to a sphere of radius 0.35.  Five
    copies of the 64 MiB field rotate between steps (320 MiB: more than the
    256 MB MALL holds), each step's backward on its forward's copy.  Kernel
    time between HIP events; the loss gradient is computed once outside the
    timed region."""
    from marching_cube_amd import _capi, diff, mesh, synth
    from marching_cube_amd.mesh import _ptr, _stream_handle
    lib = _capi.load()
    f, h = synth.bumps_volume(dim)
    copies = [torch.from_numpy(f).to(dev) for _ in range(5)]
    vol = copies[0]
    Z, Y, X = vol.shape
    ex = diff.DiffExtractor(vol.shape, dev)
    pos = ex.forward(vol, h)
    count = int(pos.shape[0])
    c = torch.tensor([0.5, 0.5, 0.5], dtype=torch.float32, device=dev)
    d = pos - c
    nrm = d.norm(dim=1, keepdim=True)
    grad_pos = (2 * (nrm - 0.35) / nrm * d).contiguous()
    gvol = torch.empty_like(vol)
    ws = mesh.Workspace(dev, lib.mcx_extract_workspace_size(X, Y, Z))
    pbuf = torch.empty((count, 3), dtype=torch.float32, device=dev)
    dc = torch.zeros(1, dtype=torch.int64, device=dev)
    stream = _stream_handle(dev)
    # same positions from both forwards ("the same vertex accuracy")
    hc = ctypes.c_int64(0)
    _capi.check(lib.mcx_extract(_ptr(vol), _capi.ELEM_F32, X, Y, Z, float(h), None, _ptr(pbuf),
                                None, None, count, None, ctypes.byref(hc), ws.ptr, ws.nbytes,
                                stream), "mcx_extract")
    assert hc.value == count
    assert pbuf.cpu().numpy().tobytes() == pos.cpu().numpy().tobytes()

    def plain(i):
        _capi.check(lib.mcx_extract(_ptr(copies[i % 5]), _capi.ELEM_F32, X, Y, Z, float(h), None,
                                    _ptr(pbuf), None, None, count, _ptr(dc), None, ws.ptr,
                                    ws.nbytes, stream), "mcx_extract")

    def fwd_diff(i):
        _capi.check(lib.mcx_extract_diff(_ptr(copies[i % 5]), X, Y, Z, float(h), _ptr(ex.pos),
                                         ex.pos.shape[0], _ptr(dc), None, _ptr(ex.ws),
                                         ex.ws.numel(), stream), "mcx_extract_diff")

    def bwd(i):
        _capi.check(lib.mcx_extract_backward(_ptr(copies[i % 5]), X, Y, Z, float(h),
                                             _ptr(grad_pos), _ptr(gvol), _ptr(ex.ws),
                                             ex.ws.numel(), _ptr(ex.scratch), ex.scratch.numel(),
                                             stream), "mcx_extract_backward")

    def fwd_bwd(i):
        fwd_diff(i)
        bwd(i)

    plain_ms, _ = _event_ms(torch, plain, steps, warmup)
    fdiff_ms, _ = _event_ms(torch, fwd_diff, steps, warmup)
    fb_ms, _ = _event_ms(torch, fwd_bwd, steps, warmup)
    ws.check()
    _capi.check(lib.mcx_workspace_status(_ptr(ex.ws), ex.ws.numel(), stream), "diff workspace")
    ex.forward(vol, h)
    ref = ex.backward(vol, grad_pos)
    fwd_bwd(0)
    torch.cuda.synchronize()
    assert gvol.cpu().numpy().tobytes() == ref.cpu().numpy().tobytes()  # deterministic
    # end-to-end autograd step (forward, loss, backward) with a reused extractor
    vg = vol.clone().requires_grad_(True)

    def autograd_step(i):
        p = diff.extract_positions(vg, h, extractor=ex)
        loss = ((torch.linalg.vector_norm(p - c, dim=1) - 0.35) ** 2).sum()
        vg.grad = None
        loss.backward()
    ag_ms, _ = _event_ms(torch, autograd_step, max(3, steps // 2), 2)
    nvox = X * Y * Z

    def roof(alg, ms):
        gbs = alg / (ms * 1e-3) / 1e9
        return {"bound": "hbm", "algorithmic_bytes": alg, "achieved": round(gbs, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)}
    bwd_ms = fb_ms - fdiff_ms
    return {"workload": f"bumps{dim}_fp32_fwd_bwd", "vertices": count,
            "fwd_ms": round(plain_ms, 4), "fwd_diff_ms": round(fdiff_ms, 4),
            "bwd_ms": round(bwd_ms, 4), "fwd_bwd_ms": round(fb_ms, 4),
            "fwd_bwd_over_fwd": round(fb_ms / plain_ms, 3),
            "fwd_bwd_over_fwd_diff": round(fb_ms / fdiff_ms, 3),
            "volume_copies": 5, "calls": "C ABI (mcx_extract, mcx_extract_diff, "
                                          "mcx_extract_backward), HIP events",
            "fwd_gvoxel_per_s": round(nvox / (plain_ms * 1e-3) / 1e9, 1),
            # algorithmic bytes: forward 4 X Y Z + 12 V (float3 positions);
            # backward 12 V (dL/dv) + 4 X Y Z (dL/dtsdf) + 4 X Y Z (volume re-read)
            "roofline_fwd": roof(4 * nvox + 12 * count, plain_ms),
            "roofline_bwd": roof(12 * count + 8 * nvox, bwd_ms),
            "autograd_step_ms": round(ag_ms, 4)}


def bench_batch(torch, dev, count, steps, warmup, threads_all, no_cpu, tune):
    """3,000 uneven short2 shape volumes (sides log-uniform in [16,128] per
    axis; sphere/box/torus at random poses): one batched launch sequence
    (planes per column and emit workgroups picked by the autotuner) against
    one extraction call per volume.  Both outputs are checked byte for byte
    after a call of each into buffers filled with -1.  Roofline: algorithmic
    bytes 4 * voxels + 28 * vertices per batched call over its HIP-event
    time.  CPU baseline: the oracle on the first 300 volumes."""
    from marching_cube_amd import _capi, autotune, batch, mesh, synth
    lib = _capi.load()
    vols, hs = synth.batch_shapes(count, device=dev)
    torch.cuda.synchronize()
    saved = autotune.get()
    tuned = None
    if tune:
        plan0 = batch.BatchPlan(vols, hs)
        n0, vb0, out0 = plan0.extract(want=("vv",))
        vv0 = out0["vv"]
        del plan0
        t = autotune.tune_batch(lambda: batch.BatchPlan(vols, hs),
                                lambda pl: pl.extract_into(vv=vv0, cap=n0, vbase=vb0, sync=False))
        tuned = {"batch_slab": t["batch_slab"], "emit_groups_per_cu": t["emit_groups_per_cu"],
                 "candidates": len(t["table"])}
        del vv0, out0
    try:
        plan = batch.BatchPlan(vols, hs)
        n, vbase, out = plan.extract(want=("vv",))
        vv = out["vv"]
        vb = vbase.cpu().numpy()
        dc = torch.zeros(1, dtype=torch.int64, device=dev)   # batched calls' total
        dcp = torch.zeros(1, dtype=torch.int64, device=dev)  # per-volume calls' counts

        def batched(i):
            plan.extract_into(vv=vv, cap=n, vbase=vbase, d_count=dc, sync=False)

        vv2 = torch.empty_like(vv)
        ws = mesh.Workspace(dev)
        lib_sizes = [mesh.volume_kind(v)[1:] for v in vols]
        ws.ensure(max(lib.mcx_extract_workspace_size(X, Y, Z) for X, Y, Z in lib_sizes))
        segs = [vv2[vb[b]:vb[b + 1]] for b in range(count)]

        def per_volume(i):
            for b in range(count):
                mesh.extract_into(vols[b], float(hs[b]), vv=segs[b], cap=int(vb[b + 1] - vb[b]),
                                  d_count=dcp, workspace=ws, sync=False)

        b_ms, _ = _event_ms(torch, batched, steps, warmup)
        p_ms, _ = _event_ms(torch, per_volume, max(2, steps // 4), 1)
        # one call of each into outputs filled with -1: every vertex written
        vv.fill_(-1)
        vv2.fill_(-1)
        batched(0)
        per_volume(0)
        torch.cuda.synchronize()
        plan.check()  # device-side status of every asynchronous call above
        ws.check()
        assert int(dc.item()) == int(vb[count])
        assert int(dcp.item()) == int(vb[count] - vb[count - 1])  # the last volume's count
        assert vv.cpu().numpy().tobytes() == vv2.cpu().numpy().tobytes()
    finally:
        autotune.set(**saved)
    nvox = int(sum(X * Y * Z for X, Y, Z in lib_sizes))
    alg = 4 * nvox + 28 * int(n)
    achieved = alg / (b_ms * 1e-3) / 1e9
    res = {"workload": f"batch{count}_shapes_16_128", "volumes": count, "voxels": nvox,
           "vertices": int(n), "batched_ms": round(b_ms, 4), "per_volume_ms": round(p_ms, 4),
           "speedup": round(p_ms / b_ms, 2),
           "batched_gvoxel_per_s": round(nvox / (b_ms * 1e-3) / 1e9, 1),
           "batched_mtri_per_s": round(n / 3 / (b_ms * 1e-3) / 1e6, 1),
           "algorithmic_bytes": alg,
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4)},
           "tuning": tuned, "cpu_baseline": None}
    if not no_cpu:
        m = min(300, count)
        host = [v.cpu().numpy().view(np.uint32) for v in vols[:m]]
        res["cpu_baseline"] = cpu_baseline_entry(
            host, [float(x) for x in hs[:m]], threads_all, 3, 3,
            f"the first {m} of the {count} volumes (same inputs); single thread and all cores "
            "(OpenMP over each volume's planes): median of 3 passes after 1 warm-up pass")
    return res


def bench_c1(torch, dev, steps, warmup, threads_all, no_cpu):
    """The reference's CPU-path case: 128^3 short2 sphere (r 0.3 m in a 1 m
    cube, trunc 3 voxels), vertex_voxel output.  GPU: extraction calls back to
    back between one HIP event pair (the 8 MiB volume stays in cache, so this
    is not an HBM figure); CPU: the oracle, single thread and all cores."""
    from marching_cube_amd import mesh, synth
    vol_np, h = synth.sphere_volume(128)
    vol = torch.from_numpy(vol_np.view(np.int32)).to(dev)
    ws = mesh.Workspace(dev)
    n = mesh.extract_into(vol, h, cap=0, workspace=ws)
    vv = torch.empty((n, mesh.VV_WORDS), dtype=torch.int32, device=dev)
    dc = torch.zeros(1, dtype=torch.int64, device=dev)

    # straight through the C ABI, arguments converted once (back-to-back calls
    # of this short extraction must not wait on the host)
    from marching_cube_amd import _capi
    lib = _capi.load()
    c_args = [ctypes.c_void_p(vol.data_ptr()), ctypes.c_int(_capi.ELEM_SHORT2), ctypes.c_int(128),
              ctypes.c_int(128), ctypes.c_int(128), ctypes.c_float(h),
              ctypes.c_void_p(vv.data_ptr()), None, None, None, ctypes.c_int64(n),
              ctypes.c_void_p(dc.data_ptr()), None, ws.ptr, ctypes.c_size_t(ws.nbytes),
              mesh._stream_handle(dev)]

    def step(i):
        st = lib.mcx_extract(*c_args)
        if st:
            _capi.check(st, "mcx_extract")

    avg, med = _event_ms(torch, step, max(steps, 20), warmup, spread=True)
    ws.check()
    assert int(dc.item()) == n
    nvox = 128 ** 3
    out = {"workload": "sphere128_extract", "vertices": n, "ms": round(avg, 4),
           "ms_median": round(med, 4), "gvoxel_per_s": round(nvox / (avg * 1e-3) / 1e9, 2),
           "mtri_per_s": round(n / 3 / (avg * 1e-3) / 1e6, 1), "cpu_baseline": None}
    if not no_cpu:
        out["cpu_baseline"] = cpu_baseline_entry(
            [vol_np], [h], threads_all, 5, 9,
            "the whole 128^3 volume; single thread: median of 5 passes, all cores: median of "
            "9, each after 1 warm-up pass")
    return out


def bench_raycast(torch, dev, vol, h, cfg, views, steps, warmup, cpu_threads, no_cpu):
    """Depth + normal maps (640x480, the frames' camera) ray-marched through
    the fused volume from `views` poses of the orbit, one launch per view,
    timed with HIP events; the CPU oracle renders one view for the baseline."""
    import numpy as np
    from marching_cube_amd import fusion, synth
    W, H = cfg["W"], cfg["H"]
    intr = [cfg["fx"], cfg["fy"], cfg["cx"], cfg["cy"]]
    poses = synth.orbit_poses(views, cfg["extent"], seed=cfg["seed"] + 1)
    depth = torch.empty((H, W), dtype=torch.float32, device=dev)
    normals = torch.empty((H, W, 3), dtype=torch.float32, device=dev)

    def cast(i):
        fusion.raycast(vol, intr, poses[i % views], W, H, h=h, depth=depth, normals=normals)

    avg, med = _event_ms(torch, cast, max(steps, views), warmup, spread=True)
    hits = 0
    for k in range(views):
        cast(k)
        hits += int((depth > 0).sum().item())
    out = {"workload": f"raycast{W}x{H}_on_fused_volume", "views": views,
           "ms_per_view": round(avg, 4), "ms_per_view_median": round(med, 4),
           "mrays_per_s": round(W * H / (avg * 1e-3) / 1e6, 1),
           "hit_fraction": round(hits / (views * W * H), 3), "cpu_baseline": None}
    if not no_cpu:
        from oracle import oracle as O
        host = vol.cpu().numpy().view(np.uint32)
        nt = host_threads(cpu_threads)
        O.raycast(host, h, intr, poses[0], W, H, origin=0.5, nthreads=nt)
        tc = time.perf_counter()
        O.raycast(host, h, intr, poses[0], W, H, origin=0.5, nthreads=nt)
        tcpu = time.perf_counter() - tc
        out["cpu_baseline"] = {"ms_per_view": round(tcpu * 1e3, 1), "kind": "port",
                               "cores": nt, "sample": "one 640x480 view (view 0), after 1 warm-up"}
    return out


def bench_mesh_io(torch, dev, vol, h, steps):
    """On the fused volume: indexed mesh (extraction with edge keys + weld),
    and the .vol codec (device volume -> device dump -> device volume)."""
    import ctypes
    import numpy as np
    from marching_cube_amd import _capi, mesh, volio
    lib = _capi.load()
    elem, X, Y, Z = mesh.volume_kind(vol)
    ws = mesh.Workspace(dev)
    n = mesh.extract_into(vol, h, cap=0, workspace=ws)
    pos = torch.empty((n, 3), dtype=torch.float32, device=dev)
    key = torch.empty(n, dtype=torch.int64, device=dev)
    upos = torch.empty((n, 3), dtype=torch.float32, device=dev)
    idx = torch.empty(n, dtype=torch.int32, device=dev)
    wws = torch.empty(lib.mcx_weld_workspace_size(n), dtype=torch.uint8, device=dev)
    st = mesh._stream_handle(dev)

    def extract_keys(i):
        _capi.check(lib.mcx_extract_ex(mesh._ptr(vol), elem, X, Y, Z, float(h), None,
                                       mesh._ptr(pos), None, None, mesh._ptr(key), n, None, None,
                                       ws.ptr, ws.nbytes, st), "mcx_extract_ex")

    def weld(i):
        _capi.check(lib.mcx_weld(mesh._ptr(key), mesh._ptr(pos), 3, n, mesh._ptr(upos),
                                 mesh._ptr(idx), None, None, mesh._ptr(wws), wws.numel(), st),
                    "mcx_weld")

    ek_ms, _ = _event_ms(torch, extract_keys, steps, 2)
    weld_ms, _ = _event_ms(torch, weld, steps, 2)
    nu = ctypes.c_int64(0)
    _capi.check(lib.mcx_weld(mesh._ptr(key), mesh._ptr(pos), 3, n, mesh._ptr(upos),
                             mesh._ptr(idx), None, ctypes.byref(nu), mesh._ptr(wws), wws.numel(),
                             st), "mcx_weld")
    # codec: whole volume in one call each way (the calls synchronise)
    ews = torch.empty(lib.mcx_rle_encode_workspace_size(X, Y, Z), dtype=torch.uint8, device=dev)
    dump = torch.empty(lib.mcx_rle_encode_bound(X, Y, Z), dtype=torch.uint8, device=dev)
    total = ctypes.c_int64(0)

    def encode():
        _capi.check(lib.mcx_rle_encode(mesh._ptr(vol), X, Y, Z, 0, Z, mesh._ptr(dump),
                                       dump.numel(), None, ctypes.byref(total), mesh._ptr(ews),
                                       ews.numel(), st), "mcx_rle_encode")

    encode()
    reps = max(3, steps // 4)
    t0 = time.perf_counter()
    for _ in range(reps):
        encode()
    enc_ms = (time.perf_counter() - t0) * 1e3 / reps
    del ews
    nb = total.value
    d = dump[:nb]
    offs = volio.slice_offsets(d.cpu().numpy().tobytes(), Z)
    out = torch.empty_like(vol)
    dws = torch.empty(lib.mcx_rle_decode_workspace_size(nb, X, Y, Z), dtype=torch.uint8,
                      device=dev)

    def decode():
        _capi.check(lib.mcx_rle_decode(mesh._ptr(d), nb, offs.ctypes.data, Z, X, Y,
                                       mesh._ptr(out), mesh._ptr(dws), dws.numel(), st),
                    "mcx_rle_decode")

    decode()
    t0 = time.perf_counter()
    for _ in range(reps):
        decode()
    dec_ms = (time.perf_counter() - t0) * 1e3 / reps
    # the round trip keeps every tsdf (weights of saturated voxels become 0)
    same_t = bool(((out & 0xFFFF) == (vol & 0xFFFF)).all().item())
    res = {"vertices": n, "unique_vertices": nu.value,
           "extract_with_keys_ms": round(ek_ms, 4), "weld_ms": round(weld_ms, 4),
           "vol_dump_bytes": nb, "vol_ratio": round(4 * X * Y * Z / nb, 2),
           "encode_ms": round(enc_ms, 3), "decode_ms": round(dec_ms, 3),
           "encode_gvoxel_per_s": round(X * Y * Z / (enc_ms * 1e-3) / 1e9, 2),
           "decode_gvoxel_per_s": round(X * Y * Z / (dec_ms * 1e-3) / 1e9, 2),
           "round_trip_tsdf_exact": same_t}
    del dump, dws, out
    return res


def bench_fusion_graph(torch, dev, nvol, steps):
    """256 volumes with sides uniform in [64, 192], each fused from 16
    synthetic 320x240 uint16 depth frames (3 random spheres/boxes in front of
    a far background box, 1.5 mm noise, 2 % dropout) converted to float
    metres; integrate -> extract -> Chamfer loss to 64 points on a sphere ->
    backward to every depth pixel, recorded once in a HIP graph.  Replays
    timed with HIP events, against the same pass launched eagerly."""
    import numpy as np
    from marching_cube_amd import fusion, fusion_diff, synth
    rng = np.random.default_rng(20261202)
    W, H, K = 320, 240, 16
    intr = [262.5, 262.5, 159.5, 119.5]
    shapes, hs, depth, poses, targets = [], [], [], [], []
    bg = np.array([[1, 0.5, 0.5, 0.5, 20.0, 20.0, 20.0, 0]], np.float32)
    for b in range(nvol):
        n = int(rng.integers(64, 193))
        h = 1.0 / n
        objs = np.concatenate([synth.kinfu_scene(3, extent=1.0, seed=b), bg])
        p = synth.orbit_poses(K, extent=1.0, radius=1.6, seed=b)
        d16 = fusion.synth_depth(intr, p, objs, W, H, sigma=0.0015, dropout=0.02, seed=b,
                                 device=dev)
        depth.append((d16.to(torch.float32) * 0.001).contiguous())
        poses.append(torch.from_numpy(p).to(dev))
        i = np.arange(64) + 0.5
        phi = np.arccos(1 - 2 * i / 64)
        th = np.pi * (1 + 5 ** 0.5) * i
        pts = 0.5 + 0.25 * np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi),
                                     np.cos(phi)], -1)
        targets.append(torch.from_numpy(pts.astype(np.float32)).to(dev))
        shapes.append((n, n, n))
        hs.append(h)
    g = fusion_diff.FusionGraph(shapes, hs, depth, poses, targets, intr, trunc=0.03)
    e_ms, _ = _event_ms(torch, lambda i: g.step(), 2, 1)
    g.capture()
    g_ms, _ = _event_ms(torch, lambda i: g.replay(), max(3, steps // 4), 1)
    g.check()
    loss_g = g.loss.clone()
    grad_g = g.grad_depth_all.clone()
    vol_ms, loss_v, grad_v = per_volume_fusion(torch, dev, g)
    # the same pass one volume at a time: the same losses, and gradients equal
    # up to the fixed-point scale (per volume there, per batch here)
    assert torch.allclose(loss_v, loss_g, rtol=1e-5, atol=0)
    assert (grad_v - grad_g).abs().max().item() <= 1e-6 * grad_g.abs().max().item()
    nvox = int(sum(a * b * c for a, b, c in shapes))
    return {"workload": f"fusion{nvol}_64_192_diff_graph", "volumes": nvol, "voxels": nvox,
            "frames_per_volume": K, "frame_size": [W, H],
            "vertices": int(g.count.sum().item()), "graph_ms": round(g_ms, 3),
            "eager_ms": round(e_ms, 3), "per_volume_ms": round(vol_ms, 3),
            "graph_vs_per_volume": round(vol_ms / g_ms, 2),
            "gvoxel_frames_per_s": round(nvox * K / (g_ms * 1e-3) / 1e9, 1)}


def per_volume_fusion(torch, dev, g):
    """The pass of FusionGraph `g` one volume at a time through the
    single-volume entry points (fp32 integration, differentiable extraction,
    the loss with one volume, MC backward, integration backward), with the
    same kernels: HIP-event time of one eager pass, its losses and dL/ddepth."""
    import ctypes
    from marching_cube_amd import _capi, fusion_diff
    from marching_cube_amd.mesh import _ptr, _stream_handle
    lib = _capi.load()
    s = _stream_handle(dev)
    tsdf = torch.empty_like(g.tsdf_all)
    w = torch.empty_like(g.w_all)
    gt = torch.empty_like(g.tsdf_all)
    grad = torch.empty_like(g.grad_depth_all)
    loss = torch.zeros(g.B, dtype=torch.float32, device=dev)
    big = max(range(g.B), key=lambda b: g.tsdf[b].numel())
    Zb, Yb, Xb = g.shapes[big]
    ws = torch.empty(max(lib.mcx_extract_workspace_size_ex(Xb, Yb, Zb, 1 << 23), 4096),
                     dtype=torch.uint8, device=dev)
    _capi.check(lib.mcx_workspace_init(_ptr(ws), ws.numel(), s), "workspace_init")
    # the same workspace leaves more record room to smaller volumes
    scr = torch.empty(max(16, max(lib.mcx_backward_scratch_size(X, Y, Z, ws.numel())
                                  for Z, Y, X in g.shapes)), dtype=torch.uint8, device=dev)
    cap = int(g.count.max().item() * 2) + 4096
    pos = torch.zeros((cap, 3), dtype=torch.float32, device=dev)
    gpos = torch.zeros_like(pos)
    vb1 = torch.zeros((g.B, 2), dtype=torch.int64, device=dev)  # per volume [0, count]
    ws_l = torch.empty(max(4096, lib.mcx_chamfer_workspace_size(cap)), dtype=torch.uint8,
                       device=dev)
    integ = [fusion_diff._Integrator(g.shapes[b], g.W, g.H, g.K, g.hs[b], g.trunc, g.max_w,
                                     list(g.intr), dev) for b in range(g.B)]

    def one(i):
        for b in range(g.B):
            Z, Y, X = g.shapes[b]
            n = Z * Y * X
            tb = tsdf[g.vox_begin[b]:g.vox_begin[b] + n].view(Z, Y, X)
            wb = w[g.vox_begin[b]:g.vox_begin[b] + n].view(Z, Y, X)
            gb = gt[g.vox_begin[b]:g.vox_begin[b] + n].view(Z, Y, X)
            _capi.check(lib.mcx_tsdf_init(_ptr(tb), _ptr(wb), n, g.init_value, s), "tsdf_init")
            integ[b].forward(tb, wb, g.poses[b], g.depth[b])
            _capi.check(lib.mcx_extract_diff(_ptr(tb), X, Y, Z, g.hs[b], _ptr(pos), cap,
                                             _ptr(vb1[b, 1:]), None, _ptr(ws), ws.numel(), s),
                        "mcx_extract_diff")
            _capi.check(lib.mcx_chamfer_batched(_ptr(pos), _ptr(vb1[b]), 1, _ptr(g.targets[b]),
                                                g.M, cap, _ptr(loss[b:]), _ptr(gpos),
                                                _ptr(ws_l), ws_l.numel(), s), "chamfer")
            _capi.check(lib.mcx_extract_backward(_ptr(tb), X, Y, Z, g.hs[b], _ptr(gpos),
                                                 _ptr(gb), _ptr(ws), ws.numel(), _ptr(scr),
                                                 scr.numel(), s), "mcx_extract_backward")
            integ[b].backward(gb, None, g.poses[b], g.depth[b], grad[b])

    ms, _ = _event_ms(torch, one, 1, 1)
    _capi.check(lib.mcx_workspace_status(_ptr(ws), ws.numel(), s), "per-volume workspace")
    return ms, loss, grad


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: the extraction runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by z-slabs and by volume, which is not built)",
              file=sys.stderr)
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
    # the reference's call pattern: one integration call per frame (K = 1)
    from marching_cube_amd import _capi
    from marching_cube_amd.mesh import _ptr, _stream_handle
    lib = _capi.load()
    vox1 = fusion.voxel_new((dim, dim, dim), cfg["trunc"], h, device=dev)
    dposes = torch.as_tensor(poses, dtype=torch.float32).reshape(-1, 12).to(dev).contiguous()
    iws = torch.empty(lib.mcx_integrate_workspace_size(cfg["W"], cfg["H"]), dtype=torch.uint8,
                      device=dev)
    intr_c = (ctypes.c_float * 4)(*intr)
    st = _stream_handle(dev)
    frame_px = cfg["W"] * cfg["H"]
    torch.cuda.synchronize()
    e0.record()
    for k in range(args.frames):
        _capi.check(lib.mcx_integrate(_ptr(vox1.data), dim, dim, dim, float(h),
                                      float(cfg["trunc"]), int(cfg["max_w"]), intr_c,
                                      ctypes.c_void_p(dposes.data_ptr() + 48 * k),
                                      ctypes.c_void_p(depth.data_ptr() + 2 * frame_px * k),
                                      0.001, cfg["W"], cfg["H"], 1, _ptr(iws), iws.numel(), st),
                    "mcx_integrate")
    e1.record()
    torch.cuda.synchronize()
    per_frame_ms = e0.elapsed_time(e1) / args.frames
    assert torch.equal(vox1.data, vox.data)  # same updates in the same order
    del vox1, depth
    frame_bytes = 8 * nvox_int(dim) + 2 * cfg["W"] * cfg["H"]
    per_frame = {"workload": f"kinfu{dim}_integrate_per_frame", "frames": args.frames,
                 "ms_per_frame": round(per_frame_ms, 4),
                 "algorithmic_bytes_per_frame": frame_bytes,
                 "roofline": {"bound": "hbm",
                              "achieved": round(frame_bytes / (per_frame_ms * 1e-3) / 1e9, 1),
                              "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(frame_bytes / (per_frame_ms * 1e-3) / 1e9
                                            / HBM_PEAK_GBS, 4)},
                 "note": "one mcx_integrate call per frame; a voxel is read and written only "
                         "where the frame updates it"}
    print(f"[bench] integrate per frame {per_frame}", file=sys.stderr, flush=True)
    vols = [vox.data, vox.data.clone()]
    print(f"[bench] volume {dim}^3 fused from {args.frames} frames in {integrate_ms:.1f} ms "
          f"(setup {time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)

    # ---- launch shapes: the autotuner times its candidates on this volume
    # (part of the set-up; outputs do not depend on them)
    tuning = None
    if not args.no_tune:
        from marching_cube_amd import autotune
        t = autotune.tune_extract(vols[0], h)
        tuning = {"columns_per_cu": t["columns_per_cu"],
                  "emit_groups_per_cu": t["emit_groups_per_cu"], "candidates": len(t["table"])}
        print(f"[bench] autotune {t['columns_per_cu']} columns/CU, "
              f"{t['emit_groups_per_cu']} emit groups/CU ({t['ms']:.3f} ms)", file=sys.stderr,
              flush=True)

    # ---- size the output and check the two copies agree
    ws = mesh.Workspace(dev)
    nvox = dim ** 3
    count = mesh.extract_into(vols[0], h, cap=0, workspace=ws)
    cap = count
    vv = torch.empty((cap, mesh.VV_WORDS), dtype=torch.int32, device=dev)
    d_count = torch.zeros(1, dtype=torch.int64, device=dev)
    assert mesh.extract_into(vols[1], h, vv=vv, cap=cap, workspace=ws) == count

    # the timed calls go straight through the C ABI (mcx_extract, the drop-in
    # boundary), their arguments converted once
    lib = _capi.load()
    ws.ensure(lib.mcx_extract_workspace_size(dim, dim, dim))
    c_args = [[ctypes.c_void_p(v.data_ptr()), ctypes.c_int(_capi.ELEM_SHORT2), ctypes.c_int(dim),
               ctypes.c_int(dim), ctypes.c_int(dim), ctypes.c_float(h),
               ctypes.c_void_p(vv.data_ptr()), None, None, None, ctypes.c_int64(cap),
               ctypes.c_void_p(d_count.data_ptr()), None, ws.ptr, ctypes.c_size_t(ws.nbytes),
               mesh._stream_handle(dev)]
              for v in vols]
    extract_c = lib.mcx_extract

    def step(i):
        st = extract_c(*c_args[i & 1])
        if st:
            _capi.check(st, "mcx_extract")

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    # The timed region: the K calls back to back, one HIP event pair on the
    # launch stream around all of them (an event pair around every call adds
    # ~7 us of idle time per call on the stream), so the average extraction
    # time is the span over K, gaps between the calls included.
    e_start = torch.cuda.Event(enable_timing=True)
    e_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    e_start.record()
    for i in range(args.steps):
        step(i)
    e_end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ws.check()  # device-side status of the asynchronous calls
    assert int(d_count.item()) == count
    avg_kernel_ms = e_start.elapsed_time(e_end) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    # Per-call spread (not the timed region): an event pair around each of K
    # further calls.
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    for i in range(args.steps):
        evs[i][0].record()
        step(i)
        evs[i][1].record()
    torch.cuda.synchronize()
    per_call = sorted(a.elapsed_time(b) for a, b in evs)
    per_call_ms = {"avg": round(sum(per_call) / len(per_call), 4),
                   "median": round(per_call[len(per_call) // 2], 4),
                   "min": round(per_call[0], 4), "max": round(per_call[-1], 4),
                   "note": "an event pair around each call (adds idle time of its own)"}

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

    diff_leg = bench_diff(torch, dev, args.diff_dim, args.steps, args.warmup) if args.diff_dim else None
    if diff_leg:
        print(f"[bench] diff {diff_leg}", file=sys.stderr, flush=True)
    batch_leg = (bench_batch(torch, dev, args.batch_count, args.steps, args.warmup,
                             host_threads(args.cpu_threads), args.no_cpu, not args.no_tune)
                 if args.batch_count else None)
    if batch_leg:
        print(f"[bench] batch {batch_leg}", file=sys.stderr, flush=True)
    graph_leg = bench_fusion_graph(torch, dev, args.graph_count, args.steps) if args.graph_count else None
    if graph_leg:
        print(f"[bench] fusion graph {graph_leg}", file=sys.stderr, flush=True)

    io_leg = bench_mesh_io(torch, dev, vols[0], h, args.steps) if args.mesh_io else None
    if io_leg:
        print(f"[bench] mesh io {io_leg}", file=sys.stderr, flush=True)
    ray_leg = (bench_raycast(torch, dev, vols[0], h, cfg, args.raycast_views, args.steps,
                             args.warmup, args.cpu_threads, args.no_cpu)
               if args.raycast_views else None)
    if ray_leg:
        print(f"[bench] raycast {ray_leg}", file=sys.stderr, flush=True)

    # ---- CPU baseline: the oracle's extraction on the host cores, single
    # thread and all cores (the reference's CPU path is only declared,
    # MyMarchingCube.h:4-10; the oracle restates its GPU semantics)
    cpu = None
    if not args.no_cpu:
        host = vols[0].cpu().numpy().view(np.uint32)
        nt = host_threads(args.cpu_threads)
        cpu = cpu_baseline_entry(
            [host], [h], nt, 3, 5,
            f"full {dim}^3 volume (same input), vertex_voxel output; single thread: median of 3 "
            f"passes, all cores ({nt} OpenMP threads over z planes): median of 5, each after "
            f"1 warm-up pass")
        print(f"[bench] cpu baseline {cpu}", file=sys.stderr, flush=True)
    c1_leg = bench_c1(torch, dev, args.steps, args.warmup, host_threads(args.cpu_threads),
                      args.no_cpu)
    print(f"[bench] c1 {c1_leg}", file=sys.stderr, flush=True)

    if args.calib_copy:
        src = torch.empty(128 << 20, dtype=torch.int32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(args.calib_copy):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

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
        "sphere128": c1_leg,
        "vertices": count,
        "triangles": count // 3,
        "mtri_per_s": round(count / 3 / (ms_per_step * 1e-3) / 1e6, 1),
        "kernel_ms_avg": round(avg_kernel_ms, 4),
        "per_call_event_ms": per_call_ms,
        "algorithmic_bytes": alg_bytes,
        "full_api_ms": round(full_ms, 4),
        "integrate_ms": round(integrate_ms, 2),
        "integrate_gvoxel_frames_per_s": round(nvox * args.frames / (integrate_ms * 1e-3) / 1e9, 2),
        # the fused sweep's algorithmic bytes: 8 X Y Z (read + write) + 2 W H K
        # (depth); its time goes to VALU issue (the frames' projections and
        # divisions per voxel), not to memory
        "integrate_roofline": {
            "bound": "hbm", "algorithmic_bytes": 8 * nvox + 2 * cfg["W"] * cfg["H"] * args.frames,
            "achieved": round((8 * nvox + 2 * cfg["W"] * cfg["H"] * args.frames)
                              / (integrate_ms * 1e-3) / 1e9, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round((8 * nvox + 2 * cfg["W"] * cfg["H"] * args.frames)
                          / (integrate_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
            "note": "VALU-issue bound"},
        "integrate_per_frame": per_frame,
        "fwd_bwd": diff_leg,
        "batched": batch_leg,
        "fusion_graph": graph_leg,
        "raycast": ray_leg,
        "mesh_io": io_leg,
        "tuning": tuning,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
