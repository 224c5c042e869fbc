"""Extraction and integration past the 32-bit limits, byte for byte against
the oracle.

2048 x 2048 x 1040 short2 volumes (4.36 G voxels, 17.4 GB): voxel indices
past 2^32, 16 MiB planes.  At this plane size a column of the 128-bit count
kernel may take at most 124 planes (its buffer offsets stay below 2^31), so
the volume is cut into nine 116-plane slabs instead of the two that 56
columns per CU on 256 CUs would give; the integration kernels find voxel
coordinates past 2^32 with 64-bit divisions.  Volumes and frames are built on
the device and copied to the host for the oracle, so both see the same bytes.
"""
import numpy as np
import pytest
import torch

from marching_cube_amd import fusion, mesh, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

X, Y, Z = 2048, 2048, 1040
H_GRID = 1.0 / 1024
TSDF_SCALE = 32767


def _shapes_volume(dev, trunc, dims=(X, Y, Z),
                   spheres=(((0.55, 0.5, 0.5), 0.3), ((1.45, 1.45, 0.75), 0.2)),
                   box=((1.35, 0.55, 0.3), (0.25, 0.2, 0.12))):
    """Packed short2 (Z, Y, X) volume of the union of two spheres and a box:
    in the band |sdf| < trunc, tsdf = rint(sdf / trunc * 32767) (fp32 on the
    device) and weight 1 + (x + 2y + 3z) % 7; elsewhere tsdf +32767, weight
    0.  Built 16 planes at a time."""
    h = H_GRID
    X, Y, Z = dims
    vol = torch.empty((Z, Y, X), dtype=torch.int32, device=dev)
    xs = ((torch.arange(X, device=dev, dtype=torch.float32) + 0.5) * h).view(1, 1, X)
    ys = ((torch.arange(Y, device=dev, dtype=torch.float32) + 0.5) * h).view(1, Y, 1)
    xi = torch.arange(X, device=dev, dtype=torch.int32).view(1, 1, X)
    yi = torch.arange(Y, device=dev, dtype=torch.int32).view(1, Y, 1)
    bc, bh = box
    for z0 in range(0, Z, 16):
        z1 = min(z0 + 16, Z)
        zs = ((torch.arange(z0, z1, device=dev, dtype=torch.float32) + 0.5) * h).view(-1, 1, 1)
        zi = torch.arange(z0, z1, device=dev, dtype=torch.int32).view(-1, 1, 1)
        sdf = None
        for (cx, cy, cz), r in spheres:
            d = torch.sqrt((xs - cx) ** 2 + (ys - cy) ** 2 + (zs - cz) ** 2) - r
            sdf = d if sdf is None else torch.minimum(sdf, d)
        qx = (xs - bc[0]).abs() - bh[0]
        qy = (ys - bc[1]).abs() - bh[1]
        qz = (zs - bc[2]).abs() - bh[2]
        outside = torch.sqrt(qx.clamp(min=0) ** 2 + qy.clamp(min=0) ** 2 + qz.clamp(min=0) ** 2)
        inside = torch.maximum(torch.maximum(qx, qy), qz).clamp(max=0)
        sdf = torch.minimum(sdf, outside + inside)
        band = sdf.abs() < trunc
        t = torch.round(torch.clamp(sdf / trunc, -1.0, 1.0) * TSDF_SCALE).to(torch.int32)
        t = torch.where(band, t, torch.full_like(t, TSDF_SCALE))
        w = torch.where(band, 1 + (xi + 2 * yi + 3 * zi) % 7, torch.zeros_like(t))
        vol[z0:z1] = (t & 0xFFFF) | (w << 16)
        del sdf, band, t, w, outside, inside, qx, qy, qz
    return vol


def test_extraction_past_32bit_limits_bit_exact(dev):
    h = H_GRID
    print(f"[case] extraction, {X}x{Y}x{Z} short2 shapes volume: vertex_voxel, int3 map, "
          "ver_voxel", flush=True)
    assert X * Y * Z > 2 ** 32
    vol = _shapes_volume(dev, 3 * h)
    host = vol.cpu().numpy().view(np.uint32)
    ref = O.extract(host, h, want=("vv", "v2vox", "ver_voxel"), nthreads=16)
    del host
    n = ref["n"]
    assert n > 1_000_000
    ws = mesh.Workspace(dev)
    assert mesh.extract_into(vol, h, cap=0, workspace=ws) == n
    vv = torch.empty((n, 7), dtype=torch.int32, device=dev)
    v2 = torch.empty((n, 3), dtype=torch.int32, device=dev)
    verv = torch.empty((Z, Y, X), dtype=torch.int16, device=dev)
    assert mesh.extract_into(vol, h, vv=vv, v2vox=v2, ver_voxel=verv, cap=n, workspace=ws) == n
    del vol, ws
    assert vv.cpu().numpy().tobytes() == ref["vv"].view(np.int32).reshape(-1, 7).tobytes()
    assert np.array_equal(v2.cpu().numpy(), ref["v2vox"])
    assert np.array_equal(verv.cpu().numpy(), ref["ver_voxel"])
    del vv, v2, verv, ref
    torch.cuda.empty_cache()


def test_integration_past_32bit_limits_bit_exact(dev):
    """Two synthetic uint16 640x480 frames of a 12-object scene, fused one
    call per frame (the reference's pattern) and in one two-frame sweep;
    both volumes equal the oracle's."""
    h, trunc, max_w = H_GRID, 0.03, 128
    print(f"[case] integration, {X}x{Y}x{Z} short2, 2 frames 640x480 (per-frame calls and one "
          "sweep)", flush=True)
    intr = [525.0, 525.0, 319.5, 239.5]
    objs = synth.kinfu_scene(12, 1.0, seed=5)
    poses = synth.orbit_poses(2, 1.0, radius=1.3, seed=5)
    depth = fusion.synth_depth(intr, poses, objs, 640, 480, seed=5, device=dev)
    per_frame = fusion.voxel_new((X, Y, Z), trunc, h, device=dev)
    for k in range(2):
        fusion.integrate(per_frame, intr, poses[k:k + 1], depth[k:k + 1], trunc=trunc,
                         max_w=max_w)
    got = per_frame.data.cpu().numpy().view(np.uint32)
    del per_frame
    torch.cuda.empty_cache()
    ref = O.volume_init((Z, Y, X))
    O.integrate(ref, h, trunc, max_w, intr, poses, depth.cpu().numpy())
    assert (ref != O.volume_init((1, 1, 1))[0, 0, 0]).sum() > 10_000_000  # updated voxels
    assert np.array_equal(got, ref)
    del got
    sweep = fusion.voxel_new((X, Y, Z), trunc, h, device=dev)
    fusion.integrate(sweep, intr, poses, depth, trunc=trunc, max_w=max_w)
    assert np.array_equal(sweep.data.cpu().numpy().view(np.uint32), ref)
    del sweep, ref
    torch.cuda.empty_cache()


def test_batch_with_large_planes_bit_exact(dev):
    """A batch whose first volume has 4100 x 2048 planes (8.4 M voxels): its
    columns take 59 planes instead of the plan's 64 (so that the 128-bit
    count kernel still reaches them through one buffer resource), the small
    volumes after it 64; every segment equals the oracle's extraction."""
    from marching_cube_amd import batch
    print("[case] batch: 4100x2048x70 volume + 5 small shape volumes", flush=True)
    big = _shapes_volume(dev, 3 * H_GRID, dims=(4100, 2048, 70),
                         spheres=(((0.6, 0.5, 0.035), 0.03), ((2.5, 1.2, 0.03), 0.05)),
                         box=((3.2, 0.8, 0.04), (0.4, 0.3, 0.02)))
    smalls, hs = synth.batch_shapes(5, lo=20, hi=64, seed=21)
    vols = [big] + [torch.from_numpy(v.view(np.int32)).to(dev) for v in smalls]
    hs = [H_GRID] + [float(x) for x in hs]
    plan = batch.BatchPlan(vols, hs)
    n, vbase, out = plan.extract(want=("vv",))
    plan.check()
    vb = vbase.cpu().numpy()
    vv = out["vv"].cpu().numpy()
    hosts = [big.cpu().numpy().view(np.uint32)] + smalls
    for b, (v, h) in enumerate(zip(hosts, hs)):
        ref = O.extract(v, float(h), want=("vv",), nthreads=16)
        assert ref["n"] > (1_000_000 if b == 0 else 0), b
        assert vb[b + 1] - vb[b] == ref["n"], b
        assert vv[vb[b]:vb[b + 1]].tobytes() == ref["vv"].view(np.int32).reshape(-1, 7).tobytes(), b
    assert n == vb[-1]
