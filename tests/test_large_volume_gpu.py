"""Extraction past the 32-bit limits, byte for byte against the oracle.

2048 x 1024 x 1040 short2 (2.18 G voxels, 8.7 GB): voxel indices past 2^31,
byte offsets past 2^33, 8 MiB planes.  At this plane size a column of the
128-bit count kernel may take at most 252 planes (its buffer offsets stay
below 2^31), so the volume is cut into five 208-plane slabs instead of the
four that 56 columns per CU on 256 CUs would give.  The volume is built on the device
(two spheres and a box, fp32 signed distances) and copied to the host for
the oracle, so both see the same bytes.
"""
import numpy as np
import pytest
import torch

from marching_cube_amd import mesh
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TSDF_SCALE = 32767


def _shapes_volume(dev, X, Y, Z, h, trunc):
    """Packed short2 (Z, Y, X) volume of the union of two spheres and a box:
    in the band |sdf| < trunc, tsdf = rint(sdf / trunc * 32767) (fp32 on the
    device) and weight 1 + (x + 2y + 3z) % 7; elsewhere tsdf +32767, weight
    0.  Built 32 planes at a time."""
    vol = torch.empty((Z, Y, X), dtype=torch.int32, device=dev)
    xs = ((torch.arange(X, device=dev, dtype=torch.float32) + 0.5) * h).view(1, 1, X)
    ys = ((torch.arange(Y, device=dev, dtype=torch.float32) + 0.5) * h).view(1, Y, 1)
    xi = torch.arange(X, device=dev, dtype=torch.int32).view(1, 1, X)
    yi = torch.arange(Y, device=dev, dtype=torch.int32).view(1, Y, 1)
    spheres = (((0.55, 0.5, 0.5), 0.3), ((1.45, 0.45, 0.75), 0.2))
    bc, bh = (1.35, 0.55, 0.3), (0.25, 0.2, 0.12)
    for z0 in range(0, Z, 32):
        z1 = min(z0 + 32, Z)
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


def test_volume_past_32bit_limits_bit_exact(dev):
    X, Y, Z = 2048, 1024, 1040
    h = 1.0 / 1024
    print(f"[case] {X}x{Y}x{Z} short2 shapes volume, vertex_voxel + int3 map + ver_voxel",
          flush=True)
    assert X * Y * Z > 2 ** 31
    vol = _shapes_volume(dev, X, Y, Z, h, 3 * h)
    host = vol.cpu().numpy().view(np.uint32)
    ref = O.extract(host, h, want=("vv", "v2vox", "ver_voxel"), nthreads=16)
    n = ref["n"]
    assert n > 1_000_000
    ws = mesh.Workspace(dev)
    assert mesh.extract_into(vol, h, cap=0, workspace=ws) == n
    vv = torch.empty((n, 7), dtype=torch.int32, device=dev)
    v2 = torch.empty((n, 3), dtype=torch.int32, device=dev)
    verv = torch.empty((Z, Y, X), dtype=torch.int16, device=dev)
    assert mesh.extract_into(vol, h, vv=vv, v2vox=v2, ver_voxel=verv, cap=n, workspace=ws) == n
    assert vv.cpu().numpy().tobytes() == ref["vv"].view(np.int32).reshape(-1, 7).tobytes()
    assert np.array_equal(v2.cpu().numpy(), ref["v2vox"])
    assert np.array_equal(verv.cpu().numpy(), ref["ver_voxel"])
