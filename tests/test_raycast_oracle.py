"""Pins the CPU oracle of the raycast (depth + normal maps) with closed forms:
* a tilted plane seen by a rotated, off-centre camera inside and outside
  the volume: every hit pixel's
  z-depth equals the analytic ray/plane intersection, every normal equals
  R^T of the plane's unit normal (pointing to the camera), and every pixel
  whose intersection lies well inside the volume is hit;
* the same with voxel i placed at (i + 0.5) h (origin 0.5);
* a sphere of radius 20 voxels: depths within 0.05 voxel of the analytic
  ray/sphere z-depth, normals within 0.04 of the radial direction (0.012 on
  average); both errors are the trilinear field's curvature error, which
  shrinks with the radius (checked at 10, 20 and 30 voxels);
* a camera inside a solid reports nothing (back faces end the rays); free
  space reports nothing; a singular camera matrix is refused;
* the principal ray of an axis-aligned camera (two ray components exactly
  zero, replaced by 1e-15) still hits and the march ends.
Volumes use the raycast's grid: voxel i at grid coordinate i."""
import numpy as np
import pytest

from oracle import oracle as O

MAXT = 32766


def pack(tsdf_unit, weight=1):
    s = np.clip(np.rint(np.clip(tsdf_unit, -1, 1) * MAXT), -MAXT, MAXT).astype(np.int16)
    w = np.full(s.shape, weight, np.uint16)
    return (s.view(np.uint16).astype(np.uint32) | (w.astype(np.uint32) << 16))


def rot(ax, ay, az):
    cx, sx, cy, sy, cz, sz = np.cos(ax), np.sin(ax), np.cos(ay), np.sin(ay), np.cos(az), np.sin(az)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def camera_rays(intr, R, W, H):
    u, v = np.meshgrid(np.arange(W, dtype=np.float64), np.arange(H, dtype=np.float64))
    dc = np.stack([(u - intr[2]) / intr[0], (v - intr[3]) / intr[1], np.ones_like(u)], -1)
    return dc @ R.T  # world directions with camera z = 1


def grid(N):
    g = np.arange(N, dtype=np.float64)
    return np.meshgrid(g, g, g, indexing="ij")  # z, y, x


INTR = np.array([80.0, 84.0, 33.3, 22.7], np.float32)
W, H = 64, 48


def plane_case(origin, back, N=48, T=3.0):
    h = 0.01
    n = np.array([0.3, -0.2, 1.0])
    n /= np.linalg.norm(n)
    z, y, x = grid(N)
    d = n @ np.array([N / 2, N / 2, N / 2 + 4])
    # grid coordinates of voxel i are i; tsdf > 0 on the camera side (n.g < d)
    vol = pack((d - (n[0] * x + n[1] * y + n[2] * z)) / T)
    R = rot(0.15, -0.2, 0.3)
    fwd = R[:, 2]  # camera z axis in the world
    cam_g = np.array([N / 2, N / 2, N / 2]) - fwd * back  # looking at the plane
    pose = np.concatenate([R.reshape(-1), (cam_g + origin) * h]).astype(np.float32)
    return vol, h, pose, n, d, N


@pytest.mark.parametrize("origin", [0.0, 0.5])
@pytest.mark.parametrize("back", [14, 40])  # camera inside / outside the volume
def test_plane_depth_and_normal_closed_form(origin, back):
    vol, h, pose, n, d, N = plane_case(origin, back)
    depth, normals = O.raycast(vol, h, INTR, pose, W, H, origin=origin)
    R = pose[:9].reshape(3, 3).astype(np.float64)
    cc = pose[9:].astype(np.float64) / h - origin
    dw = camera_rays(INTR.astype(np.float64), R, W, H)
    s = (d - n @ cc) / (dw @ n)  # camera-z distance in voxels
    hitp = cc + s[..., None] * dw
    hit = depth > 0
    inner = (s > 0) & np.all((hitp > 4) & (hitp < N - 5), axis=-1)
    assert inner.sum() > 0.5 * W * H
    assert hit[inner].all()
    err = np.abs(depth[hit] - s[hit] * h)
    assert err.max() < 2e-3 * h, err.max() / h
    nc = R.T @ (-n)
    assert np.abs(normals[hit] - nc).max() < 1e-3
    assert (normals[~hit] == 0).all()


def sphere_case(N=64, radius=20.0, T=3.0):
    c = np.array([N / 2 + 1.3, N / 2 - 0.7, N / 2 + 0.4])
    z, y, x = grid(N)
    dist = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - radius
    return pack(dist / T), c, radius


def test_sphere_depth_and_normals():
    N = 64
    vol, c, rad = sphere_case(N)
    h = 0.005
    R = rot(0.0, 0.25, 0.05)
    cam_g = c - R[:, 2] * 28
    pose = np.concatenate([R.reshape(-1), cam_g * h]).astype(np.float32)
    depth, normals = O.raycast(vol, h, INTR, pose, W, H, origin=0.0)
    dw = camera_rays(INTR.astype(np.float64), R, W, H)
    oc = cam_g - c
    a = (dw * dw).sum(-1)
    b = 2 * (dw @ oc)
    cq = oc @ oc - rad ** 2
    disc = b * b - 4 * a * cq
    s = np.where(disc > 0, (-b - np.sqrt(np.maximum(disc, 0))) / (2 * a), np.nan)
    hit = depth > 0
    solid = disc > 0.05 * (4 * a * rad ** 2)  # rays not grazing the silhouette
    assert solid.sum() > 200
    assert hit[solid].all()
    assert not hit[disc < 0].any()
    err = np.abs(depth[hit] / h - s[hit])
    assert err.max() < 0.05, err.max()
    p = cam_g + s[..., None] * dw
    radial = (p - c) / rad
    nc = radial @ R  # R^T radial
    nerr = np.abs(normals[solid] - nc[solid])
    assert nerr.max() < 0.04 and nerr.mean() < 0.012, (nerr.max(), nerr.mean())


def test_sphere_error_shrinks_with_radius():
    errs = []
    for rad in (10.0, 20.0, 30.0):
        N = 96
        vol, c, _ = sphere_case(N, radius=rad)
        h = 0.005
        R = rot(0.0, 0.25, 0.05)
        cam_g = c - R[:, 2] * (rad + 20)
        pose = np.concatenate([R.reshape(-1), cam_g * h]).astype(np.float32)
        depth, normals = O.raycast(vol, h, INTR, pose, W, H)
        dw = camera_rays(INTR.astype(np.float64), R, W, H)
        oc = cam_g - c
        a, b = (dw * dw).sum(-1), 2 * (dw @ oc)
        disc = b * b - 4 * a * (oc @ oc - rad ** 2)
        solid = disc > 0.05 * (4 * a * rad ** 2)
        s = (-b - np.sqrt(np.maximum(disc, 0))) / (2 * a)
        nc = ((cam_g + s[..., None] * dw - c) / rad) @ R
        errs.append((np.abs(depth[solid] / h - s[solid]).mean(),
                     np.abs(normals[solid] - nc[solid]).mean()))
    for k in range(2):
        assert errs[0][k] > errs[1][k] > errs[2][k], errs


def test_camera_inside_solid_sees_nothing():
    N = 48
    vol, c, rad = sphere_case(N, radius=16.0)
    h = 0.01
    R = rot(0.1, 0.2, 0.3)
    pose = np.concatenate([R.reshape(-1), c * h]).astype(np.float32)  # at the sphere centre
    depth, normals = O.raycast(vol, h, INTR, pose, W, H)
    assert (depth == 0).all() and (normals == 0).all()


def test_free_space_and_singular_camera():
    N = 32
    vol = pack(np.ones((N, N, N)))
    pose = np.concatenate([np.eye(3).reshape(-1), [0.16, 0.16, 0.02]]).astype(np.float32)
    depth, normals = O.raycast(vol, 0.01, INTR, pose, W, H)
    assert (depth == 0).all()
    bad = INTR.copy()
    bad[0] = 0.0
    assert O.raycast(vol, 0.01, bad, pose, W, H) is None
    assert O.raycast_setup(bad, pose) is None


def test_axis_aligned_principal_ray():
    N = 40
    z, y, x = grid(N)
    z0 = 27.3
    vol = pack((z0 - z) / 3.0)
    intr = np.array([50.0, 50.0, 16.0, 12.0], np.float32)
    h = 0.02
    cam = np.array([20.0, 20.0, 5.0])
    pose = np.concatenate([np.eye(3).reshape(-1), cam * h]).astype(np.float32)
    im = O.raycast_setup(intr, pose)
    ray = im @ np.array([16.0, 12.0, 1.0], np.float32)
    assert ray[0] == 0 and ray[1] == 0  # the principal ray is exactly axis-aligned
    depth, normals = O.raycast(vol, h, intr, pose, 32, 24)
    assert (depth > 0).all()
    assert np.abs(depth - (z0 - cam[2]) * h).max() < 1e-3 * h
    assert np.abs(normals - np.array([0, 0, -1])).max() < 1e-3
