"""Synthetic TSDF inputs of the benchmark shapes (numpy, host side).

These generate the inputs the extraction and integration entry points are
measured and tested on; nothing here is on the device path.

* sphere_volume   -- 128^3 short2 TSDF of a sphere r=0.3 m in a 1 m cube,
                     trunc = 3 voxels, w=1 inside the band, (+32767, 0) outside.
* noise_volume    -- dense random short2 volume (every cube case appears).
* bumps_volume    -- fp32 field: sum of seeded Gaussian bumps + small noise.
* batch_shapes    -- uneven short2 volumes holding a sphere / box / torus SDF
                     at a random pose (sides log-uniform per axis).
* kinfu_scene / orbit_poses -- seeded spheres/boxes scene and an orbiting
                     camera for the depth-frame synthesis kernel.

Packing of a short2 voxel into one uint32: low 16 bits = int16 tsdf, high 16
bits = uint16 weight (cu_voxel_t, structure.h:14-18).
"""
import math

import numpy as np

TSDF_SCALE = 32767.0


def pack_short2(tsdf, w):
    t = np.asarray(tsdf).astype(np.int16).view(np.uint16).astype(np.uint32)
    return t | (np.asarray(w).astype(np.uint32) << 16)


def unpack_short2(vol):
    v = np.asarray(vol, dtype=np.uint32)
    return (v & 0xFFFF).astype(np.uint16).view(np.int16), (v >> 16).astype(np.uint16)


def _node_coords(n, h):
    return (np.arange(n, dtype=np.float64) + 0.5) * h


def sdf_to_short2(sdf, trunc):
    """Band |sdf| < trunc: w=1, tsdf = lrint(clamp(sdf/trunc,-1,1)*32767) in
    FP64; outside: tsdf=+32767, w=0."""
    band = np.abs(sdf) < trunc
    t = np.rint(np.clip(sdf / trunc, -1.0, 1.0) * TSDF_SCALE)
    tsdf = np.where(band, t, TSDF_SCALE).astype(np.int16)
    w = band.astype(np.uint16)
    return pack_short2(tsdf, w)


def sphere_volume(n=128, radius=0.3, extent=1.0, trunc_vox=3.0, center=None):
    h = extent / n
    c = (0.5 * extent,) * 3 if center is None else center
    x = _node_coords(n, h)
    zz, yy, xx = np.meshgrid(x, x, x, indexing="ij")
    sdf = np.sqrt((xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2) - radius
    return sdf_to_short2(sdf, trunc_vox * h), h


def noise_volume(shape, seed=0, p_zero_w=0.0, max_w=128):
    """Dense random short2 volume; tsdf uniform in [-32767, 32767]."""
    rng = np.random.default_rng(seed)
    Z, Y, X = shape
    t = rng.integers(-32767, 32768, size=(Z, Y, X), dtype=np.int32)
    w = rng.integers(1, max_w + 1, size=(Z, Y, X), dtype=np.int32)
    if p_zero_w > 0:
        w[rng.random((Z, Y, X)) < p_zero_w] = 0
    return pack_short2(t, w)


def bumps_volume(n=256, nbumps=64, seed=20261202, noise_vox=0.1, dtype=np.float32):
    """fp32 implicit field on [0,1]^3 (h = 1/n): f = 0.5 - sum_k a_k exp(-|p-c_k|^2/(2 s_k^2))
    plus band-limited noise of amplitude noise_vox voxels (in distance units)."""
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    x = _node_coords(n, h).astype(np.float32)
    c = rng.uniform(0.2, 0.8, size=(nbumps, 3)).astype(np.float32)
    s = rng.uniform(0.05, 0.12, size=nbumps).astype(np.float32)
    a = rng.uniform(0.4, 1.0, size=nbumps).astype(np.float32)
    f = np.full((n, n, n), 0.5, np.float32)
    for k in range(nbumps):
        gx = np.exp(-((x - c[k, 0]) ** 2) / (2 * s[k] ** 2))
        gy = np.exp(-((x - c[k, 1]) ** 2) / (2 * s[k] ** 2))
        gz = np.exp(-((x - c[k, 2]) ** 2) / (2 * s[k] ** 2))
        f -= a[k] * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    # band-limited noise: a few low-frequency cosines, amplitude noise_vox*h
    k = rng.integers(1, 6, size=(8, 3))
    ph = rng.uniform(0, 2 * math.pi, size=8)
    noise = np.zeros_like(f)
    for i in range(8):
        noise += (np.cos(2 * math.pi * k[i, 2] * x + ph[i])[:, None, None]
                  * np.cos(2 * math.pi * k[i, 1] * x)[None, :, None]
                  * np.cos(2 * math.pi * k[i, 0] * x)[None, None, :])
    f += (noise_vox * h / 8.0) * noise
    return f.astype(dtype), h


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def shape_params(dims, kind, rng):
    """Random pose and size of one shape in a (Z,Y,X) volume with h = 1/max(dims):
    centre c (x,y,z), rotation R (world = R @ shape frame) and size parameters."""
    Z, Y, X = dims
    h = 1.0 / max(dims)
    ext = np.array([X, Y, Z], np.float64) * h
    c = ext * rng.uniform(0.35, 0.65, size=3)
    R = _random_rotation(rng)
    m = 0.5 * ext.min()
    if kind == "sphere":
        size = np.array([m * rng.uniform(0.4, 0.8), 0.0, 0.0])
    elif kind == "box":
        size = m * rng.uniform(0.3, 0.7, size=3)
    else:  # torus: major radius, minor radius
        R0 = m * rng.uniform(0.4, 0.6)
        size = np.array([R0, R0 * rng.uniform(0.25, 0.5), 0.0])
    return h, c, R, size


def _shape_sdf(xp, p, kind, size):
    """SDF of a shape at the origin of its frame; p (..., 3); xp = numpy or torch."""
    if kind == "sphere":
        return xp.sqrt((p * p).sum(-1)) - size[0]
    if kind == "box":
        q = xp.abs(p) - xp.asarray(size) if xp is np else p.abs() - p.new_tensor(size)
        qp = q.clip(0, None) if xp is np else q.clamp(min=0)
        qmax = q.max(-1) if xp is np else q.max(-1).values
        return xp.sqrt((qp * qp).sum(-1)) + (qmax.clip(None, 0) if xp is np else qmax.clamp(max=0))
    qx = xp.sqrt(p[..., 0] ** 2 + p[..., 1] ** 2) - size[0]
    return xp.sqrt(qx ** 2 + p[..., 2] ** 2) - size[1]


def shape_volume(dims, kind, rng, trunc_vox=3.0):
    """One (Z,Y,X) short2 volume with h = 1/max(dims) holding a shape SDF."""
    Z, Y, X = dims
    h, c, R, size = shape_params(dims, kind, rng)
    zz, yy, xx = np.meshgrid(_node_coords(Z, h), _node_coords(Y, h), _node_coords(X, h),
                             indexing="ij")
    p = np.stack([xx - c[0], yy - c[1], zz - c[2]], axis=-1) @ R  # into shape frame
    return sdf_to_short2(_shape_sdf(np, p, kind, size), trunc_vox * h), h


def shape_volume_device(dims, kind, rng, device, trunc_vox=3.0):
    """shape_volume computed with torch (FP64) on `device`: a (Z,Y,X) int32
    tensor of packed short2 voxels.  Same shape parameters as shape_volume for
    the same rng state; values may differ from it by one tsdf unit where the
    FP64 arithmetic rounds differently."""
    import torch
    Z, Y, X = dims
    h, c, R, size = shape_params(dims, kind, rng)
    f64 = dict(dtype=torch.float64, device=device)
    x = (torch.arange(X, **f64) + 0.5) * h - c[0]
    y = (torch.arange(Y, **f64) + 0.5) * h - c[1]
    z = (torch.arange(Z, **f64) + 0.5) * h - c[2]
    zz, yy, xx = torch.meshgrid(z, y, x, indexing="ij")
    p = torch.stack([xx, yy, zz], dim=-1) @ torch.as_tensor(R, **f64)
    sdf = _shape_sdf(torch, p, kind, size)
    trunc = trunc_vox * h
    band = sdf.abs() < trunc
    t = torch.round((sdf / trunc).clamp(-1.0, 1.0) * TSDF_SCALE)
    tsdf = torch.where(band, t, torch.full_like(t, TSDF_SCALE)).to(torch.int32)
    return ((tsdf & 0xFFFF) | (band.to(torch.int32) << 16)), h


def batch_dims(count, lo=16, hi=128, seed=20261202):
    rng = np.random.default_rng(seed)
    d = np.exp(rng.uniform(math.log(lo), math.log(hi), size=(count, 3)))
    return np.clip(np.rint(d), lo, hi).astype(np.int64)  # rows are (Z, Y, X)


def batch_shapes(count, lo=16, hi=128, seed=20261202, device=None):
    """`count` uneven short2 volumes (numpy uint32 arrays, or int32 device
    tensors when `device` is given) and their grid sizes."""
    dims = batch_dims(count, lo, hi, seed)
    rng = np.random.default_rng(seed + 1)
    kinds = ("sphere", "box", "torus")
    vols, hs = [], []
    for i in range(count):
        d, kind = tuple(int(a) for a in dims[i]), kinds[rng.integers(0, 3)]
        if device is None:
            v, h = shape_volume(d, kind, rng)
        else:
            v, h = shape_volume_device(d, kind, rng, device)
        vols.append(v)
        hs.append(h)
    return vols, np.array(hs, np.float32)


def kinfu_scene(nobj=40, extent=3.0, seed=20261202):
    """Scene of random spheres and axis-aligned boxes inside [0, extent]^3.
    Returns float32 array (nobj, 8): kind (0 sphere, 1 box), cx, cy, cz,
    r (sphere radius) or half sizes hx, hy, hz (box), padding."""
    rng = np.random.default_rng(seed)
    s = extent / 3.0
    objs = np.zeros((nobj, 8), np.float32)
    for i in range(nobj):
        kind = rng.integers(0, 2)
        c = rng.uniform(0.2 * extent, 0.8 * extent, size=3)
        objs[i, 0] = kind
        objs[i, 1:4] = c
        if kind == 0:
            objs[i, 4] = rng.uniform(0.08, 0.3) * s
        else:
            objs[i, 4:7] = rng.uniform(0.06, 0.25, size=3) * s
    return objs


def orbit_poses(nframes, extent=3.0, radius=2.6, seed=20261202):
    """Camera-to-world poses (nframes, 12): R row-major (columns = camera x, y,
    z axes in world) and t, orbiting the cube centre and looking at it."""
    rng = np.random.default_rng(seed + 7)
    c = np.full(3, extent / 2.0)
    poses = np.zeros((nframes, 12), np.float32)
    for k in range(nframes):
        th = 2 * math.pi * k / max(nframes, 1)
        el = 0.35 * math.sin(3 * th) + rng.normal(0, 0.01)
        eye = c + radius * np.array([math.cos(th) * math.cos(el), math.sin(el),
                                     math.sin(th) * math.cos(el)])
        fwd = c - eye
        fwd /= np.linalg.norm(fwd)
        up = np.array([0.0, 1.0, 0.0])
        right = np.cross(fwd, up)
        right /= np.linalg.norm(right)
        down = np.cross(fwd, right)
        R = np.stack([right, down, fwd], axis=1)  # columns: camera x, y, z in world
        poses[k, :9] = R.reshape(-1)
        poses[k, 9:] = eye
    return poses
