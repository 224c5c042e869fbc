"""Host-side mirror of the reference's TSDF volume calls, on torch device
memory, calling libmcx through the C ABI.

  cu_voxel_new(v, size, min_t, max_t, grid_s)   cu_kinect_fusion_v1.cu:381-398
      -> voxel_new(size, min_t, grid_s, max_t=...)   allocates and fills (-32766, 0)
  _cu_voxel_volumetric_integration(v, ci, ce, size, depth, ...)   :534-549
      -> integrate(volume, intr, poses, depth_frames)  K frames in one sweep
  _cu_voxel_extract_depth_and_normal_map(v, ci, ce, size, d, n, c)   :972-986
      -> raycast(volume, intr, pose, W, H)  depth + normal maps (no colour)
Camera model (cu_geometry_functions.h:96-129): intrinsics {fx, fy, cx, cy};
poses camera-to-world {R row-major, t}; q = R^T (p - t).
"""
import ctypes

import torch

from . import _capi
from .mesh import CuVoxel, _ptr, _stream_handle

def _lib():
    return _capi.load()


def _host_f32(a):
    a = torch.as_tensor(a, dtype=torch.float32).contiguous().cpu()
    return a, ctypes.cast(a.data_ptr(), ctypes.c_void_p)


def voxel_init(vol):
    """Fill a packed short2 volume with (tsdf -32766, w 0) (d_voxel_initialize)."""
    Z, Y, X = vol.shape[:3]
    _capi.check(_lib().mcx_volume_init(_ptr(vol), X, Y, Z, _stream_handle(vol.device)),
                "mcx_volume_init")
    return vol


def voxel_new(size, trunc, grid_s, device="cuda", max_t=None):
    """cu_voxel_new: a (Z, Y, X) packed short2 volume initialised to (-32766, 0),
    truncated at min_t = trunc behind the surface and max_t (default trunc)
    in front of it."""
    X, Y, Z = size
    vol = torch.empty((Z, Y, X), dtype=torch.int32, device=device)
    voxel_init(vol)
    return CuVoxel(data=vol, grid_s=grid_s, min_t=trunc, max_t=trunc if max_t is None else max_t)


def integrate(volume, intr, poses, depth, *, trunc=None, min_t=None, max_t=None, max_w=128,
              depth_scale=0.001, h=None):
    """Fuse K depth frames ((K, H, W) device tensor: uint16, or float32 in
    units of depth_scale metres) into `volume` (CuVoxel or packed int32
    tensor) in one sweep.  poses: (K, 12) float32, camera-to-world.
    Truncation: min_t behind the surface and max_t in front of it, as the
    reference's cu_voxel fields (cu_kinect_fusion_v1.cu:455-463); `trunc`
    sets both, else they come from the arguments or from the CuVoxel."""
    if isinstance(volume, CuVoxel):
        vol, h = volume.data, volume.grid_s if h is None else h
        min_t = volume.min_t if min_t is None else min_t
        max_t = volume.max_t if max_t is None else max_t
    else:
        vol = volume
    if trunc is not None:
        min_t = max_t = trunc
    if h is None:
        raise ValueError("grid size h is required")
    if min_t is None or max_t is None:
        raise ValueError("truncation (trunc, or min_t and max_t) is required")
    Z, Y, X = vol.shape
    if depth.dtype in (torch.uint16, torch.int16):
        kind = _capi.DEPTH_U16
    elif depth.dtype == torch.float32:
        kind = _capi.DEPTH_F32
    else:
        raise ValueError("depth frames must be uint16 or float32 (K, H, W)")
    K, H, W = depth.shape
    poses = torch.as_tensor(poses, dtype=torch.float32).reshape(K, 12).to(vol.device).contiguous()
    lib = _lib()
    ws = torch.empty(max(1, lib.mcx_integrate_workspace_size(W, H)), dtype=torch.uint8,
                     device=vol.device)
    hi, hp = _host_f32(intr)
    _capi.check(lib.mcx_integrate_ex(_ptr(vol), X, Y, Z, float(h), float(min_t), float(max_t),
                                     int(max_w), hp, _ptr(poses), _ptr(depth.contiguous()), kind,
                                     float(depth_scale), W, H, K, _ptr(ws), ws.numel(),
                                     _stream_handle(vol.device)),
                "mcx_integrate_ex")
    del hi
    return volume


def raycast(volume, intr, pose, W, H, *, h=None, origin=0.5, depth=None, normals=None):
    """Ray-march a depth map (H, W) float32 (metres, 0 = no hit) and
    camera-frame normals (H, W, 3) from a packed short2 volume seen from
    `pose` (12 floats camera-to-world {R row-major, t}).  origin: voxel i at
    world (i + origin) h -- 0.5 for volumes fused by `integrate`, 0 for the
    reference's own convention.  Raises on a singular camera matrix."""
    if isinstance(volume, CuVoxel):
        vol, h = volume.data, volume.grid_s if h is None else h
    else:
        vol = volume
    if h is None:
        raise ValueError("grid size h is required")
    Z, Y, X = vol.shape[:3]
    dev = vol.device
    if depth is None:
        depth = torch.empty((H, W), dtype=torch.float32, device=dev)
    if normals is None:
        normals = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
    hi, hp = _host_f32(intr)
    pi, pp = _host_f32(torch.as_tensor(pose, dtype=torch.float32).reshape(12))
    _capi.check(_lib().mcx_raycast(_ptr(vol), X, Y, Z, float(h), float(origin), hp, pp, W, H,
                                   _ptr(depth), _ptr(normals), _stream_handle(dev)),
                "mcx_raycast")
    del hi, pi
    return depth, normals


def synth_depth(intr, poses, objs, W, H, *, sigma=0.0015, dropout=0.02, seed=20261202,
                depth_unit=0.001, device="cuda"):
    """Ray-cast synthetic uint16 depth frames (K, H, W) of a sphere/box scene."""
    poses = torch.as_tensor(poses, dtype=torch.float32).reshape(-1, 12).to(device).contiguous()
    objs = torch.as_tensor(objs, dtype=torch.float32).reshape(-1, 8).to(device).contiguous()
    K = poses.shape[0]
    depth = torch.empty((K, H, W), dtype=torch.uint16, device=device)
    hi, hp = _host_f32(intr)
    _capi.check(_lib().mcx_synth_depth(_ptr(depth), W, H, K, hp, _ptr(poses), _ptr(objs),
                                       objs.shape[0], float(sigma), float(dropout),
                                       int(seed) & 0xFFFFFFFF, float(depth_unit),
                                       _stream_handle(device)), "mcx_synth_depth")
    del hi
    return depth


# The KinectFusion-scale workload: 512^3 short2 over a 3 m cube, fused from
# 300 640x480 frames of a 40-object scene, trunc 30 mm, max weight 128.
KINFU = dict(dim=512, extent=3.0, frames=300, W=640, H=480, fx=525.0, fy=525.0, cx=319.5,
             cy=239.5, trunc=0.03, max_w=128, sigma=0.0015, dropout=0.02, nobj=40,
             seed=20261202)


def kinfu_volume(dim=None, frames=None, device="cuda", cfg=KINFU, return_frames=False):
    """Build the fused KinFu-scale volume on the device (synthesise frames,
    init, integrate).  Returns (CuVoxel, frames, poses) when return_frames."""
    from . import synth
    dim = cfg["dim"] if dim is None else dim
    frames = cfg["frames"] if frames is None else frames
    h = cfg["extent"] / dim
    intr = [cfg["fx"], cfg["fy"], cfg["cx"], cfg["cy"]]
    objs = synth.kinfu_scene(cfg["nobj"], cfg["extent"], cfg["seed"])
    poses = synth.orbit_poses(frames, cfg["extent"], seed=cfg["seed"])
    depth = synth_depth(intr, poses, objs, cfg["W"], cfg["H"], sigma=cfg["sigma"],
                        dropout=cfg["dropout"], seed=cfg["seed"], device=device)
    vox = voxel_new((dim, dim, dim), cfg["trunc"], h, device=device)
    integrate(vox, intr, poses, depth, trunc=cfg["trunc"], max_w=cfg["max_w"])
    if return_frames:
        return vox, depth, poses
    return vox
