"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It loads oracle/liboracle.so (built by oracle/Makefile) and
wraps the C restatement of the reference's marching cubes
(cu_marching_cubes_v1.cu), normals (cu_kinect_fusion_v1.cu:3146-3242) and
TSDF integration (cu_kinect_fusion_v1.cu:364-549) for numpy arrays.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ELEM_SHORT2 = 0
ELEM_F32 = 1
ELEM_F64 = 2  # finite-difference checks only

# vertex_voxel (structure.h:29-35), 28 bytes
VERTEX_VOXEL = np.dtype([("pos", "<f4", (3,)), ("vox", "<i4", (3,)),
                         ("tsdf", "<i2"), ("w", "<i2")])
assert VERTEX_VOXEL.itemsize == 28

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, i32, f32, f64 = (ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                  ctypes.c_float, ctypes.c_double)
        L.orc_classify.restype = i64
        L.orc_classify.argtypes = [vp, i32, i32, i32, i32, vp, vp]
        L.orc_extract.restype = i64
        L.orc_extract.argtypes = [vp, i32, i32, i32, i32, f32, vp, vp, vp, vp, i64, i32]
        L.orc_extract_f64.restype = i64
        L.orc_extract_f64.argtypes = [vp, i32, i32, i32, i32, f64, vp, vp, vp, i64]
        L.orc_extract_backward_f64.restype = None
        L.orc_extract_backward_f64.argtypes = [vp, i32, i32, i32, f64, vp, i64, vp]
        L.orc_normals.restype = None
        L.orc_normals.argtypes = [vp, i32, i32, i32, f32, vp, i64, i64, vp]
        L.orc_volume_init.restype = None
        L.orc_volume_init.argtypes = [vp, i64]
        L.orc_integrate_tsdf.restype = None
        L.orc_integrate_tsdf.argtypes = [vp, vp, i32, i32, i32, f32, f32, i32, vp, vp, vp, f32,
                                         i32, i32, i32]
        L.orc_integrate_tsdf_f64.restype = None
        L.orc_integrate_tsdf_f64.argtypes = [vp, vp, i32, i32, i32, f32, f32, i32, vp, vp, vp,
                                             ctypes.c_double, i32, i32, i32]
        L.orc_integrate_ex.restype = None
        L.orc_integrate_ex.argtypes = [vp, i32, i32, i32, f32, f32, f32, i32, vp, vp, vp, i32,
                                       f32, i32, i32, i32]
        L.orc_integrate.restype = None
        L.orc_integrate.argtypes = [vp, i32, i32, i32, f32, f32, i32, vp, vp, vp, f32,
                                    i32, i32, i32]
        L.orc_raycast_setup.restype = i32
        L.orc_raycast_setup.argtypes = [vp, vp, vp]
        L.orc_raycast.restype = i32
        L.orc_raycast.argtypes = [vp, i32, i32, i32, f32, f32, vp, vp, i32, i32, vp, vp, i32]
        L.orc_rle_encode_slice.restype = i64
        L.orc_rle_encode_slice.argtypes = [vp, i64, vp]
        L.orc_rle_decode_slice.restype = i32
        L.orc_rle_decode_slice.argtypes = [vp, i64, vp, i64]
        L.orc_rle_encode_dump.restype = i64
        L.orc_rle_encode_dump.argtypes = [vp, i32, i32, i32, i32, vp]
        L.orc_rle_decode_dump.restype = i32
        L.orc_rle_decode_dump.argtypes = [vp, i64, i32, i32, i32, i32, vp]
        for name in ("orc_tables_tri", "orc_tables_num_verts", "orc_tables_edge"):
            getattr(L, name).restype = ctypes.POINTER(ctypes.c_int)
            getattr(L, name).argtypes = []
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def _elem(vol):
    if vol.dtype == np.float32:
        return ELEM_F32
    if vol.dtype == np.float64:
        return ELEM_F64
    if vol.dtype in (np.uint32, np.int32) or (vol.dtype == np.int16 and vol.shape[-1] == 2):
        return ELEM_SHORT2
    raise TypeError(f"unsupported volume dtype {vol.dtype}")


def _dims(vol):
    if vol.dtype == np.int16:
        Z, Y, X, _ = vol.shape
    else:
        Z, Y, X = vol.shape
    return X, Y, Z


def tables():
    L = lib()
    tri = np.ctypeslib.as_array(L.orc_tables_tri(), shape=(256 * 16,)).reshape(256, 16).copy()
    nv = np.ctypeslib.as_array(L.orc_tables_num_verts(), shape=(256,)).copy()
    edge = np.ctypeslib.as_array(L.orc_tables_edge(), shape=(256,)).copy()
    return edge, tri, nv


def classify(vol):
    vol = np.ascontiguousarray(vol)
    X, Y, Z = _dims(vol)
    ci = np.zeros(X * Y * Z, np.uint8)
    nv = np.zeros(X * Y * Z, np.uint8)
    active = lib().orc_classify(_ptr(vol), _elem(vol), X, Y, Z, _ptr(ci), _ptr(nv))
    return ci.reshape(Z, Y, X), nv.reshape(Z, Y, X), int(active)


def extract(vol, h, want=("vv",), cap=None, nthreads=1):
    """fp32 reference-formula extraction.  Returns dict with the requested
    outputs among vv (structured VERTEX_VOXEL), pos (V,3), v2vox (V,3),
    ver_voxel (Z,Y,X int16) and n (total vertex count)."""
    vol = np.ascontiguousarray(vol)
    X, Y, Z = _dims(vol)
    L = lib()
    n = L.orc_extract(_ptr(vol), _elem(vol), X, Y, Z, h, None, None, None, None, 0, nthreads)
    c = n if cap is None else cap
    out = {"n": int(n)}
    vv = np.zeros(c, VERTEX_VOXEL) if "vv" in want else None
    pos = np.zeros((c, 3), np.float32) if "pos" in want else None
    v2 = np.zeros((c, 3), np.int32) if "v2vox" in want else None
    verv = np.zeros((Z, Y, X), np.int16) if "ver_voxel" in want else None
    L.orc_extract(_ptr(vol), _elem(vol), X, Y, Z, h, _ptr(vv), _ptr(pos), _ptr(v2),
                  _ptr(verv), c, nthreads)
    for k, a in (("vv", vv), ("pos", pos), ("v2vox", v2), ("ver_voxel", verv)):
        if a is not None:
            out[k] = a
    return out


def count(vol, nthreads=1):
    vol = np.ascontiguousarray(vol)
    X, Y, Z = _dims(vol)
    return int(lib().orc_extract(_ptr(vol), _elem(vol), X, Y, Z, 1.0, None, None, None,
                                 None, 0, nthreads))


def extract_f64(vol, h):
    vol = np.ascontiguousarray(vol)
    X, Y, Z = _dims(vol)
    L = lib()
    n = L.orc_extract_f64(_ptr(vol), _elem(vol), X, Y, Z, h, None, None, None, 0)
    pos = np.zeros((n, 3), np.float64)
    cube = np.zeros(n, np.int64)
    edge = np.zeros(n, np.int8)
    L.orc_extract_f64(_ptr(vol), _elem(vol), X, Y, Z, h, _ptr(pos), _ptr(cube), _ptr(edge), n)
    return pos, cube, edge


def extract_backward_f64(vol, h, gpos):
    vol = np.ascontiguousarray(vol, dtype=np.float32)
    Z, Y, X = vol.shape
    gpos = np.ascontiguousarray(gpos, dtype=np.float64)
    grad = np.zeros(X * Y * Z, np.float64)
    lib().orc_extract_backward_f64(_ptr(vol), X, Y, Z, h, _ptr(gpos), gpos.shape[0], _ptr(grad))
    return grad.reshape(Z, Y, X)


def normals(vol, h, pos):
    vol = np.ascontiguousarray(vol)
    X, Y, Z = _dims(vol)
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    out = np.zeros((pos.shape[0], 3), np.float32)
    lib().orc_normals(_ptr(vol), X, Y, Z, h, _ptr(pos), pos.shape[1], pos.shape[0], _ptr(out))
    return out


def volume_init(shape):
    vol = np.zeros(shape, np.uint32)
    lib().orc_volume_init(_ptr(vol), vol.size)
    return vol


def integrate(vol, h, trunc, max_w, intr, poses, depth, depth_scale=0.001):
    """In-place integration of uint16 depth frames (K,H,W) into vol (Z,Y,X uint32)."""
    assert vol.dtype == np.uint32 and vol.flags.c_contiguous
    Z, Y, X = vol.shape
    intr = np.ascontiguousarray(intr, np.float32)
    poses = np.ascontiguousarray(poses, np.float32).reshape(-1, 12)
    depth = np.ascontiguousarray(depth, np.uint16)
    K, H, W = depth.shape
    assert poses.shape[0] == K
    lib().orc_integrate(_ptr(vol), X, Y, Z, h, trunc, max_w, _ptr(intr), _ptr(poses),
                        _ptr(depth), depth_scale, W, H, K)
    return vol


def integrate_ex(vol, h, trunc_neg, trunc_pos, max_w, intr, poses, depth, depth_scale=0.001):
    """In-place integration with two truncation distances (the reference's
    min_t / max_t) of uint16 or float32 depth frames (K,H,W)."""
    assert vol.dtype == np.uint32 and vol.flags.c_contiguous
    Z, Y, X = vol.shape
    intr = np.ascontiguousarray(intr, np.float32)
    poses = np.ascontiguousarray(poses, np.float32).reshape(-1, 12)
    f32 = depth.dtype == np.float32
    depth = np.ascontiguousarray(depth, np.float32 if f32 else np.uint16)
    K, H, W = depth.shape
    assert poses.shape[0] == K
    lib().orc_integrate_ex(_ptr(vol), X, Y, Z, h, trunc_neg, trunc_pos, max_w, _ptr(intr),
                           _ptr(poses), _ptr(depth), int(f32), depth_scale, W, H, K)
    return vol


def integrate_tsdf(tsdf, w, h, trunc, max_w, intr, poses, depth, depth_scale=1.0):
    """In-place fp32 TSDF integration (mcx_integrate_tsdf semantics): tsdf
    float32 and w int32 (Z,Y,X), float depth frames (K,H,W)."""
    assert tsdf.dtype == np.float32 and w.dtype == np.int32
    assert tsdf.flags.c_contiguous and w.flags.c_contiguous and tsdf.shape == w.shape
    Z, Y, X = tsdf.shape
    intr = np.ascontiguousarray(intr, np.float32)
    poses = np.ascontiguousarray(poses, np.float32).reshape(-1, 12)
    depth = np.ascontiguousarray(depth, np.float32)
    K, H, W = depth.shape
    lib().orc_integrate_tsdf(_ptr(tsdf), _ptr(w), X, Y, Z, h, trunc, max_w, _ptr(intr),
                             _ptr(poses), _ptr(depth), depth_scale, W, H, K)
    return tsdf, w


def integrate_tsdf_f64(tsdf, w, h, trunc, max_w, intr, poses, depth, depth_scale=1.0):
    """FP64 form of integrate_tsdf (pixels chosen as in fp32; depth-dependent
    arithmetic in double): tsdf float64, depth float64."""
    assert tsdf.dtype == np.float64 and w.dtype == np.int32
    Z, Y, X = tsdf.shape
    intr = np.ascontiguousarray(intr, np.float32)
    poses = np.ascontiguousarray(poses, np.float32).reshape(-1, 12)
    depth = np.ascontiguousarray(depth, np.float64)
    K, H, W = depth.shape
    lib().orc_integrate_tsdf_f64(_ptr(tsdf), _ptr(w), X, Y, Z, h, trunc, max_w, _ptr(intr),
                                 _ptr(poses), _ptr(depth), depth_scale, W, H, K)
    return tsdf, w


def raycast(vol, h, intr, pose, W, H, origin=0.0, nthreads=8):
    """Depth (H,W) and camera-frame normals (H,W,3) ray-marched through a
    packed short2 volume (Z,Y,X uint32); None when K R^T is singular."""
    vol = np.ascontiguousarray(vol).view(np.uint32)
    Z, Y, X = vol.shape[:3]
    intr = np.ascontiguousarray(intr, np.float32)
    pose = np.ascontiguousarray(pose, np.float32).reshape(12)
    depth = np.zeros((H, W), np.float32)
    normals = np.zeros((H, W, 3), np.float32)
    if lib().orc_raycast(_ptr(vol), X, Y, Z, h, origin, _ptr(intr), _ptr(pose), W, H,
                         _ptr(depth), _ptr(normals), nthreads):
        return None
    return depth, normals


def raycast_setup(intr, pose):
    """IM = (K R^T)^-1 as the raycast computes it (row-major 3x3), or None."""
    im = np.zeros(9, np.float32)
    if lib().orc_raycast_setup(_ptr(np.ascontiguousarray(intr, np.float32)),
                               _ptr(np.ascontiguousarray(pose, np.float32).reshape(12)),
                               _ptr(im)):
        return None
    return im.reshape(3, 3)


def rle_encode_slice(vox):
    """Reference run-length bytes of one slice (uint32 voxels, any shape)."""
    vox = np.ascontiguousarray(vox).view(np.uint32).reshape(-1)
    out = np.zeros(6 * max(1, vox.size), np.uint8)
    n = lib().orc_rle_encode_slice(_ptr(vox), vox.size, _ptr(out))
    return out[:n].tobytes()


def rle_decode_slice(data, n):
    """Voxels (uint32, n) of one slice, or None for a malformed slice."""
    buf = np.frombuffer(bytes(data), np.uint8).copy()
    out = np.zeros(n, np.uint32)
    if lib().orc_rle_decode_slice(_ptr(buf), buf.size, _ptr(out), n):
        return None
    return out


def rle_encode_dump(vol, z0=0, nz=None):
    """Dump (per slice: uint32 byte count, records) of slices z0.. of vol (Z,Y,X)."""
    vol = np.ascontiguousarray(vol).view(np.uint32)
    Z, Y, X = vol.shape
    nz = Z - z0 if nz is None else nz
    out = np.zeros(nz * (4 + 6 * X * Y), np.uint8)
    n = lib().orc_rle_encode_dump(_ptr(vol), X, Y, z0, nz, _ptr(out))
    return out[:n].tobytes()


def rle_decode_dump(dump, X, Y, z0, nz):
    """Slices z0..z0+nz-1 (nz,Y,X uint32) of a dump, or None if malformed."""
    buf = np.frombuffer(bytes(dump), np.uint8).copy()
    out = np.zeros((nz, Y, X), np.uint32)
    if lib().orc_rle_decode_dump(_ptr(buf), buf.size, X, Y, z0, nz, _ptr(out)):
        return None
    return out
