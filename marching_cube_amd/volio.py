"""`.vol` volume files: the reference's run-length format, encoded and
decoded by libmcx's GPU codec (mcx_rle_encode / mcx_rle_decode).

Reference interface mirrored here (cu_kinect_fusion_v1.cu, compiled out in
the reference, written there for a voxel with colour; here for the live
4-byte voxel {int16 tsdf; uint16 w} of structure.h:14-18):
  cu_voxel_run_length_write_header / _read_header   :1303-1326
      -> write_header / read_header: "VOLUME_TSDF_RGB\\0", int32 size[3],
         float min_t, max_t, grid_s (40 bytes, little-endian)
  cu_voxel_run_length_encode(v_local, file, begin_local, begin_global, size)
      :1451-1531 -> run_length_encode: a missing file gets the whole volume;
         an existing one has the region [begin_global, +size) of its volume
         replaced by [begin_local, +size) of v_local (same grid_s, in bounds)
  cu_voxel_run_length_decode(v_local, file, begin_local, begin_global, size)
      :1535-1600 -> run_length_decode: v_local None (or without data) loads
         the whole file; otherwise the region is copied into v_local
Errors: a missing file for decoding and out-of-bound regions raise (the
reference prints a message and returns, :1542-1544, :1565-1579); a
malformed file raises McxError(MCX_ERR_FORMAT).
"""
import ctypes
import os
import struct

import numpy as np
import torch

from . import _capi
from .mesh import CuVoxel, _ptr, _stream_handle

VOLUME_FORMAT = b"VOLUME_TSDF_RGB\x00"
HEADER = struct.Struct("<16s3i3f")
ENCODE_SLAB_VOXELS = 1 << 24  # voxels per encode call (16 bytes of workspace each)


def _lib():
    return _capi.load()


def write_header(fp, size, min_t, max_t, grid_s):
    fp.write(HEADER.pack(VOLUME_FORMAT, *[int(a) for a in size], float(min_t), float(max_t),
                         float(grid_s)))


def read_header(fp):
    """(size (X, Y, Z), min_t, max_t, grid_s); raises on another format."""
    raw = fp.read(HEADER.size)
    if len(raw) < HEADER.size:
        raise ValueError("run_length_read_header: truncated header")
    fmt, X, Y, Z, min_t, max_t, grid_s = HEADER.unpack(raw)
    if fmt.split(b"\x00")[0] != VOLUME_FORMAT.split(b"\x00")[0]:
        raise ValueError("run_length_read_header: invalid format")
    return (X, Y, Z), min_t, max_t, grid_s


def encode(vol, z0=0, nz=None, slab_voxels=ENCODE_SLAB_VOXELS):
    """Dump bytes (per slice: uint32 byte count, records) of slices
    [z0, z0+nz) of a packed short2 device volume (Z, Y, X) int32."""
    Z, Y, X = vol.shape[:3]
    nz = Z - z0 if nz is None else nz
    lib = _lib()
    per = max(1, min(nz, slab_voxels // max(1, X * Y)))
    ws = torch.empty(max(1, lib.mcx_rle_encode_workspace_size(X, Y, per)), dtype=torch.uint8,
                     device=vol.device)
    out = torch.empty(lib.mcx_rle_encode_bound(X, Y, per), dtype=torch.uint8, device=vol.device)
    parts = []
    for z in range(z0, z0 + nz, per):
        k = min(per, z0 + nz - z)
        total = ctypes.c_int64(0)
        _capi.check(lib.mcx_rle_encode(_ptr(vol), X, Y, Z, z, k, _ptr(out), out.numel(), None,
                                       ctypes.byref(total), _ptr(ws), ws.numel(),
                                       _stream_handle(vol.device)), "mcx_rle_encode")
        parts.append(out[:total.value].cpu().numpy().tobytes())
    return b"".join(parts)


def slice_offsets(dump, nslices):
    """Byte offsets of the nslices uint32 byte counts of a dump."""
    offs = np.zeros(nslices, np.int64)
    off = 0
    n = len(dump)
    for z in range(nslices):
        if off + 4 > n:
            raise _capi.McxError(_capi.ERR_FORMAT, "run_length dump: truncated")
        offs[z] = off
        (b,) = struct.unpack_from("<I", dump, off)
        off += 4 + b
    if off > n:
        raise _capi.McxError(_capi.ERR_FORMAT, "run_length dump: truncated")
    return offs


def decode(dump, X, Y, offsets, device, out=None):
    """Slices at `offsets` (byte offsets of their counts) of a dump into an
    (len(offsets), Y, X) int32 device tensor."""
    nz = len(offsets)
    if out is None:
        out = torch.empty((nz, Y, X), dtype=torch.int32, device=device)
    if isinstance(dump, torch.Tensor):
        d = dump
    else:
        d = torch.from_numpy(np.frombuffer(dump, np.uint8).copy()).to(device)
    lib = _lib()
    ws = torch.empty(max(1, lib.mcx_rle_decode_workspace_size(d.numel(), X, Y, nz)),
                     dtype=torch.uint8, device=device)
    offs = np.ascontiguousarray(offsets, np.int64)
    _capi.check(lib.mcx_rle_decode(_ptr(d), d.numel(), offs.ctypes.data, nz, X, Y, _ptr(out),
                                   _ptr(ws), ws.numel(), _stream_handle(device)), "mcx_rle_decode")
    return out


def save(path, volume):
    """Write a whole CuVoxel to `path` (header + dump)."""
    X, Y, Z = volume.size
    dump = encode(volume.data)
    with open(path, "wb") as fp:
        write_header(fp, (X, Y, Z), volume.min_t, volume.max_t, volume.grid_s)
        fp.write(dump)


def load(path, device="cuda"):
    """Read a whole `.vol` file into a CuVoxel on `device`."""
    with open(path, "rb") as fp:
        (X, Y, Z), min_t, max_t, grid_s = read_header(fp)
        dump = fp.read()
    offs = slice_offsets(dump, Z)
    data = decode(dump, X, Y, offs, device)
    return CuVoxel(data=data, grid_s=grid_s, min_t=min_t, max_t=max_t)


def _check_region(local_size, begin_local, global_size, begin_global, size):
    for a in range(3):
        if begin_local[a] < 0 or begin_global[a] < 0 or size[a] < 0:
            return False
        if begin_local[a] + size[a] > local_size[a] or begin_global[a] + size[a] > global_size[a]:
            return False
    return True


def run_length_encode(v_local, file_name, begin_local=(0, 0, 0), begin_global=(0, 0, 0),
                      size=(0, 0, 0)):
    """cu_voxel_run_length_encode: write v_local whole when the file does
    not exist, else replace a region of the file's volume by v_local's."""
    if not os.path.exists(file_name):
        save(file_name, v_local)
        return
    with open(file_name, "rb") as fp:
        gsize, min_t, max_t, grid_s = read_header(fp)
        dump = fp.read()
    if (not _check_region(v_local.size, begin_local, gsize, begin_global, size)
            or np.float32(v_local.grid_s) != np.float32(grid_s)):
        raise ValueError("cu_voxel_run_length_encode - out of bound.")
    X, Y, Z = gsize
    offs = slice_offsets(dump, Z)
    (bx, by, bz), (gx, gy, gz), (sx, sy, sz) = begin_local, begin_global, size
    dev = v_local.data.device
    if sz > 0:
        part = decode(dump, X, Y, offs[gz:gz + sz], dev)
        part[:, gy:gy + sy, gx:gx + sx] = v_local.data[bz:bz + sz, by:by + sy, bx:bx + sx]
        new = encode(part)
        end = offs[gz + sz] if gz + sz < Z else len(dump)
        dump = dump[:offs[gz]] + new + dump[end:]
    with open(file_name, "wb") as fp:
        write_header(fp, gsize, min_t, max_t, grid_s)
        fp.write(dump)


def run_length_decode(v_local, file_name, begin_local=(0, 0, 0), begin_global=(0, 0, 0),
                      size=(0, 0, 0), device="cuda"):
    """cu_voxel_run_length_decode: with no local volume (None, or no data)
    load the whole file and return it; else copy the file's region
    [begin_global, +size) into v_local at begin_local and return v_local."""
    if not os.path.exists(file_name):
        raise FileNotFoundError(f"cu_voxel_run_length_decode - fail to open {file_name} file.")
    if v_local is None or v_local.data is None:
        return load(file_name, device)
    with open(file_name, "rb") as fp:
        gsize, _, _, grid_s = read_header(fp)
        dump = fp.read()
    if (not _check_region(v_local.size, begin_local, gsize, begin_global, size)
            or np.float32(v_local.grid_s) != np.float32(grid_s)):
        raise ValueError("cu_voxel_run_length_decode - out of bound.")
    X, Y, Z = gsize
    (bx, by, bz), (gx, gy, gz), (sx, sy, sz) = begin_local, begin_global, size
    if sz > 0:
        offs = slice_offsets(dump, Z)
        part = decode(dump, X, Y, offs[gz:gz + sz], v_local.data.device)
        v_local.data[bz:bz + sz, by:by + sy, bx:bx + sx] = part[:, gy:gy + sy, gx:gx + sx]
    return v_local
