"""Host-side mirror of the reference's mesh-extraction interface, on torch
device memory, calling libmcx through the C ABI.

Reference interface mirrored here:
  class MarchingCubes (cu_marching_cubes_v1.h:28-65)
      run(tsdf, triangles_buffer, vertexindex_to_voxel, ver_voxel)
      -> view of the first total_vertexes records of triangles_buffer
         (cu_marching_cubes_v1.cu:417-467); empty buffers are created with
         DEFAULT_TRIANGLES_BUFFER_SIZE entries (:423-438); an empty mesh
         returns an empty view (:453-456).
  cu_voxel (structure.h:20-27)            -> CuVoxel
  vertex_voxel (structure.h:29-35)        -> int32 tensor (V, 7): words 0-2
                                             float position, 3-5 vox_pos,
                                             6 = tsdf | w << 16

Volumes are torch tensors on a ROCm device: packed short2 as int32/uint32
(Z, Y, X) (low half int16 tsdf, high half uint16 weight), int16 (Z, Y, X, 2),
or float32 (Z, Y, X) for the weightless differentiable form.
"""
import ctypes
from dataclasses import dataclass

import torch

from . import _capi

VV_WORDS = 7  # vertex_voxel is 28 bytes


def _stream_handle(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def volume_kind(vol):
    """(elem, X, Y, Z) of a volume tensor; raises on an unsupported layout."""
    if not vol.is_cuda:
        raise ValueError("volume must be a device tensor")
    if not vol.is_contiguous():
        raise ValueError("volume must be contiguous ([z][y][x], x fastest)")
    if vol.dtype in (torch.int32, getattr(torch, "uint32", torch.int32)) and vol.dim() == 3:
        Z, Y, X = vol.shape
        return _capi.ELEM_SHORT2, X, Y, Z
    if vol.dtype == torch.int16 and vol.dim() == 4 and vol.shape[-1] == 2:
        Z, Y, X, _ = vol.shape
        return _capi.ELEM_SHORT2, X, Y, Z
    if vol.dtype == torch.float32 and vol.dim() == 3:
        Z, Y, X = vol.shape
        return _capi.ELEM_F32, X, Y, Z
    raise ValueError(f"unsupported volume tensor {vol.dtype} {tuple(vol.shape)}")


class Workspace:
    """Caller-owned device workspace for the look-back state (one per stream)."""

    def __init__(self, device=None, nbytes=0, grow=True):
        self.device = torch.device(device if device is not None else "cuda")
        self.buf = None
        self.grow = True
        self.ensure(nbytes)
        self.grow = grow

    def ensure(self, nbytes):
        if self.buf is not None and (self.buf.numel() >= nbytes or not self.grow):
            return self
        nbytes = max(int(nbytes), 4096)
        self.buf = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        lib = _capi.load()
        _capi.check(lib.mcx_workspace_init(_ptr(self.buf), self.buf.numel(),
                                           _stream_handle(self.device)), "workspace_init")
        return self

    @property
    def ptr(self):
        return _ptr(self.buf)

    def check(self):
        """Raise McxError if an extraction on this workspace reported a
        device-side error since the last check (synchronises the stream; for
        calls made with sync=False and for graph replays)."""
        _capi.check(_capi.load().mcx_workspace_status(self.ptr, self.nbytes,
                                                      _stream_handle(self.device)),
                    "extraction workspace")

    @property
    def nbytes(self):
        return self.buf.numel()


_default_ws = {}


def default_workspace(device):
    device = torch.device(device)
    key = (device.index, torch.cuda.current_stream(device).cuda_stream)
    ws = _default_ws.get(key)
    if ws is None:
        ws = _default_ws[key] = Workspace(device)
    return ws


@dataclass
class ExtractResult:
    count: int
    vv: torch.Tensor = None         # (count, 7) int32
    pos: torch.Tensor = None        # (count, 3) float32
    v2vox: torch.Tensor = None      # (count, 3) int32
    ver_voxel: torch.Tensor = None  # (Z, Y, X) int16


def extract_into(vol, h, *, vv=None, pos=None, v2vox=None, ver_voxel=None, cap=None,
                 d_count=None, workspace=None, sync=True):
    """Launch the fused extraction into caller buffers.  Returns the host
    vertex count when sync, else None (the count is then in d_count)."""
    elem, X, Y, Z = volume_kind(vol)
    lib = _capi.load()
    ws = workspace or default_workspace(vol.device)
    ws.ensure(lib.mcx_extract_workspace_size(X, Y, Z))
    if cap is None:
        caps = [t.shape[0] for t in (vv, pos, v2vox) if t is not None]
        cap = min(caps) if caps else 0
    hc = ctypes.c_int64(0)
    st = lib.mcx_extract(_ptr(vol), elem, X, Y, Z, float(h), _ptr(vv), _ptr(pos), _ptr(v2vox),
                         _ptr(ver_voxel), int(cap), _ptr(d_count),
                         ctypes.byref(hc) if sync else None, ws.ptr, ws.nbytes,
                         _stream_handle(vol.device))
    if st == _capi.ERR_CAPACITY:
        return int(hc.value)
    _capi.check(st, "mcx_extract")
    return int(hc.value) if sync else None


def extract(vol, h, *, want=("vv",), cap=None, workspace=None):
    """Extract the mesh of `vol`; returns ExtractResult with tensors sized to
    the vertex count.  `want` selects among vv, pos, v2vox, ver_voxel.  The
    output is allocated at `cap` (default: a size estimate) and regrown once if
    the mesh is larger."""
    elem, X, Y, Z = volume_kind(vol)
    dev = vol.device
    if cap is None:
        cap = max(4096, min(27_000_000, (X * Y * Z) // 16))
    for _ in range(2):
        bufs = {
            "vv": torch.empty((cap, VV_WORDS), dtype=torch.int32, device=dev) if "vv" in want else None,
            "pos": torch.empty((cap, 3), dtype=torch.float32, device=dev) if "pos" in want else None,
            "v2vox": torch.empty((cap, 3), dtype=torch.int32, device=dev) if "v2vox" in want else None,
            "ver_voxel": torch.empty((Z, Y, X), dtype=torch.int16, device=dev) if "ver_voxel" in want else None,
        }
        n = extract_into(vol, h, cap=cap, workspace=workspace, **bufs)
        if n <= cap:
            break
        cap = n
    r = ExtractResult(count=n)
    for k in ("vv", "pos", "v2vox"):
        if bufs[k] is not None:
            setattr(r, k, bufs[k][:n])
    r.ver_voxel = bufs["ver_voxel"]
    return r


def indexed_mesh(vol, h, workspace=None):
    """Indexed mesh of the iso-0 surface: (vertices (U, 3) float32, faces
    (F, 3) int32).  Vertices are shared exactly by grid edge (each the first
    occurrence in the canonical extraction order), faces keep the canonical
    triangle order and winding."""
    elem, X, Y, Z = volume_kind(vol)
    lib = _capi.load()
    ws = workspace or default_workspace(vol.device)
    ws.ensure(lib.mcx_extract_workspace_size(X, Y, Z))
    n = extract_into(vol, h, cap=0, workspace=ws)
    dev = vol.device
    pos = torch.empty((max(n, 1), 3), dtype=torch.float32, device=dev)
    key = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
    hc = ctypes.c_int64(0)
    _capi.check(lib.mcx_extract_ex(_ptr(vol), elem, X, Y, Z, float(h), None, _ptr(pos), None,
                                   None, _ptr(key), n, None, ctypes.byref(hc), ws.ptr,
                                   ws.nbytes, _stream_handle(dev)), "mcx_extract_ex")
    upos = torch.empty((max(n, 1), 3), dtype=torch.float32, device=dev)
    idx = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    wws = torch.empty(max(4096, lib.mcx_weld_workspace_size(n)), dtype=torch.uint8, device=dev)
    nu = ctypes.c_int64(0)
    _capi.check(lib.mcx_weld(_ptr(key), _ptr(pos), 3, n, _ptr(upos), _ptr(idx), None,
                             ctypes.byref(nu), _ptr(wws), wws.numel(), _stream_handle(dev)),
                "mcx_weld")
    return upos[: nu.value], idx[:n].view(-1, 3)


def classify(vol):
    """Per-cube (cube index, numVerts) as uint8 (Z, Y, X) tensors."""
    elem, X, Y, Z = volume_kind(vol)
    ci = torch.empty((Z, Y, X), dtype=torch.uint8, device=vol.device)
    nv = torch.empty((Z, Y, X), dtype=torch.uint8, device=vol.device)
    _capi.check(_capi.load().mcx_classify(_ptr(vol), elem, X, Y, Z, _ptr(ci), _ptr(nv),
                                          _stream_handle(vol.device)), "mcx_classify")
    return ci, nv


def normals(vol, h, positions):
    """Per-vertex normals (V, 3) (d_voxel_extract_surface_normals,
    cu_kinect_fusion_v1.cu:3146-3238); `positions` is a vertex_voxel tensor
    (V, 7) or a float32 (V, 3) tensor."""
    elem, X, Y, Z = volume_kind(vol)
    if elem != _capi.ELEM_SHORT2:
        raise ValueError("normals need a short2 volume")
    stride = positions.shape[1]
    n = positions.shape[0]
    out = torch.empty((n, 3), dtype=torch.float32, device=vol.device)
    _capi.check(_capi.load().mcx_normals(_ptr(vol), X, Y, Z, float(h), _ptr(positions), stride, n,
                                         _ptr(out), _stream_handle(vol.device)), "mcx_normals")
    return out


def vv_positions(vv):
    """float32 (V, 3) view of the positions in a vertex_voxel tensor."""
    return vv[:, :3].view(torch.float32)


def vv_fields(vv):
    """(pos float32 (V,3), vox int32 (V,3), tsdf int16 (V,), w uint16-as-int32 (V,))."""
    pos = vv[:, :3].contiguous().view(torch.float32)
    vox = vv[:, 3:6]
    tw = vv[:, 6]
    tsdf = (tw & 0xFFFF).to(torch.int32)
    tsdf = torch.where(tsdf >= 32768, tsdf - 65536, tsdf).to(torch.int16)
    w = (tw >> 16) & 0xFFFF
    return pos, vox, tsdf, w


@dataclass
class CuVoxel:
    """cu_voxel (structure.h:20-27): size, truncation, grid size, device data."""
    data: torch.Tensor
    grid_s: float
    min_t: float = 0.0
    max_t: float = 0.0

    @property
    def size(self):
        _, X, Y, Z = volume_kind(self.data)
        return (X, Y, Z)

    @property
    def size_xy(self):
        X, Y, _ = self.size
        return X * Y


class MarchingCubes:
    """Drop-in for class MarchingCubes (cu_marching_cubes_v1.h:28-65)."""
    POINTS_PER_TRIANGLE = 3
    DEFAULT_TRIANGLES_BUFFER_SIZE = 9 * 1000 * 1000 * POINTS_PER_TRIANGLE

    def __init__(self, device=None):
        self.workspace = Workspace(device)

    def run(self, tsdf, triangles_buffer=None, vertexindex_to_voxel=None, ver_voxel=None):
        """Returns (triangles view, triangles_buffer, vertexindex_to_voxel,
        ver_voxel): Python cannot rebind the caller's buffers in place, so the
        possibly created buffers are returned beside the view."""
        vol = tsdf.data
        elem, X, Y, Z = volume_kind(vol)
        dev = vol.device
        if triangles_buffer is None or triangles_buffer.numel() == 0:
            triangles_buffer = torch.empty((self.DEFAULT_TRIANGLES_BUFFER_SIZE, VV_WORDS),
                                           dtype=torch.int32, device=dev)
        if vertexindex_to_voxel is None or vertexindex_to_voxel.numel() == 0:
            vertexindex_to_voxel = torch.empty((self.DEFAULT_TRIANGLES_BUFFER_SIZE, 3),
                                               dtype=torch.int32, device=dev)
        if ver_voxel is None or ver_voxel.numel() == 0:
            ver_voxel = torch.empty((Z, Y, X), dtype=torch.int16, device=dev)
        cap = min(triangles_buffer.shape[0], vertexindex_to_voxel.shape[0])
        n = extract_into(vol, tsdf.grid_s, vv=triangles_buffer, v2vox=vertexindex_to_voxel,
                         ver_voxel=ver_voxel, cap=cap, workspace=self.workspace)
        if n > cap:
            raise _capi.McxError(_capi.ERR_CAPACITY,
                                 f"MarchingCubes.run: {n} vertices exceed buffer of {cap}")
        return triangles_buffer[:n], triangles_buffer, vertexindex_to_voxel, ver_voxel
