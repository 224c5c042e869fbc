"""Batched extraction of many uneven volumes (no counterpart in the
reference): one plan (device descriptors, built once), one launch sequence
per extraction, one concatenated output with per-volume vertex ranges."""
import ctypes

import numpy as np
import torch

from . import _capi
from .mesh import VV_WORDS, _ptr, _stream_handle, volume_kind


class BatchPlan:
    def __init__(self, vols, hs):
        if len(vols) == 0:
            raise ValueError("empty batch")
        kinds = {volume_kind(v)[0] for v in vols}
        if len(kinds) != 1:
            raise ValueError("all volumes of a batch must have the same voxel kind")
        self.elem = kinds.pop()
        self.vols = list(vols)  # keep the device memory alive
        self.device = self.vols[0].device
        dims = np.array([volume_kind(v)[1:] for v in vols], dtype=np.int32)  # X, Y, Z
        self.dims = dims
        h = np.asarray(hs, dtype=np.float32)
        ptrs = (ctypes.c_void_p * len(vols))(*[v.data_ptr() for v in vols])
        lib = _capi.load()
        handle = ctypes.c_void_p()
        _capi.check(lib.mcx_batch_create(ctypes.byref(handle), len(vols),
                                         dims.ctypes.data, h.ctypes.data, ptrs, self.elem),
                    "mcx_batch_create")
        self.handle = handle
        self.n_tiles = lib.mcx_batch_tiles(handle)
        self.ws = None

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _capi is not None:  # modules may be gone at exit
            _capi.load().mcx_batch_destroy(h)
            self.handle = None

    def workspace(self, records=-1):
        lib = _capi.load()
        nbytes = max(4096, lib.mcx_batch_workspace_size(self.handle, records))
        if self.ws is None or self.ws.numel() < nbytes:
            self.ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            _capi.check(lib.mcx_workspace_init(_ptr(self.ws), self.ws.numel(),
                                               _stream_handle(self.device)), "workspace_init")
        return self.ws

    def check(self):
        """Raise McxError if a batched extraction reported a device-side error
        since the last check (synchronises the stream)."""
        if self.ws is not None:
            _capi.check(_capi.load().mcx_workspace_status(_ptr(self.ws), self.ws.numel(),
                                                          _stream_handle(self.device)),
                        "batched extraction workspace")

    def extract_into(self, *, vv=None, pos=None, v2vox=None, vbase=None, cap=None, d_count=None,
                     sync=True):
        lib = _capi.load()
        ws = self.workspace()
        if vbase is None:
            vbase = torch.empty(len(self.vols) + 1, dtype=torch.int64, device=self.device)
        if cap is None:
            caps = [t.shape[0] for t in (vv, pos, v2vox) if t is not None]
            cap = min(caps) if caps else 0
        hc = ctypes.c_int64(0)
        st = lib.mcx_extract_batched(self.handle, _ptr(vv), _ptr(pos), _ptr(v2vox), int(cap),
                                     _ptr(vbase), _ptr(d_count), ctypes.byref(hc) if sync else None,
                                     _ptr(ws), ws.numel(), _stream_handle(self.device))
        if st != _capi.ERR_CAPACITY:
            _capi.check(st, "mcx_extract_batched")
        return (int(hc.value) if sync else None), vbase

    def extract(self, want=("vv",), cap=None):
        dev = self.device
        if cap is None:
            cap = max(4096, int(sum(int(np.prod(d)) for d in self.dims) // 16))
        for _ in range(2):
            bufs = {
                "vv": torch.empty((cap, VV_WORDS), dtype=torch.int32, device=dev) if "vv" in want else None,
                "pos": torch.empty((cap, 3), dtype=torch.float32, device=dev) if "pos" in want else None,
                "v2vox": torch.empty((cap, 3), dtype=torch.int32, device=dev) if "v2vox" in want else None,
            }
            n, vbase = self.extract_into(cap=cap, **bufs)
            if n <= cap:
                break
            cap = n
        return n, vbase, {k: (v[:n] if v is not None else None) for k, v in bufs.items()}
