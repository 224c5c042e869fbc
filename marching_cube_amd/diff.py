"""Differentiable marching-cubes extraction behind torch.autograd.Function.

No counterpart in the reference.  Forward: mcx_extract_diff on an fp32 field
(Z, Y, X) -> vertex positions (V, 3), identical to the plain extraction's
positions (reference formulas, canonical order).  Backward: the hand-written
mcx_extract_backward kernels push dL/dv back to dL/df through each edge's
interpolation t = f_a / (f_a - f_b), with fixed-order sums (no atomics), so
gradients are bitwise reproducible.  The compacted per-cube records of the
forward are kept in its workspace and reused; the volume is not re-read.
"""
import ctypes

import torch

from . import _capi
from .mesh import _ptr, _stream_handle


class DiffExtractor:
    """Reusable device buffers (workspace, positions, backward scratch) for
    differentiable extractions of fp32 volumes of one shape.  The workspace
    holds the last forward's per-cube records until the next forward, so a
    backward must follow its own forward on the same extractor."""

    def __init__(self, shape, device, records=None, cap=None, recompute=False):
        Z, Y, X = (int(a) for a in shape)
        self.X, self.Y, self.Z = X, Y, Z
        self.device = torch.device(device)
        cubes = max(0, (X - 1) * (Y - 1) * (Z - 1))
        self.cubes = cubes
        self.records = records or max(65536, cubes // 16)
        self.cap = cap or max(4096, min(60_000_000, cubes // 8))
        self.ws = None
        self.pos = None
        self.scratch = None
        self.count = 0
        # recompute: the backward rebuilds the forward's per-cube records from
        # the volume (count + scan kernels) instead of reading them from the
        # workspace, so the workspace may serve other extractions in between
        self.recompute = recompute
        self.generation = 0  # forwards run so far (a backward must match the last)
        self._alloc()

    def _alloc(self):
        lib = _capi.load()
        nbytes = max(4096, lib.mcx_extract_workspace_size_ex(self.X, self.Y, self.Z, self.records))
        if self.ws is None or self.ws.numel() < nbytes:
            self.ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            _capi.check(lib.mcx_workspace_init(_ptr(self.ws), self.ws.numel(),
                                               _stream_handle(self.device)), "workspace_init")
            self.scratch = None
        if self.pos is None or self.pos.shape[0] < self.cap:
            self.pos = torch.empty((self.cap, 3), dtype=torch.float32, device=self.device)
        if self.scratch is None:
            sb = lib.mcx_backward_scratch_size(self.X, self.Y, self.Z, self.ws.numel())
            self.scratch = torch.empty(max(sb, 16), dtype=torch.uint8, device=self.device)

    def forward(self, vol, h):
        """Vertex positions (count, 3), a view of the extractor's buffer."""
        lib = _capi.load()
        assert tuple(vol.shape) == (self.Z, self.Y, self.X) and vol.dtype == torch.float32
        self.h = float(h)
        for _ in range(4):
            hc = ctypes.c_int64(0)
            st = lib.mcx_extract_diff(_ptr(vol), self.X, self.Y, self.Z, self.h, _ptr(self.pos),
                                      self.pos.shape[0], None, ctypes.byref(hc), _ptr(self.ws),
                                      self.ws.numel(), _stream_handle(self.device))
            if st == _capi.ERR_WORKSPACE:
                self.records = self.cubes  # every cube may be recorded: no column overflows
                self._alloc()
                continue
            if st == _capi.ERR_CAPACITY:
                self.cap = int(hc.value)
                self._alloc()
                continue
            _capi.check(st, "mcx_extract_diff")
            self.count = int(hc.value)
            self.generation += 1
            return self.pos[: self.count]
        raise _capi.McxError(_capi.ERR_INTERNAL, "mcx_extract_diff did not converge")

    def backward(self, vol, grad_pos, out=None):
        """dL/dvol from dL/dpos of the preceding forward (same volume)."""
        lib = _capi.load()
        grad_pos = grad_pos.contiguous().to(torch.float32)
        assert grad_pos.shape == (self.count, 3)
        gvol = torch.empty_like(vol) if out is None else out
        fn = lib.mcx_extract_backward_recompute if self.recompute else lib.mcx_extract_backward
        _capi.check(fn(_ptr(vol), self.X, self.Y, self.Z, self.h, _ptr(grad_pos), _ptr(gvol),
                       _ptr(self.ws), self.ws.numel(), _ptr(self.scratch), self.scratch.numel(),
                       _stream_handle(self.device)), "mcx_extract_backward")
        return gvol


class ExtractFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vol, h, extractor):
        if vol.dtype != torch.float32 or vol.dim() != 3 or not vol.is_cuda:
            raise ValueError("differentiable extraction takes a float32 (Z, Y, X) device tensor")
        vol = vol.contiguous()
        ex = extractor or DiffExtractor(vol.shape, vol.device)
        pos = ex.forward(vol, h)
        ctx.ex = ex
        ctx.generation = ex.generation
        ctx.h = float(h)
        ctx.save_for_backward(vol)
        # the extractor's buffer is reused by its next forward: hand out a copy
        # only when the caller shares an extractor across calls
        return pos.clone() if extractor is not None else pos

    @staticmethod
    def backward(ctx, grad_pos):
        (vol,) = ctx.saved_tensors
        if ctx.ex.recompute:
            ctx.ex.h = ctx.h
            ctx.ex.count = grad_pos.shape[0]
        elif ctx.ex.generation != ctx.generation:
            raise RuntimeError("the DiffExtractor ran another forward before this backward; "
                               "its per-cube records are gone")
        return ctx.ex.backward(vol, grad_pos), None, None


def extract_positions(vol, h, extractor=None):
    """Vertex positions (V, 3) of the iso-0 surface of fp32 field `vol`,
    differentiable with respect to `vol`.  `extractor` (a DiffExtractor of
    the volume's shape) reuses device buffers across calls; its backward must
    run before its next forward unless it was made with recompute=True (then
    the backward rebuilds the per-cube records from the volume and any number
    of forwards may run in between)."""
    return ExtractFn.apply(vol, h, extractor)
