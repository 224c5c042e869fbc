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


class _Forward:
    """Workspace + outputs of one forward call (kept for its backward)."""

    def __init__(self, vol, h, cap=None):
        lib = _capi.load()
        Z, Y, X = vol.shape
        self.X, self.Y, self.Z, self.h = X, Y, Z, float(h)
        cubes = max(0, (X - 1) * (Y - 1) * (Z - 1))
        records = max(65536, cubes // 16)
        cap = cap or max(4096, min(60_000_000, cubes // 8))
        for _ in range(4):
            nbytes = lib.mcx_extract_workspace_size_ex(X, Y, Z, records)
            ws = torch.zeros(max(nbytes, 4096), dtype=torch.uint8, device=vol.device)
            pos = torch.empty((cap, 3), dtype=torch.float32, device=vol.device)
            hc = ctypes.c_int64(0)
            st = lib.mcx_extract_diff(_ptr(vol), X, Y, Z, self.h, _ptr(pos), cap, None,
                                      ctypes.byref(hc), _ptr(ws), ws.numel(),
                                      _stream_handle(vol.device))
            if st == _capi.ERR_WORKSPACE:
                records = cubes  # every cube may be recorded: no column overflows
                continue
            if st == _capi.ERR_CAPACITY:
                cap = int(hc.value)
                continue
            _capi.check(st, "mcx_extract_diff")
            break
        else:
            raise _capi.McxError(_capi.ERR_INTERNAL, "mcx_extract_diff did not converge")
        self.ws = ws
        self.count = int(hc.value)
        self.pos = pos[: self.count]

    def backward(self, vol, grad_pos):
        lib = _capi.load()
        grad_pos = grad_pos.contiguous().to(torch.float32)
        assert grad_pos.shape == (self.count, 3)
        gvol = torch.empty_like(vol)
        sb = lib.mcx_backward_scratch_size(self.X, self.Y, self.Z, self.ws.numel())
        scratch = torch.empty(max(sb, 16), dtype=torch.uint8, device=vol.device)
        _capi.check(lib.mcx_extract_backward(_ptr(vol), self.X, self.Y, self.Z, self.h,
                                             _ptr(grad_pos), _ptr(gvol), _ptr(self.ws),
                                             self.ws.numel(), _ptr(scratch), scratch.numel(),
                                             _stream_handle(vol.device)),
                    "mcx_extract_backward")
        return gvol


class ExtractFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vol, h):
        if vol.dtype != torch.float32 or vol.dim() != 3 or not vol.is_cuda:
            raise ValueError("differentiable extraction takes a float32 (Z, Y, X) device tensor")
        vol = vol.contiguous()
        fw = _Forward(vol, h)
        ctx.fw = fw
        ctx.save_for_backward(vol)
        return fw.pos

    @staticmethod
    def backward(ctx, grad_pos):
        (vol,) = ctx.saved_tensors
        return ctx.fw.backward(vol, grad_pos), None


def extract_positions(vol, h):
    """Vertex positions (V, 3) of the iso-0 surface of fp32 field `vol`,
    differentiable with respect to `vol`."""
    return ExtractFn.apply(vol, h)
