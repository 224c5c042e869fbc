"""Differentiable TSDF fusion: depth frames -> fp32 TSDF -> mesh, and back.

No counterpart in the reference (its integration kernels are dead code and
fixed-point).  The forward is the projective observation of
cu_kinect_fusion_v1.cu:427-532 (ray-distance sdf, nearest pixel, truncation
window; see mcx_integrate.hip), stored unpacked in fp32 with int32 weights:

    f <- fma(f, w, t) / (w + 1),   w <- min(w + 1, max_w),   t = clamp(sdf / trunc, -1, 1)

* ``IntegrateTsdfFn`` -- torch.autograd.Function depth (K, H, W) -> tsdf
  (Z, Y, X); its backward is mcx_integrate_tsdf_backward (deterministic:
  int64 fixed-point sums).
* ``FusionGraph`` -- a batch of volumes, each fused from its own K frames,
  extracted, scored with a Chamfer-style vertex loss, and differentiated back
  to every depth pixel, with the whole sequence (integrate -> extract -> loss
  -> backward) recorded once in a HIP graph and replayed.
"""
import ctypes

import numpy as np
import torch

from . import _capi
from .mesh import _ptr, _stream_handle


def _intr_arr(intr):
    return (ctypes.c_float * 4)(*[float(v) for v in intr])


def tsdf_init(shape, device, value=1.0):
    """(tsdf, weight) of a volume (Z, Y, X): every voxel (value, 0)."""
    tsdf = torch.empty(shape, dtype=torch.float32, device=device)
    w = torch.empty(shape, dtype=torch.int32, device=device)
    _capi.check(_capi.load().mcx_tsdf_init(_ptr(tsdf), _ptr(w), tsdf.numel(), float(value),
                                           _stream_handle(tsdf.device)), "mcx_tsdf_init")
    return tsdf, w


class _Integrator:
    """Launch helper with its own workspaces (ray factors, accumulators)."""

    def __init__(self, shape, W, H, K, h, trunc, max_w, intr, device):
        self.Z, self.Y, self.X = (int(a) for a in shape)
        self.W, self.H, self.K = int(W), int(H), int(K)
        self.h, self.trunc, self.max_w = float(h), float(trunc), int(max_w)
        self.intr = _intr_arr(intr)
        lib = _capi.load()
        self.ws_f = torch.empty(max(4096, lib.mcx_integrate_workspace_size(W, H)),
                                dtype=torch.uint8, device=device)
        self.ws_b = torch.empty(max(4096, lib.mcx_integrate_tsdf_backward_workspace_size(W, H, K)),
                                dtype=torch.uint8, device=device)

    def forward(self, tsdf, w, poses, depth, depth_scale=1.0):
        _capi.check(_capi.load().mcx_integrate_tsdf(
            _ptr(tsdf), _ptr(w), self.X, self.Y, self.Z, self.h, self.trunc, self.max_w,
            self.intr, _ptr(poses), _ptr(depth), float(depth_scale), self.W, self.H, self.K,
            _ptr(self.ws_f), self.ws_f.numel(), _stream_handle(tsdf.device)), "mcx_integrate_tsdf")

    def backward(self, grad_tsdf, w0, poses, depth, grad_depth, depth_scale=1.0):
        _capi.check(_capi.load().mcx_integrate_tsdf_backward(
            _ptr(grad_tsdf), _ptr(w0), self.X, self.Y, self.Z, self.h, self.trunc, self.max_w,
            self.intr, _ptr(poses), _ptr(depth), float(depth_scale), self.W, self.H, self.K,
            _ptr(grad_depth), _ptr(self.ws_b), self.ws_b.numel(), _stream_handle(grad_tsdf.device)),
            "mcx_integrate_tsdf_backward")


class IntegrateTsdfFn(torch.autograd.Function):
    """tsdf = integrate(tsdf0, w0, depth); differentiable w.r.t. depth."""

    @staticmethod
    def forward(ctx, depth, tsdf0, w0, poses, h, trunc, max_w, intr):
        depth = depth.contiguous()
        poses = poses.contiguous().to(torch.float32)
        K, H, W = depth.shape
        integ = _Integrator(tsdf0.shape, W, H, K, h, trunc, max_w, intr, depth.device)
        tsdf = tsdf0.clone()
        w = w0.clone()
        integ.forward(tsdf, w, poses, depth)
        ctx.integ = integ
        ctx.save_for_backward(depth, w0, poses)
        ctx.mark_non_differentiable(w)
        return tsdf, w

    @staticmethod
    def backward(ctx, grad_tsdf, grad_w):
        depth, w0, poses = ctx.saved_tensors
        gd = torch.empty_like(depth)
        ctx.integ.backward(grad_tsdf.contiguous(), w0, poses, depth, gd)
        return gd, None, None, None, None, None, None, None


def integrate_tsdf(depth, tsdf0, w0, poses, h, trunc, max_w, intr):
    """(tsdf, w) after fusing depth frames (K, H, W, metres) taken from poses
    (K, 12: camera-to-world R row-major, t) into (tsdf0, w0); tsdf is
    differentiable with respect to depth."""
    return IntegrateTsdfFn.apply(depth, tsdf0, w0, poses, h, trunc, max_w, intr)


def _nearest_sq(pos, targets):
    """(squared distance, index) of the nearest target of every row of pos,
    from explicit differences (no |a|^2 + |b|^2 - 2ab cancellation)."""
    d2 = (pos.unsqueeze(1) - targets.unsqueeze(0)).pow(2).sum(dim=2)
    return d2.min(dim=1)


def chamfer_loss(pos, targets, count=None):
    """sum over vertices of the squared distance to the nearest target point;
    with a device `count`, only the first `count` rows of pos are used (no
    host synchronisation: capturable).  Rows are summed in blocks of 1024
    (per-row reductions), so no reduction needs a cross-block buffer."""
    m = _nearest_sq(pos, targets).values
    if count is not None:
        m = m * (torch.arange(pos.shape[0], device=pos.device) < count)
    n = m.shape[0]
    pad = (-n) % 1024
    if pad:
        m = torch.cat([m, m.new_zeros(pad)])
    return m.view(-1, 1024).sum(dim=1).sum()


def chamfer_grad(pos, targets, count):
    """d chamfer_loss / d pos (rows past count are 0)."""
    j = _nearest_sq(pos, targets).indices
    g = 2.0 * (pos - targets[j])
    mask = (torch.arange(pos.shape[0], device=pos.device) < count).unsqueeze(1)
    return (g * mask).contiguous()


class FusionGraph:
    """B volumes (own shape, grid size, K frames of H x W each, own targets):
    integrate -> extract -> Chamfer loss -> backward to dL/d(depth), recorded
    in one HIP graph.

    Static buffers: ``depth[b]`` (K, H, W) float metres, ``poses[b]`` (K, 12),
    ``targets[b]`` (M, 3); outputs ``loss`` (B,), ``grad_depth[b]``,
    ``count[b]``.  Every volume starts from (tsdf_init_value, 0): -1 makes
    never-observed voxels (object interiors behind the truncation band) count
    as inside, so they add no surface.  The
    extraction workspace is shared (the backward recomputes the forward's
    records), sized by an eager run; ``cap`` bounds the vertices per volume.
    """

    def __init__(self, shapes, hs, depth, poses, targets, intr, trunc, max_w=128, cap=None,
                 tsdf_init_value=-1.0):
        lib = _capi.load()
        self.B = len(shapes)
        self.dev = depth[0].device
        self.shapes = [tuple(int(a) for a in s) for s in shapes]
        self.hs = [float(h) for h in hs]
        self.depth = [d.contiguous() for d in depth]
        self.poses = [p.contiguous().to(torch.float32) for p in poses]
        self.targets = [t.contiguous().to(torch.float32) for t in targets]
        self.init_value = float(tsdf_init_value)
        self.integ = [_Integrator(s, d.shape[2], d.shape[1], d.shape[0], h, trunc, max_w, intr,
                                  self.dev) for s, h, d in zip(self.shapes, self.hs, self.depth)]
        self.tsdf = [torch.empty(s, dtype=torch.float32, device=self.dev) for s in self.shapes]
        self.w = [torch.empty(s, dtype=torch.int32, device=self.dev) for s in self.shapes]
        self.w0 = [torch.zeros(s, dtype=torch.int32, device=self.dev) for s in self.shapes]
        self.gtsdf = [torch.empty(s, dtype=torch.float32, device=self.dev) for s in self.shapes]
        self.grad_depth = [torch.empty_like(d) for d in self.depth]
        self.count = torch.zeros(self.B, dtype=torch.int64, device=self.dev)
        self.loss = torch.zeros(self.B, dtype=torch.float32, device=self.dev)
        # extraction workspace and output capacity from an eager run
        cubes = max((X - 1) * (Y - 1) * (Z - 1) for Z, Y, X in self.shapes)
        records = max(65536, cubes // 8)
        need_cap = 0
        for b in range(self.B):
            self._integrate(b)
            Z, Y, X = self.shapes[b]
            for _ in range(3):
                ws_bytes = lib.mcx_extract_workspace_size_ex(X, Y, Z, records)
                ws = torch.empty(max(ws_bytes, 4096), dtype=torch.uint8, device=self.dev)
                _capi.check(lib.mcx_workspace_init(_ptr(ws), ws.numel(), _stream_handle(self.dev)),
                            "workspace_init")
                hc = ctypes.c_int64(0)
                st = lib.mcx_extract_diff(_ptr(self.tsdf[b]), X, Y, Z, self.hs[b], None, 0, None,
                                          ctypes.byref(hc), _ptr(ws), ws.numel(),
                                          _stream_handle(self.dev))
                if st == _capi.ERR_WORKSPACE:
                    records = 2 * records
                    continue
                if st not in (_capi.OK, _capi.ERR_CAPACITY):
                    _capi.check(st, "mcx_extract_diff")
                break
            need_cap = max(need_cap, int(hc.value))
        self.cap = int(cap or max(4096, int(need_cap * 1.25)))
        self.cap = (self.cap + 1023) // 1024 * 1024  # whole blocks of the loss sum
        self.ws_bytes = max(lib.mcx_extract_workspace_size_ex(X, Y, Z, records)
                            for Z, Y, X in self.shapes)
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=self.dev)
        _capi.check(lib.mcx_workspace_init(_ptr(self.ws), self.ws.numel(),
                                           _stream_handle(self.dev)), "workspace_init")
        self.pos = [torch.zeros((self.cap, 3), dtype=torch.float32, device=self.dev)
                    for _ in range(self.B)]
        self.gpos = [torch.zeros((self.cap, 3), dtype=torch.float32, device=self.dev)
                     for _ in range(self.B)]
        sb = max(lib.mcx_backward_scratch_size(X, Y, Z, self.ws_bytes) for Z, Y, X in self.shapes)
        self.scratch = torch.empty(max(sb, 16), dtype=torch.uint8, device=self.dev)
        self.graph = None

    def _integrate(self, b):
        _capi.check(_capi.load().mcx_tsdf_init(_ptr(self.tsdf[b]), _ptr(self.w[b]),
                                               self.tsdf[b].numel(), self.init_value,
                                               _stream_handle(self.dev)), "mcx_tsdf_init")
        self.integ[b].forward(self.tsdf[b], self.w[b], self.poses[b], self.depth[b])

    def step(self):
        """One pass over every volume (eager, or recorded when capturing).
        Launch-only: no allocation, no host synchronisation."""
        lib = _capi.load()
        s = _stream_handle(self.dev)
        for b in range(self.B):
            Z, Y, X = self.shapes[b]
            self._integrate(b)
            _capi.check(lib.mcx_extract_diff(_ptr(self.tsdf[b]), X, Y, Z, self.hs[b],
                                             _ptr(self.pos[b]), self.cap,
                                             _ptr(self.count[b:b + 1]), None, _ptr(self.ws),
                                             self.ws.numel(), s), "mcx_extract_diff")
            n = self.count[b]
            self.loss[b:b + 1].copy_(chamfer_loss(self.pos[b], self.targets[b], n).reshape(1))
            self.gpos[b].copy_(chamfer_grad(self.pos[b], self.targets[b], n))
            _capi.check(lib.mcx_extract_backward_recompute(
                _ptr(self.tsdf[b]), X, Y, Z, self.hs[b], _ptr(self.gpos[b]), _ptr(self.gtsdf[b]),
                _ptr(self.ws), self.ws.numel(), _ptr(self.scratch), self.scratch.numel(), s),
                "mcx_extract_backward_recompute")
            self.integ[b].backward(self.gtsdf[b], self.w0[b], self.poses[b], self.depth[b],
                                   self.grad_depth[b])

    def capture(self):
        """Record step() in a HIP graph (after one eager warm-up pass)."""
        s = torch.cuda.Stream(self.dev)
        s.wait_stream(torch.cuda.current_stream(self.dev))
        with torch.cuda.stream(s):
            self.step()
        torch.cuda.current_stream(self.dev).wait_stream(s)
        torch.cuda.synchronize(self.dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.step()
        torch.cuda.synchronize(self.dev)
        return self.graph

    def replay(self):
        self.graph.replay()

    def check(self):
        """Raise McxError if an extraction of the pass reported a device-side
        error since the last check (synchronises the stream)."""
        _capi.check(_capi.load().mcx_workspace_status(_ptr(self.ws), self.ws.numel(),
                                                      _stream_handle(self.dev)),
                    "fusion graph extraction workspace")
