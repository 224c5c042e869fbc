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
  to every depth pixel, every stage one batched launch sequence, the whole
  pass (integrate -> extract -> loss -> backward) recorded once in a HIP
  graph and replayed.
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
    """B volumes (own shape and grid size; K frames of H x W each, M targets
    each, the same K, H, W and M for all): integrate -> extract -> Chamfer
    loss -> backward to dL/d(depth), each stage one launch sequence over the
    whole batch (batched integration, batched differentiable extraction,
    batched loss, batched MC backward, batched integration backward), the
    whole pass recorded once in a HIP graph and replayed.

    Static buffers: ``depth`` (B, K, H, W) float metres, ``poses`` (B, K, 12),
    ``targets`` (B, M, 3) (``depth[b]``... also as per-volume views); outputs
    ``loss`` (B,), ``grad_depth`` (B, K, H, W), ``count`` (B,), ``vbase``
    (B + 1: volume b's vertices are rows vbase[b]:vbase[b+1] of ``pos``).
    Every volume starts from (tsdf_init_value, 0): -1 makes never-observed
    voxels (object interiors behind the truncation band) count as inside, so
    they add no surface.  The fp32 TSDF of all volumes is one concatenated
    array (volume b after volumes 0..b-1); the extraction workspace is sized
    by an eager run; ``cap`` bounds the vertices of the whole batch.
    """

    def __init__(self, shapes, hs, depth, poses, targets, intr, trunc, max_w=128, cap=None,
                 tsdf_init_value=-1.0):
        from .batch import BatchPlan
        lib = _capi.load()
        self.B = len(shapes)
        self.dev = depth[0].device
        self.shapes = [tuple(int(a) for a in s) for s in shapes]
        self.hs = [float(h) for h in hs]
        if len({tuple(d.shape) for d in depth}) != 1 or len({tuple(t.shape) for t in targets}) != 1:
            raise ValueError("every volume needs the same K x H x W frames and M targets")
        self.K, self.H, self.W = (int(a) for a in depth[0].shape)
        self.M = int(targets[0].shape[0])
        self.depth_all = torch.stack([d.to(torch.float32) for d in depth]).contiguous()
        self.poses_all = torch.stack([p.to(torch.float32).reshape(self.K, 12)
                                      for p in poses]).contiguous()
        self.targets_all = torch.stack([t.to(torch.float32) for t in targets]).contiguous()
        self.depth = list(self.depth_all.unbind(0))
        self.poses = list(self.poses_all.unbind(0))
        self.targets = list(self.targets_all.unbind(0))
        self.intr = _intr_arr(intr)
        self.trunc, self.max_w = float(trunc), int(max_w)
        self.init_value = float(tsdf_init_value)
        nvox = [Z * Y * X for Z, Y, X in self.shapes]
        self.vox_begin = np.concatenate([[0], np.cumsum(nvox)]).astype(np.int64)
        total = int(self.vox_begin[-1])
        self.tsdf_all = torch.empty(total, dtype=torch.float32, device=self.dev)
        self.w_all = torch.empty(total, dtype=torch.int32, device=self.dev)
        self.gtsdf_all = torch.empty(total, dtype=torch.float32, device=self.dev)
        self.tsdf = [self.tsdf_all[self.vox_begin[b]:self.vox_begin[b + 1]].view(self.shapes[b])
                     for b in range(self.B)]
        self.w = [self.w_all[self.vox_begin[b]:self.vox_begin[b + 1]].view(self.shapes[b])
                  for b in range(self.B)]
        self.gtsdf = [self.gtsdf_all[self.vox_begin[b]:self.vox_begin[b + 1]].view(self.shapes[b])
                      for b in range(self.B)]
        self.plan = BatchPlan(self.tsdf, self.hs)
        if lib.mcx_batch_voxels(self.plan.handle) != total:
            raise RuntimeError("batch plan voxel count does not match the concatenated TSDF")
        self.grad_depth_all = torch.empty_like(self.depth_all)
        self.grad_depth = list(self.grad_depth_all.unbind(0))
        self.vbase = torch.zeros(self.B + 1, dtype=torch.int64, device=self.dev)
        self.total = torch.zeros(1, dtype=torch.int64, device=self.dev)
        self.loss = torch.zeros(self.B, dtype=torch.float32, device=self.dev)
        self.maxabs = torch.zeros(1, dtype=torch.int32, device=self.dev)
        self.ws_f = torch.empty(max(4096, lib.mcx_integrate_workspace_size(self.W, self.H)),
                                dtype=torch.uint8, device=self.dev)
        self.ws_b = torch.empty(max(4096, lib.mcx_integrate_tsdf_backward_batched_workspace_size(
            self.plan.handle, self.W, self.H, self.K)), dtype=torch.uint8, device=self.dev)
        # extraction workspace (every record kept: the backward needs them) and
        # output capacity from an eager run
        s = _stream_handle(self.dev)
        self._integrate()
        records = max(65536, sum((X - 1) * (Y - 1) * (Z - 1) for Z, Y, X in self.shapes) // 8)
        for _ in range(4):
            ws_bytes = max(4096, lib.mcx_batch_workspace_size(self.plan.handle, records))
            self.ws = torch.empty(ws_bytes, dtype=torch.uint8, device=self.dev)
            _capi.check(lib.mcx_workspace_init(_ptr(self.ws), self.ws.numel(), s),
                        "workspace_init")
            hc = ctypes.c_int64(0)
            st = lib.mcx_extract_diff_batched(self.plan.handle, None, 0, _ptr(self.vbase), None,
                                              ctypes.byref(hc), _ptr(self.ws), self.ws.numel(), s)
            if st == _capi.ERR_WORKSPACE:
                records *= 2
                continue
            if st not in (_capi.OK, _capi.ERR_CAPACITY):
                _capi.check(st, "mcx_extract_diff_batched")
            break
        else:
            raise RuntimeError("FusionGraph: record scratch still too small")
        self.cap = int(cap or max(4096, int(hc.value * 1.25)))
        self.pos = torch.zeros((self.cap, 3), dtype=torch.float32, device=self.dev)
        self.gpos = torch.zeros((self.cap, 3), dtype=torch.float32, device=self.dev)
        self.ws_l = torch.empty(max(4096, lib.mcx_chamfer_workspace_size(self.cap)),
                                dtype=torch.uint8, device=self.dev)
        sb = lib.mcx_batch_backward_scratch_size(self.plan.handle, self.ws.numel())
        self.scratch = torch.empty(max(sb, 16), dtype=torch.uint8, device=self.dev)
        self.graph = None

    @property
    def count(self):
        """Vertices per volume (B,), from the vertex ranges of the last pass."""
        return self.vbase[1:] - self.vbase[:-1]

    def _integrate(self):
        _capi.check(_capi.load().mcx_integrate_tsdf_batched(
            self.plan.handle, _ptr(self.tsdf_all), _ptr(self.w_all), 1, self.init_value,
            self.trunc, self.max_w, self.intr, _ptr(self.poses_all), _ptr(self.depth_all), 1.0,
            self.W, self.H, self.K, _ptr(self.ws_f), self.ws_f.numel(),
            _stream_handle(self.dev)), "mcx_integrate_tsdf_batched")

    def step(self):
        """One pass over every volume (eager, or recorded when capturing).
        Launch-only: no allocation, no host synchronisation."""
        lib = _capi.load()
        s = _stream_handle(self.dev)
        h = self.plan.handle
        self._integrate()
        _capi.check(lib.mcx_extract_diff_batched(h, _ptr(self.pos), self.cap, _ptr(self.vbase),
                                                 _ptr(self.total), None, _ptr(self.ws),
                                                 self.ws.numel(), s), "mcx_extract_diff_batched")
        _capi.check(lib.mcx_chamfer_batched(_ptr(self.pos), _ptr(self.vbase), self.B,
                                            _ptr(self.targets_all), self.M, self.cap,
                                            _ptr(self.loss), _ptr(self.gpos), _ptr(self.ws_l),
                                            self.ws_l.numel(), s), "mcx_chamfer_batched")
        _capi.check(lib.mcx_extract_backward_batched(h, _ptr(self.gpos), _ptr(self.gtsdf_all),
                                                     _ptr(self.maxabs), _ptr(self.ws),
                                                     self.ws.numel(), _ptr(self.scratch),
                                                     self.scratch.numel(), s),
                    "mcx_extract_backward_batched")
        _capi.check(lib.mcx_integrate_tsdf_backward_batched(
            h, _ptr(self.gtsdf_all), None, _ptr(self.maxabs), self.trunc, self.max_w, self.intr,
            _ptr(self.poses_all), _ptr(self.depth_all), 1.0, self.W, self.H, self.K,
            _ptr(self.grad_depth_all), _ptr(self.ws_b), self.ws_b.numel(), s),
            "mcx_integrate_tsdf_backward_batched")

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
        error since the last check (synchronises the stream), or if the vertices
        exceeded the capacity."""
        _capi.check(_capi.load().mcx_workspace_status(_ptr(self.ws), self.ws.numel(),
                                                      _stream_handle(self.dev)),
                    "fusion graph extraction workspace")
        if int(self.total.item()) > self.cap:
            raise _capi.McxError(_capi.ERR_CAPACITY, f"FusionGraph: {int(self.total.item())} "
                                                     f"vertices exceed the capacity {self.cap}")
