"""GPU parity of volume init, projective integration and the synthetic depth
frames against the oracle: integration is bit-exact (same fp32 sequence),
the depth synthesis matches an FP64 numpy ray cast."""
import numpy as np
import pytest
import torch

from marching_cube_amd import fusion, mesh, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def small_scene(dev, K=8, W=96, H=72, dim=40, seed=5):
    intr = [80.0, 80.0, 47.5, 35.5]
    extent = 1.0
    objs = synth.kinfu_scene(6, extent, seed)
    objs[:, 4:7] *= 2.0  # objects of 2-8 voxels radius at 40^3
    poses = synth.orbit_poses(K, extent, radius=1.2, seed=seed)
    depth = fusion.synth_depth(intr, poses, objs, W, H, sigma=0.001, dropout=0.02, seed=seed,
                               device=dev)
    return intr, objs, poses, depth, extent / dim, dim


def test_volume_init(dev):
    print("[case] volume_init 5x6x7", flush=True)
    v = torch.zeros((5, 6, 7), dtype=torch.int32, device=dev)
    fusion.voxel_init(v)
    ref = O.volume_init((5, 6, 7))
    assert v.cpu().numpy().view(np.uint32).tobytes() == ref.tobytes()


def test_integrate_bit_exact(dev):
    print("[case] integrate 40^3 x 8 frames", flush=True)
    intr, objs, poses, depth, h, dim = small_scene(dev)
    vox = fusion.voxel_new((dim, dim, dim), 0.04, h, device=dev)
    fusion.integrate(vox, intr, poses, depth, trunc=0.04, max_w=5)
    got = vox.data.cpu().numpy().view(np.uint32)
    ref = O.volume_init((dim, dim, dim))
    O.integrate(ref, h, 0.04, 5, np.array(intr, np.float32), poses, depth.cpu().numpy())
    t, w = synth.unpack_short2(ref)
    assert (w > 0).mean() > 0.05  # the scene was actually observed
    assert (w == 5).any()         # and the weight cap was reached
    assert got.tobytes() == ref.tobytes()


def test_integrate_split_truncation_float_frames_per_frame_calls(dev):
    # min_t 0.08 behind the surface, max_t 0.03 in front (the volume's own
    # fields); uint16 frames in one sweep, and float frames in metres one call
    # per frame (the reference's call pattern): both equal the oracle's
    # two-distance integration and each other
    print("[case] integrate 40^3 x 8 frames, min_t 0.08, max_t 0.03, u16 fused and f32 per frame",
          flush=True)
    intr, objs, poses, depth, h, dim = small_scene(dev, seed=13)
    K = depth.shape[0]
    intr_np = np.array(intr, np.float32)
    vox = fusion.voxel_new((dim, dim, dim), 0.08, h, device=dev, max_t=0.03)
    assert (vox.min_t, vox.max_t) == (0.08, 0.03)
    fusion.integrate(vox, intr, poses, depth, max_w=128)
    ref = O.volume_init((dim, dim, dim))
    O.integrate_ex(ref, h, 0.08, 0.03, 128, intr_np, poses, depth.cpu().numpy())
    got = vox.data.cpu().numpy().view(np.uint32)
    assert got.tobytes() == ref.tobytes()
    sym = O.volume_init((dim, dim, dim))
    O.integrate_ex(sym, h, 0.03, 0.03, 128, intr_np, poses, depth.cpu().numpy())
    assert (sym != ref).mean() > 0.001  # the two distances change the result
    metres = (depth.to(torch.float32) * 0.001).contiguous()
    vox2 = fusion.voxel_new((dim, dim, dim), 0.08, h, device=dev, max_t=0.03)
    for k in range(K):
        fusion.integrate(vox2, intr, poses[k:k + 1], metres[k:k + 1], depth_scale=1.0, max_w=128)
    ref2 = O.volume_init((dim, dim, dim))
    O.integrate_ex(ref2, h, 0.08, 0.03, 128, intr_np, poses, metres.cpu().numpy(),
                   depth_scale=1.0)
    got2 = vox2.data.cpu().numpy().view(np.uint32)
    assert got2.tobytes() == ref2.tobytes()
    assert got2.tobytes() == got.tobytes()


def test_integrate_then_extract_matches_oracle(dev):
    print("[case] integrate + extract 40^3", flush=True)
    intr, objs, poses, depth, h, dim = small_scene(dev, K=12, seed=9)
    vox = fusion.voxel_new((dim, dim, dim), 0.05, h, device=dev)
    fusion.integrate(vox, intr, poses, depth, trunc=0.05, max_w=128)
    host = vox.data.cpu().numpy().view(np.uint32)
    r = mesh.extract(vox.data, h, want=("vv",))
    ref = O.extract(host, h, want=("vv",))
    assert r.count == ref["n"] and r.count > 100
    assert r.vv.cpu().numpy().tobytes() == ref["vv"].view(np.int32).reshape(-1, 7).tobytes()


def test_synth_depth_matches_numpy_raycast(dev):
    print("[case] synth depth", flush=True)
    intr = [60.0, 60.0, 31.5, 23.5]
    objs = np.zeros((2, 8), np.float32)
    objs[0, :5] = [0, 0.5, 0.5, 0.5, 0.2]             # sphere
    objs[1, :7] = [1, 0.8, 0.3, 0.6, 0.1, 0.1, 0.15]   # box
    poses = synth.orbit_poses(3, 1.0, radius=1.5, seed=2)
    W, H = 64, 48
    d = fusion.synth_depth(intr, poses, objs, W, H, sigma=0.0, dropout=0.0, device=dev)
    d = d.cpu().numpy().astype(np.float64) * 1e-3
    fx, fy, cx, cy = intr
    for k in range(3):
        R = poses[k, :9].reshape(3, 3).astype(np.float64)
        o = poses[k, 9:].astype(np.float64)
        uu, vv = np.meshgrid(np.arange(W), np.arange(H))
        dc = np.stack([(uu - cx) / fx, (vv - cy) / fy, np.ones_like(uu, dtype=np.float64)], -1)
        dw = dc @ R.T
        best = np.full((H, W), np.inf)
        # sphere
        L = o - objs[0, 1:4]
        A = (dw * dw).sum(-1)
        B = (dw * L).sum(-1)
        C = (L * L).sum() - objs[0, 4] ** 2
        disc = B * B - A * C
        s = np.where(disc >= 0, (-B - np.sqrt(np.maximum(disc, 0))) / A, np.inf)
        best = np.minimum(best, np.where(s > 1e-4, s, np.inf))
        # box
        lo = objs[1, 1:4] - objs[1, 4:7]
        hi = objs[1, 1:4] + objs[1, 4:7]
        with np.errstate(divide="ignore", invalid="ignore"):
            t0 = (lo - o) / dw
            t1 = (hi - o) / dw
        tmin = np.minimum(t0, t1).max(-1)
        tmax = np.maximum(t0, t1).min(-1)
        sb = np.where((tmax >= tmin) & (tmin > 1e-4), tmin, np.inf)
        best = np.minimum(best, sb)
        hit = np.isfinite(best)
        assert hit.mean() > 0.05
        assert ((d[k] > 0) == hit).mean() > 0.995  # silhouettes agree up to edge pixels
        both = hit & (d[k] > 0)
        assert np.abs(d[k][both] - best[both]).max() <= 0.0006  # rounding to 1 mm


@pytest.mark.parametrize("dims", [(40, 40, 40), (7, 5, 300), (300, 3, 20), (256, 2, 9),
                                  (1, 1, 700), (513, 1, 7)])
def test_per_frame_calls_on_odd_shapes(dev, dims):
    """One call per frame (the reference's pattern) into volumes whose
    256-voxel step carries into y, into z, or not at all (X below, at and
    above 256, X * Y below 256): each equals the oracle's integration."""
    X, Y, Z = dims
    print(f"[case] per-frame integration {X}x{Y}x{Z}, 8 frames", flush=True)
    intr, objs, poses, depth, h, _ = small_scene(dev, seed=21)
    # the cameras moved by the offset of the volume's centre from the scene's
    # (the same frames), so that the thin volumes cut through the scene
    poses = poses.copy()
    poses[:, 9:] -= (np.array([0.5, 0.5, 0.5]) - 0.5 * h * np.array(dims)).astype(np.float32)
    vox = fusion.voxel_new((X, Y, Z), 0.05, h, device=dev)
    for k in range(depth.shape[0]):
        fusion.integrate(vox, intr, poses[k:k + 1], depth[k:k + 1], trunc=0.05, max_w=128)
    ref = O.volume_init((Z, Y, X))
    O.integrate(ref, h, 0.05, 128, np.array(intr, np.float32), poses, depth.cpu().numpy())
    assert (ref != O.volume_init((1, 1, 1))[0, 0, 0]).sum() >= 30  # voxels the frames update
    assert vox.data.cpu().numpy().view(np.uint32).tobytes() == ref.tobytes()
