"""GPU checks of the batched differentiable fusion pass (FusionGraph: batched
fp32 integration, batched differentiable extraction, batched Chamfer loss,
batched MC backward, batched integration backward):

* every stage of every volume against its single-volume entry point and the
  oracle: fused TSDF and weights bit-exact against the oracle's fp32
  integration, positions bit-exact against the oracle's extraction, dL/dpos
  as the nearest-target closed form, loss against an FP64 sum, dL/dtsdf
  bit-exact against mcx_extract_backward (and within 1e-5 of the max of the
  oracle's FP64 gradient), dL/d(depth) against mcx_integrate_tsdf_backward
  (the two differ only by their fixed-point scale);
* at the sizes and input type of the batched end-to-end workload (64^3 -
  96^3 volumes, 16 synthetic uint16 320x240 frames each, in metres): the
  captured graph's dL/d(depth) against FP64 central differences of the
  oracle's FP64 chain (integration -> extraction -> Chamfer loss) on sampled
  topology-stable pixels, relative error <= 1e-4."""
import numpy as np
import pytest
import torch

from marching_cube_amd import _capi, diff, fusion, fusion_diff, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def fib_sphere(n, center, radius):
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    th = np.pi * (1 + 5 ** 0.5) * i
    return (center + radius * np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi),
                                        np.cos(phi)], -1)).astype(np.float32)


def make_batch(dev, shapes, K, W, H, M, seed=0):
    """Per volume: a unit-cube scene of 3 random spheres/boxes in front of a
    far background box, K orbit poses, uint16 depth frames (1 mm units, 1.5 mm
    noise, 2 % dropout) converted to float metres, M targets on a sphere."""
    intr = np.array([0.82 * W, 0.82 * W, (W - 1) / 2, (H - 1) / 2], np.float32)
    bg = np.array([[1, 0.5, 0.5, 0.5, 20.0, 20.0, 20.0, 0]], np.float32)
    hs, depth, poses, targets = [], [], [], []
    for b, (Z, Y, X) in enumerate(shapes):
        hs.append(1.0 / max(X, Y, Z))
        objs = np.concatenate([synth.kinfu_scene(3, extent=1.0, seed=seed + b), bg])
        p = synth.orbit_poses(K, extent=1.0, radius=1.6, seed=seed + b)
        d16 = fusion.synth_depth(intr, p, objs, W, H, sigma=0.0015, dropout=0.02,
                                 seed=seed + b, device=dev)
        depth.append((d16.to(torch.float32) * 0.001).contiguous())
        poses.append(torch.from_numpy(p).to(dev))
        targets.append(torch.from_numpy(fib_sphere(M, np.full(3, 0.5), 0.25)).to(dev))
    return intr, hs, depth, poses, targets


def test_batched_stages_match_single_volume_and_oracle(dev):
    print("[case] FusionGraph 4 uneven volumes x 6 frames 96x72 vs single-volume + oracle",
          flush=True)
    shapes = [(40, 44, 48), (36, 40, 32), (33, 33, 33), (50, 30, 45)]
    K, W, H, M, trunc = 6, 96, 72, 48, 0.03
    intr, hs, depth, poses, targets = make_batch(dev, shapes, K, W, H, M)
    g = fusion_diff.FusionGraph(shapes, hs, depth, poses, targets, intr, trunc=trunc)
    g.step()
    g.check()
    vb = g.vbase.cpu().numpy()
    pos_all = g.pos.cpu().numpy()
    gpos_all = g.gpos.cpu().numpy()
    assert (np.diff(vb) > 0).all()
    for b, (Z, Y, X) in enumerate(shapes):
        h = hs[b]
        dnp = depth[b].cpu().numpy()
        pnp = poses[b].cpu().numpy()
        # fused fp32 TSDF and weights: the oracle's integration from (-1, 0)
        rt = np.full((Z, Y, X), -1.0, np.float32)
        rw = np.zeros((Z, Y, X), np.int32)
        O.integrate_tsdf(rt, rw, h, trunc, 128, intr, pnp, dnp)
        tsdf = g.tsdf[b].cpu().numpy()
        assert tsdf.tobytes() == rt.tobytes(), b
        assert (g.w[b].cpu().numpy() == rw).all(), b
        # positions: the oracle's extraction
        ref = O.extract(rt, h, want=("pos",))
        pos = pos_all[vb[b]:vb[b + 1]]
        assert pos.tobytes() == ref["pos"].tobytes(), b
        # dL/dpos: 2 (p - nearest target); loss: FP64 sum of squared distances
        t = targets[b].cpu().numpy()
        d2 = ((pos[:, None, :].astype(np.float64) - t[None].astype(np.float64)) ** 2).sum(-1)
        j = d2.argmin(1)
        srt = np.sort(d2, 1)
        clear = srt[:, 1] - srt[:, 0] > 1e-6 * srt[:, 1]  # no near-tie between targets
        expect = (2.0 * (pos - t[j])).astype(np.float32)
        gp = gpos_all[vb[b]:vb[b + 1]]
        assert clear.mean() > 0.99
        assert gp[clear].tobytes() == expect[clear].tobytes(), b
        L = d2[np.arange(len(pos)), j].sum()
        assert abs(float(g.loss[b]) - L) <= 1e-5 * L, b
        # dL/dtsdf: the single-volume backward on the same dL/dpos
        vt = torch.from_numpy(rt).to(dev)
        ex = diff.DiffExtractor((Z, Y, X), dev)
        p1 = ex.forward(vt, h)
        assert p1.cpu().numpy().tobytes() == pos.tobytes()
        gv = ex.backward(vt, torch.from_numpy(gp.copy()).to(dev))
        gt = g.gtsdf[b].cpu().numpy()
        assert gt.tobytes() == gv.cpu().numpy().tobytes(), b
        g64 = O.extract_backward_f64(rt, h, gp.astype(np.float64))
        assert np.abs(gt - g64).max() <= 1e-5 * np.abs(g64).max(), b
        # dL/d(depth): the single-volume integration backward (weights from 0)
        integ = fusion_diff._Integrator((Z, Y, X), W, H, K, h, trunc, 128, intr, dev)
        gd1 = torch.empty_like(depth[b])
        integ.backward(torch.from_numpy(gt).to(dev), torch.zeros((Z, Y, X), dtype=torch.int32,
                                                                  device=dev),
                       poses[b], depth[b], gd1)
        a, r = g.grad_depth[b].cpu().numpy(), gd1.cpu().numpy()
        assert np.abs(r).max() > 0
        assert np.abs(a - r).max() <= 1e-6 * np.abs(r).max(), b
    # deterministic: a second pass gives the same bits
    keep = (g.loss.clone(), g.grad_depth_all.clone(), g.gtsdf_all.clone())
    g.step()
    torch.cuda.synchronize()
    assert torch.equal(g.loss, keep[0])
    assert torch.equal(g.grad_depth_all, keep[1])
    assert torch.equal(g.gtsdf_all, keep[2])


def f64_chain(h, trunc, poses, depth64, targets, shape, intr):
    t = np.full(shape, -1.0, np.float64)
    w = np.zeros(shape, np.int32)
    O.integrate_tsdf_f64(t, w, h, trunc, 128, intr, poses, depth64)
    pos, cube, edge = O.extract_f64(t, h)
    d2 = ((pos[:, None, :] - targets[None, :, :].astype(np.float64)) ** 2).sum(-1)
    j = d2.argmin(1)
    return d2[np.arange(len(pos)), j].sum(), (cube, edge, j)


def test_graph_depth_gradient_matches_fp64_fd_at_batch_sizes(dev):
    print("[case] FusionGraph graph replay, 4 volumes 64^3-96^3 x 16 uint16 320x240 frames, "
          "dL/ddepth vs FP64 central differences", flush=True)
    shapes = [(64, 64, 64), (80, 80, 80), (96, 96, 96), (72, 72, 72)]
    K, W, H, M, trunc = 16, 320, 240, 64, 0.03
    intr, hs, depth, poses, targets = make_batch(dev, shapes, K, W, H, M, seed=7)
    g = fusion_diff.FusionGraph(shapes, hs, depth, poses, targets, intr, trunc=trunc)
    g.capture()
    g.replay()
    g.replay()
    torch.cuda.synchronize()
    g.check()
    grad = g.grad_depth_all.cpu().numpy().astype(np.float64)
    loss = g.loss.cpu().numpy()
    rng = np.random.default_rng(20261203)
    delta = 1e-6
    checked, worst = 0, 0.0
    for b in range(len(shapes)):
        h = hs[b]
        d64 = depth[b].cpu().numpy().astype(np.float64)
        pnp = poses[b].cpu().numpy()
        tnp = targets[b].cpu().numpy()
        L0, topo0 = f64_chain(h, trunc, pnp, d64, tnp, shapes[b], intr)
        assert abs(float(loss[b]) - L0) <= 1e-5 * abs(L0), b
        gb = grad[b]
        cand = np.argwhere(np.abs(gb) > 1e-2 * np.abs(gb).max())
        done_b = 0
        for idx in rng.permutation(len(cand)):
            k, v, u = (int(a) for a in cand[idx])
            dp, dm = d64.copy(), d64.copy()
            dp[k, v, u] += delta
            dm[k, v, u] -= delta
            Lp, tp = f64_chain(h, trunc, pnp, dp, tnp, shapes[b], intr)
            Lm, tm = f64_chain(h, trunc, pnp, dm, tnp, shapes[b], intr)
            stable = all(a.shape == c.shape and (a == c).all() for t in (tp, tm)
                         for a, c in zip(t, topo0))
            if not stable:
                continue  # a case, an edge or a nearest target changes under +-delta
            fd = (Lp - Lm) / (2 * delta)
            rel = abs(gb[k, v, u] - fd) / abs(fd)
            worst = max(worst, rel)
            done_b += 1
            if done_b >= 8:
                break
        checked += done_b
    print(f"[fd] checked {checked} pixels over {len(shapes)} volumes, worst relative error "
          f"{worst:.3e}", flush=True)
    assert checked >= 24
    assert worst <= 1e-4
