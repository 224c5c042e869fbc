"""GPU checks of the differentiable extraction (ExtractFn / mcx_extract_diff /
mcx_extract_backward):
* forward positions bit-identical to the plain extraction and the oracle;
* backward against the oracle's exact FP64 gradient of sum(g . p);
* bitwise reproducible backward (fixed-order sums, no atomics);
* FP64 central differences of the sphere-distance loss on topology-stable
  voxels of the 256^3 bumps field, relative error <= 1e-4."""
import numpy as np
import pytest
import torch

from marching_cube_amd import diff, mesh, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

R_TARGET = 0.35
CENTER = np.array([0.5, 0.5, 0.5])


def loss_and_grad_np(p):
    d = p - CENTER
    n = np.linalg.norm(d, axis=1)
    r = n - R_TARGET
    return (r * r).sum(), (2 * r / n)[:, None] * d


def test_forward_matches_plain_and_oracle(dev):
    print("[case] diff forward 48^3", flush=True)
    f, h = synth.bumps_volume(48, nbumps=16)
    v = torch.from_numpy(f).to(dev)
    p = diff.extract_positions(v, h)
    plain = mesh.extract(v, h, want=("pos",))
    ref = O.extract(f, h, want=("pos",))
    assert p.shape[0] == ref["n"] == plain.count
    assert p.detach().cpu().numpy().tobytes() == ref["pos"].tobytes()
    assert p.detach().cpu().numpy().tobytes() == plain.pos.cpu().numpy().tobytes()


@pytest.mark.parametrize("n,nb", [(40, 12), (96, 24)])
def test_backward_matches_fp64_gradient(dev, n, nb):
    print(f"[case] diff backward {n}^3", flush=True)
    f, h = synth.bumps_volume(n, nbumps=nb, seed=n)
    v = torch.from_numpy(f).to(dev).requires_grad_(True)
    p = diff.extract_positions(v, h)
    g = torch.randn(p.shape, generator=torch.Generator().manual_seed(n)).to(dev)
    (p * g).sum().backward()
    got = v.grad.cpu().numpy().astype(np.float64)
    ref = O.extract_backward_f64(f, h, g.cpu().numpy().astype(np.float64))
    scale = np.abs(ref).max()
    assert scale > 0
    err = np.abs(got - ref)
    assert err.max() <= 1e-5 * scale, err.max() / scale
    big = np.abs(ref) > 1e-3 * scale
    assert (err[big] / np.abs(ref[big])).max() < 1e-3
    assert (got[~(np.abs(ref) > 0)] == 0).all()  # no gradient far from the surface


def test_backward_is_deterministic(dev):
    print("[case] diff determinism", flush=True)
    f, h = synth.bumps_volume(64, nbumps=20, seed=3)
    outs = []
    for _ in range(2):
        v = torch.from_numpy(f).to(dev).requires_grad_(True)
        p = diff.extract_positions(v, h)
        d = p - torch.tensor(CENTER, dtype=torch.float32, device=dev)
        loss = ((d.norm(dim=1) - R_TARGET) ** 2).sum()
        loss.backward()
        outs.append(v.grad.cpu().numpy().tobytes())
    assert outs[0] == outs[1]


def local_loss_f64(f, h, idx, val):
    """Sphere-distance loss of the vertices of the 8 cubes around voxel idx,
    with f[idx] replaced by val, in FP64 (oracle extraction of the 3^3
    neighbourhood; vertices off the voxel's edges cancel in differences)."""
    z, y, x = idx
    sub = f[z - 1:z + 2, y - 1:y + 2, x - 1:x + 2].astype(np.float64).copy()
    sub[1, 1, 1] = val
    # oracle FP64 extraction takes fp32 input: evaluate the positions in FP64
    # directly with the same formula on the 2^3 cubes
    pos = []
    CXv = [0, 1, 1, 0, 0, 1, 1, 0]
    CYv = [0, 0, 1, 1, 0, 0, 1, 1]
    CZv = [0, 0, 0, 0, 1, 1, 1, 1]
    EA = [0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3]
    EB = [1, 2, 3, 0, 5, 6, 7, 4, 4, 5, 6, 7]
    edge, tri, nvt = TABLES
    for cz in range(2):
        for cy in range(2):
            for cx in range(2):
                fc = [sub[cz + CZv[k], cy + CYv[k], cx + CXv[k]] for k in range(8)]
                ci = sum(1 << k for k in range(8) if fc[k] < 0)
                for e in tri[ci]:
                    if e < 0:
                        break
                    a, b = EA[e], EB[e]
                    t = (0.0 - fc[a]) / (fc[b] - fc[a] + 1e-15)
                    pa = np.array([x - 1 + cx + CXv[a], y - 1 + cy + CYv[a], z - 1 + cz + CZv[a]])
                    pb = np.array([x - 1 + cx + CXv[b], y - 1 + cy + CYv[b], z - 1 + cz + CZv[b]])
                    pa = (pa + 0.5) * h
                    pb = (pb + 0.5) * h
                    pos.append(pa + t * (pb - pa))
    if not pos:
        return 0.0
    return loss_and_grad_np(np.array(pos))[0]


TABLES = None


def test_fd_config3_256(dev):
    """256^3 bumps field: FP64 central differences on 2,000 random
    topology-stable voxels (no sign change under +-delta at the voxel)."""
    global TABLES
    TABLES = O.tables()
    print("[case] diff FD 256^3", flush=True)
    f, h = synth.bumps_volume(256)
    v = torch.from_numpy(f).to(dev).requires_grad_(True)
    p = diff.extract_positions(v, h)
    d = p - torch.tensor(CENTER, dtype=torch.float32, device=dev)
    loss = ((d.norm(dim=1) - R_TARGET) ** 2).sum()
    loss.backward()
    grad = v.grad.cpu().numpy()
    # voxels that are corners of an active cube: the gradient is nonzero there
    cand = np.argwhere(grad != 0)
    inner = cand[(cand > 1).all(axis=1) & (cand < 254).all(axis=1)]
    rng = np.random.default_rng(20261202)
    delta = 1e-6
    checked, worst = 0, 0.0
    for i in rng.permutation(len(inner)):
        idx = tuple(int(a) for a in inner[i])
        fv = float(f[idx])
        if abs(fv) <= 2 * delta or abs(fv) < 1e-4:
            continue  # topology must not change under +-delta
        lp = local_loss_f64(f, h, idx, fv + delta)
        lm = local_loss_f64(f, h, idx, fv - delta)
        fd = (lp - lm) / (2 * delta)
        g = float(grad[idx])
        if abs(fd) < 1e-3 * np.abs(grad).max():
            continue  # relative error of tiny gradients is meaningless
        rel = abs(g - fd) / abs(fd)
        worst = max(worst, rel)
        checked += 1
        if checked >= 2000:
            break
    print(f"[fd] checked {checked} voxels, worst relative error {worst:.3e}", flush=True)
    assert checked >= 2000
    assert worst <= 1e-4


def test_shared_extractor_reuse_and_order_check(dev):
    """One DiffExtractor across calls: same positions and gradients as fresh
    buffers; a backward after another forward on the extractor raises."""
    print("[case] diff shared extractor", flush=True)
    f, h = synth.bumps_volume(40, nbumps=10, seed=9)
    ex = diff.DiffExtractor(f.shape, dev)
    ref_v = torch.from_numpy(f).to(dev).requires_grad_(True)
    ref_p = diff.extract_positions(ref_v, h)
    (ref_p ** 2).sum().backward()
    for _ in range(2):
        v = torch.from_numpy(f).to(dev).requires_grad_(True)
        p = diff.extract_positions(v, h, extractor=ex)
        assert p.detach().cpu().numpy().tobytes() == ref_p.detach().cpu().numpy().tobytes()
        (p ** 2).sum().backward()
        assert v.grad.cpu().numpy().tobytes() == ref_v.grad.cpu().numpy().tobytes()
    v1 = torch.from_numpy(f).to(dev).requires_grad_(True)
    p1 = diff.extract_positions(v1, h, extractor=ex)
    v2 = torch.from_numpy(f).to(dev).requires_grad_(True)
    diff.extract_positions(v2, h, extractor=ex)
    with pytest.raises(RuntimeError, match="another forward"):
        p1.sum().backward()


def test_recompute_backward_matches_stored_records(dev):
    """recompute=True: forwards of other volumes run between a forward and its
    backward on one shared extractor; the gradient equals the stored-record
    backward bit for bit."""
    print("[case] diff recompute backward", flush=True)
    f1, h = synth.bumps_volume(48, nbumps=14, seed=21)
    f2, _ = synth.bumps_volume(48, nbumps=9, seed=22)
    ref = []
    for f in (f1, f2):
        v = torch.from_numpy(f).to(dev).requires_grad_(True)
        p = diff.extract_positions(v, h)
        (p ** 3).sum().backward()
        ref.append(v.grad.cpu().numpy().tobytes())
    ex = diff.DiffExtractor(f1.shape, dev, recompute=True)
    v1 = torch.from_numpy(f1).to(dev).requires_grad_(True)
    v2 = torch.from_numpy(f2).to(dev).requires_grad_(True)
    p1 = diff.extract_positions(v1, h, extractor=ex)
    p2 = diff.extract_positions(v2, h, extractor=ex)  # overwrites the workspace
    (p1 ** 3).sum().backward()
    (p2 ** 3).sum().backward()
    assert v1.grad.cpu().numpy().tobytes() == ref[0]
    assert v2.grad.cpu().numpy().tobytes() == ref[1]
