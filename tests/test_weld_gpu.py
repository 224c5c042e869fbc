"""Indexed mesh (mcx_extract_ex edge keys + mcx_weld) against a numpy
restatement built on the oracle's per-vertex (cube, edge) ids:
* edge keys = (lower corner voxel) * 3 + axis of the grid edge, bit-exact;
* unique vertices in order of first occurrence, each at its first
  occurrence's position; indices map every vertex to its unique vertex;
* on a closed surface the welded mesh is a closed, consistently oriented
  2-manifold with Euler characteristic 2."""
import numpy as np
import pytest
import torch

from marching_cube_amd import mesh, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CX = np.array([0, 1, 1, 0, 0, 1, 1, 0])
CY = np.array([0, 0, 1, 1, 0, 0, 1, 1])
CZ = np.array([0, 0, 0, 0, 1, 1, 1, 1])
EA = np.array([0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3])
EB = np.array([1, 2, 3, 0, 5, 6, 7, 4, 4, 5, 6, 7])


def ref_keys(cube, edge, X, Y):
    x = cube % X
    y = (cube // X) % Y
    z = cube // (X * Y)
    a, b = EA[edge], EB[edge]
    ax = np.where(CX[a] != CX[b], 0, np.where(CY[a] != CY[b], 1, 2))
    mx = x + np.minimum(CX[a], CX[b])
    my = y + np.minimum(CY[a], CY[b])
    mz = z + np.minimum(CZ[a], CZ[b])
    return ((mz * Y + my) * X + mx) * 3 + ax


def ref_weld(keys, pos):
    _, first, inv = np.unique(keys, return_index=True, return_inverse=True)
    order = np.argsort(first)            # unique keys by first occurrence
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    return pos[first[order]], rank[inv].astype(np.int32)


def gpu_keys_and_weld(f, h, dev):
    v = torch.from_numpy(f.view(np.int32) if f.dtype == np.uint32 else f).to(dev)
    ws = mesh.Workspace(dev)
    n = mesh.extract_into(v, h, cap=0, workspace=ws)
    from marching_cube_amd import _capi
    import ctypes
    lib = _capi.load()
    elem, X, Y, Z = mesh.volume_kind(v)
    pos = torch.empty((max(n, 1), 3), dtype=torch.float32, device=dev)
    key = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
    hc = ctypes.c_int64(0)
    _capi.check(lib.mcx_extract_ex(mesh._ptr(v), elem, X, Y, Z, float(h), None, mesh._ptr(pos),
                                   None, None, mesh._ptr(key), n, None, ctypes.byref(hc),
                                   ws.ptr, ws.nbytes, mesh._stream_handle(dev)), "extract_ex")
    upos, faces = mesh.indexed_mesh(v, h)
    return n, key[:n].cpu().numpy(), pos[:n].cpu().numpy(), upos.cpu().numpy(), faces.cpu().numpy()


@pytest.mark.parametrize("case", ["sphere", "noise", "bumps"])
def test_edge_keys_and_weld_match_reference(dev, case):
    print(f"[case] weld {case}", flush=True)
    if case == "sphere":
        f, h = synth.sphere_volume(96)
    elif case == "noise":
        f, h = synth.noise_volume((23, 19, 31), seed=7, p_zero_w=0.05), 0.1
    else:
        f, h = synth.bumps_volume(40, nbumps=10, seed=4)
    n, key, pos, upos, faces = gpu_keys_and_weld(f, h, dev)
    ref = O.extract(f, h, want=("pos",))
    _, cube, edge = O.extract_f64(f, h)
    Z, Y, X = f.shape[:3]
    assert n == ref["n"] == len(cube)
    assert (key == ref_keys(cube, edge.astype(np.int64), X, Y)).all()
    rupos, ridx = ref_weld(key, ref["pos"])
    assert upos.tobytes() == rupos.astype(np.float32).tobytes()
    assert faces.reshape(-1).tobytes() == ridx.tobytes()
    assert len(upos) < n  # vertices are shared


def test_welded_sphere_is_closed_manifold(dev):
    print("[case] welded sphere manifold", flush=True)
    f, h = synth.sphere_volume(128)
    v = torch.from_numpy(f.view(np.int32)).to(dev)
    upos, faces = mesh.indexed_mesh(v, h)
    faces = faces.cpu().numpy().astype(np.int64)
    U, F = len(upos), len(faces)
    d = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
    # every directed edge once (consistent orientation), its reverse once (closed)
    dk = d[:, 0] * U + d[:, 1]
    assert len(np.unique(dk)) == len(dk)
    rk = d[:, 1] * U + d[:, 0]
    assert np.isin(rk, dk).all()
    E = len(dk) // 2
    assert U - E + F == 2
    assert (faces[:, 0] != faces[:, 1]).sum() > 0.99 * F  # few degenerate triangles


@pytest.mark.parametrize("pattern,V", [("one_key", 1000), ("few", 1), ("few", 63), ("few", 65),
                                       ("few", 4097), ("pairs", 640), ("runs", 100003),
                                       ("mesh_like", 300001), ("negative", 5000)])
def test_weld_synthetic_keys(dev, pattern, V):
    """mcx_weld on its own (C ABI) with keys that repeat inside a wave, across
    waves and at wave boundaries, and sizes that are not multiples of 64:
    unique vertices in first-occurrence order and every index as the numpy
    restatement (np.unique first occurrences)."""
    print(f"[case] weld keys {pattern} V={V}", flush=True)
    from marching_cube_amd import _capi
    import ctypes
    rng = np.random.default_rng(V)
    if pattern == "one_key":
        keys = np.full(V, 123456789, np.int64)
    elif pattern == "few":
        keys = rng.integers(0, 50, V).astype(np.int64)
    elif pattern == "pairs":
        # lane 63 of a wave and lane 0 of the next share a key; so do lanes
        # 62 and 1 of one wave (the later lane is the repeat)
        keys = np.arange(V, dtype=np.int64) * 7
        for w in range(V // 64 - 1):
            keys[64 * w + 64] = keys[64 * w + 63]
            keys[64 * w + 62] = keys[64 * w + 1]
    elif pattern == "runs":
        keys = np.repeat(rng.integers(0, 1 << 40, V // 5 + 1), 5)[:V].astype(np.int64)
    elif pattern == "negative":
        # -1 (the bit pattern of an empty table slot, kept in a slot of its
        # own), the int64 extremes and other negative keys among small ones
        pool = np.array([-1, -2, np.iinfo(np.int64).min, np.iinfo(np.int64).max, 0, 1, 2, 3,
                         -(1 << 40)], np.int64)
        keys = pool[rng.integers(0, len(pool), V)]
    else:  # about six occurrences per key, spread over the array
        keys = (rng.integers(0, V // 6, V) * 3 + 1).astype(np.int64)
    pos = rng.standard_normal((V, 3)).astype(np.float32)
    lib = _capi.load()
    k = torch.from_numpy(keys).to(dev)
    p = torch.from_numpy(pos).to(dev)
    upos = torch.empty((V, 3), dtype=torch.float32, device=dev)
    idx = torch.empty(V, dtype=torch.int32, device=dev)
    wws = torch.empty(max(4096, lib.mcx_weld_workspace_size(V)), dtype=torch.uint8, device=dev)
    nu = ctypes.c_int64(0)
    _capi.check(lib.mcx_weld(mesh._ptr(k), mesh._ptr(p), 3, V, mesh._ptr(upos), mesh._ptr(idx),
                             None, ctypes.byref(nu), mesh._ptr(wws), wws.numel(),
                             mesh._stream_handle(dev)), "mcx_weld")
    rupos, ridx = ref_weld(keys, pos)
    assert nu.value == len(rupos)
    assert upos[: nu.value].cpu().numpy().tobytes() == rupos.tobytes()
    assert idx.cpu().numpy().tobytes() == ridx.tobytes()


def test_weld_empty(dev):
    f = synth.pack_short2(np.zeros((6, 7, 8)), np.zeros((6, 7, 8)))
    upos, faces = mesh.indexed_mesh(torch.from_numpy(f.view(np.int32)).to(dev), 0.1)
    assert len(upos) == 0 and len(faces) == 0
