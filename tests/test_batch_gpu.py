"""Batched extraction (mcx_extract_batched) against per-volume oracle
extraction: every volume's vertex range is bit-exact, the bases are the
prefix of the per-volume counts, including empty and degenerate volumes."""
import numpy as np
import pytest
import torch

from marching_cube_amd import batch, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_batch_matches_per_volume_oracle(dev):
    print("[case] batch of 60 shape volumes", flush=True)
    vols, hs = synth.batch_shapes(60, lo=6, hi=48, seed=11)
    # empty, degenerate (Z = 1) and noise volumes in the middle and at the end
    vols.insert(7, synth.pack_short2(np.zeros((9, 10, 11)), np.zeros((9, 10, 11))))
    hs = np.insert(hs, 7, 0.1)
    vols.insert(20, synth.noise_volume((1, 12, 13), seed=2))
    hs = np.insert(hs, 20, 0.1)
    vols.insert(33, synth.noise_volume((17, 9, 40), seed=4, p_zero_w=0.1))
    hs = np.insert(hs, 33, 0.05)
    vols.append(synth.noise_volume((5, 1, 7), seed=3))
    hs = np.append(hs, 0.2)
    tens = [torch.from_numpy(v.view(np.int32)).to(dev) for v in vols]
    plan = batch.BatchPlan(tens, hs)
    n, vbase, out = plan.extract(want=("vv", "v2vox"))
    vbase = vbase.cpu().numpy()
    vv = out["vv"].cpu().numpy()
    v2 = out["v2vox"].cpu().numpy()
    counts = []
    for b, (v, h) in enumerate(zip(vols, hs)):
        ref = O.extract(v, float(h), want=("vv", "v2vox"))
        counts.append(ref["n"])
        seg = slice(vbase[b], vbase[b + 1])
        assert vbase[b + 1] - vbase[b] == ref["n"], b
        assert vv[seg].tobytes() == ref["vv"].view(np.int32).reshape(-1, 7).tobytes(), b
        assert v2[seg].tobytes() == ref["v2vox"].tobytes(), b
    assert n == sum(counts) == vbase[-1]
    assert vbase[0] == 0


def test_batch_repeat_and_f32(dev):
    print("[case] batch repeat + fp32", flush=True)
    fs = [synth.bumps_volume(n, nbumps=6, seed=n)[0] for n in (12, 20, 33, 17)]
    hs = [1.0 / f.shape[0] for f in fs]
    plan = batch.BatchPlan([torch.from_numpy(f).to(dev) for f in fs], hs)
    first = None
    for rep in range(3):
        n, vbase, out = plan.extract(want=("pos",))
        pos = out["pos"].cpu().numpy()
        if first is None:
            first = pos.tobytes()
        assert pos.tobytes() == first
    vb = vbase.cpu().numpy()
    for b, (f, h) in enumerate(zip(fs, hs)):
        ref = O.extract(f, h, want=("pos",))
        assert pos[vb[b]:vb[b + 1]].tobytes() == ref["pos"].tobytes()


@pytest.mark.parametrize("groups", [1, 8, 32])
def test_batch_and_single_launch_shapes(dev, groups):
    # waves taking many 64-record units (few emit workgroups) and few: the
    # batched segments and one-call-per-volume outputs equal the oracle
    from marching_cube_amd import autotune, mesh
    print(f"[case] batch of 150 shape volumes, {groups} emit groups per CU", flush=True)
    vols, hs = synth.batch_shapes(150, lo=12, hi=80, seed=23)
    tens = [torch.from_numpy(v.view(np.int32)).to(dev) for v in vols]
    saved = autotune.get()
    try:
        autotune.set(emit_groups_per_cu=groups)
        plan = batch.BatchPlan(tens, hs)
        n, vbase, out = plan.extract(want=("vv",))
        vbase = vbase.cpu().numpy()
        vv = out["vv"].cpu().numpy()
        ws = mesh.Workspace(dev)
        for b, (v, h) in enumerate(zip(vols, hs)):
            ref = O.extract(v, float(h), want=("vv",))["vv"].view(np.int32).reshape(-1, 7)
            assert vv[vbase[b]:vbase[b + 1]].tobytes() == ref.tobytes(), ("batched", b)
            r = mesh.extract(tens[b], float(h), want=("vv",), workspace=ws)
            assert r.vv.cpu().numpy().tobytes() == ref.tobytes(), ("single", b)
    finally:
        autotune.set(saved["columns_per_cu"], saved["emit_groups_per_cu"], saved["batch_slab"])


@pytest.mark.parametrize("records", [0, 2000])
def test_batch_staged_path(dev, records):
    # a record scratch too small for the batch: tiles take the staged path
    # (re-read from their volume), all or some of them
    import ctypes
    from marching_cube_amd import _capi
    from marching_cube_amd.mesh import VV_WORDS
    print(f"[case] batch staged path, records={records}", flush=True)
    vols, hs = synth.batch_shapes(40, lo=10, hi=60, seed=31)
    tens = [torch.from_numpy(v.view(np.int32)).to(dev) for v in vols]
    plan = batch.BatchPlan(tens, hs)
    lib = _capi.load()
    ws = torch.empty(max(4096, lib.mcx_batch_workspace_size(plan.handle, records)),
                     dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    assert lib.mcx_workspace_init(ws.data_ptr(), ws.numel(), stream) == _capi.OK
    refs = [O.extract(v, float(h), want=("vv",))["vv"].view(np.int32).reshape(-1, 7)
            for v, h in zip(vols, hs)]
    total = sum(len(r) for r in refs)
    vv = torch.empty((total, VV_WORDS), dtype=torch.int32, device=dev)
    vbase = torch.empty(len(vols) + 1, dtype=torch.int64, device=dev)
    hc = ctypes.c_int64(0)
    st = lib.mcx_extract_batched(plan.handle, vv.data_ptr(), None, None, total, vbase.data_ptr(),
                                 None, ctypes.byref(hc), ws.data_ptr(), ws.numel(), stream)
    assert st == _capi.OK and hc.value == total
    vb = vbase.cpu().numpy()
    out = vv.cpu().numpy()
    for b, ref in enumerate(refs):
        assert out[vb[b]:vb[b + 1]].tobytes() == ref.tobytes(), b


@pytest.mark.parametrize("slab", [16, 64])
def test_batch_at_shape_batch_sizes(dev, slab):
    # the size range of the batched-shapes workload (sides 16..128 per axis),
    # plus a full 128^3 sphere and a 33 x 127 x 125 noise volume (rows that
    # end inside a lane, dense surface): vertex_voxel records and the int3
    # map of every volume against the oracle, at two plane counts per column
    from marching_cube_amd import autotune
    print(f"[case] batch of 242 volumes, sides 16..128, slab {slab}", flush=True)
    vols, hs = synth.batch_shapes(240, lo=16, hi=128, seed=41)
    sph, hsph = synth.sphere_volume(128)
    vols.append(sph)
    hs = np.append(hs, np.float32(hsph))
    vols.append(synth.noise_volume((33, 127, 125), seed=5, p_zero_w=0.02))
    hs = np.append(hs, np.float32(1.0 / 128))
    assert max(max(v.shape) for v in vols) == 128
    tens = [torch.from_numpy(v.view(np.int32)).to(dev) for v in vols]
    saved = autotune.get()
    try:
        autotune.set(batch_slab=slab)
        plan = batch.BatchPlan(tens, hs)
    finally:
        autotune.set(saved["columns_per_cu"], saved["emit_groups_per_cu"], saved["batch_slab"])
    n, vbase, out = plan.extract(want=("vv", "v2vox"))
    plan.check()
    vb = vbase.cpu().numpy()
    vv = out["vv"].cpu().numpy()
    v2 = out["v2vox"].cpu().numpy()
    for b, (v, h) in enumerate(zip(vols, hs)):
        ref = O.extract(v, float(h), want=("vv", "v2vox"))
        seg = slice(vb[b], vb[b + 1])
        assert vb[b + 1] - vb[b] == ref["n"], b
        assert vv[seg].tobytes() == ref["vv"].view(np.int32).reshape(-1, 7).tobytes(), b
        assert v2[seg].tobytes() == ref["v2vox"].tobytes(), b
    assert n == vb[-1] > 1_000_000
