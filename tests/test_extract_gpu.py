"""GPU parity of the fused extraction kernel (libmcx mcx_extract) against the
CPU oracle: counts and cube indices bit-exact, every vertex_voxel record
bit-exact (positions included: the kernel and the oracle evaluate the same
fp32 sequence), the int3 map and ver_voxel bit-exact."""
import numpy as np
import pytest
import torch

from marching_cube_amd import _capi, mesh, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def to_dev(vol, dev):
    v = np.ascontiguousarray(vol)
    if v.dtype == np.uint32:
        v = v.view(np.int32)
    return torch.from_numpy(v).to(dev)


def gpu_extract(vol_np, h, dev, want=("vv", "v2vox", "ver_voxel"), cap=None):
    print(f"[case] extract shape={vol_np.shape} dtype={vol_np.dtype} want={want}", flush=True)
    v = to_dev(vol_np, dev)
    r = mesh.extract(v, h, want=want, cap=cap)
    torch.cuda.synchronize()
    return r


def check_against_oracle(vol, h, dev, want=("vv", "v2vox", "ver_voxel")):
    ref = O.extract(vol, h, want=want)
    r = gpu_extract(vol, h, dev, want=want)
    assert r.count == ref["n"]
    if "vv" in want:
        got = r.vv.cpu().numpy()
        assert got.tobytes() == ref["vv"].view(np.int32).reshape(-1, 7).tobytes()
    if "pos" in want:
        assert r.pos.cpu().numpy().tobytes() == ref["pos"].tobytes()
    if "v2vox" in want:
        assert r.v2vox.cpu().numpy().tobytes() == ref["v2vox"].tobytes()
    if "ver_voxel" in want:
        assert r.ver_voxel.cpu().numpy().tobytes() == ref["ver_voxel"].tobytes()
    return ref["n"]


def test_sphere_128_bit_exact(dev):
    vol, h = synth.sphere_volume(128)
    n = check_against_oracle(vol, h, dev)
    assert n == 166596


@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 5, 7), (9, 10, 11), (17, 33, 65),
                                   (5, 3, 1030), (4, 700, 6), (33, 31, 129), (2, 2, 4099)])
def test_noise_volumes_bit_exact(dev, shape):
    vol = synth.noise_volume(shape, seed=sum(shape), p_zero_w=0.1)
    check_against_oracle(vol, 0.003, dev)


def test_classify_kernel_matches_oracle(dev):
    for shape in [(6, 7, 9), (20, 21, 22)]:
        vol = synth.noise_volume(shape, seed=2, p_zero_w=0.15)
        ci, nv, _ = O.classify(vol)
        gci, gnv = mesh.classify(to_dev(vol, dev))
        assert np.array_equal(gci.cpu().numpy(), ci)
        assert np.array_equal(gnv.cpu().numpy(), nv)


def test_fp32_field_positions_bit_exact(dev):
    f, h = synth.bumps_volume(48, nbumps=16)
    check_against_oracle(f, h, dev, want=("pos", "v2vox"))


def test_empty_volume(dev):
    vol = synth.pack_short2(np.zeros((8, 9, 10)), np.zeros((8, 9, 10)))
    r = gpu_extract(vol, 1.0, dev)
    assert r.count == 0
    assert not r.ver_voxel.cpu().numpy().any()


@pytest.mark.parametrize("shape", [(1, 5, 6), (5, 1, 6), (5, 6, 1), (1, 1, 1)])
def test_degenerate_shapes_have_no_cubes(dev, shape):
    # a side of one voxel: no cube, no vertex; ver_voxel is still cleared
    vol = synth.noise_volume(shape, seed=7)
    assert O.extract(vol, 0.5, want=("vv",))["n"] == 0
    r = gpu_extract(vol, 0.5, dev)
    assert r.count == 0
    assert not r.ver_voxel.cpu().numpy().any()


def test_device_error_word_reaches_every_caller(dev):
    # The look-back's error word, set here from the host (a wait never hits
    # its limit in a correct run): an asynchronous call's device count reads
    # -1, the status query reports the error once and clears it, a
    # synchronous call reports it as well, and later calls are unaffected.
    print("[case] error word set from the host, async and sync calls", flush=True)
    vol = synth.sphere_volume(64)[0]
    ref = O.extract(vol, 1.0 / 64, want=("vv",))
    n = ref["n"]
    v = to_dev(vol, dev)
    ws = mesh.Workspace(dev)
    assert mesh.extract_into(v, 1.0 / 64, cap=0, workspace=ws) == n
    err_word = ws.buf[12:16]  # WsHeader.error (after three 32-bit counters)
    dc = torch.zeros(1, dtype=torch.int64, device=dev)
    vv = torch.empty((n, 7), dtype=torch.int32, device=dev)
    err_word.copy_(torch.tensor([1, 0, 0, 0], dtype=torch.uint8, device=dev))
    mesh.extract_into(v, 1.0 / 64, vv=vv, cap=n, d_count=dc, workspace=ws, sync=False)
    torch.cuda.synchronize()
    assert int(dc.item()) == -1
    with pytest.raises(_capi.McxError) as e:
        ws.check()
    assert e.value.status == _capi.ERR_INTERNAL
    ws.check()  # reported once
    err_word.copy_(torch.tensor([1, 0, 0, 0], dtype=torch.uint8, device=dev))
    with pytest.raises(_capi.McxError) as e:
        mesh.extract_into(v, 1.0 / 64, vv=vv, cap=n, workspace=ws)
    assert e.value.status == _capi.ERR_INTERNAL
    assert mesh.extract_into(v, 1.0 / 64, vv=vv, cap=n, d_count=dc, workspace=ws) == n
    assert int(dc.item()) == n
    assert vv.cpu().numpy().tobytes() == ref["vv"].view(np.int32).reshape(-1, 7).tobytes()
    ws.check()


def test_capacity_overflow_reports_required_count(dev):
    vol = synth.noise_volume((12, 13, 14), seed=4)
    ref = O.extract(vol, 0.5, want=("vv",))
    n = ref["n"]
    cap = n // 3
    v = to_dev(vol, dev)
    buf = torch.full((cap, 7), -1, dtype=torch.int32, device=dev)
    got = mesh.extract_into(v, 0.5, vv=buf, cap=cap)
    assert got == n
    assert buf.cpu().numpy().tobytes() == ref["vv"].view(np.int32).reshape(-1, 7)[:cap].tobytes()


def test_deterministic_and_workspace_reuse(dev):
    # alternating sizes on one workspace exercises the ticket reset and the
    # parity flip of the look-back descriptors
    ws = mesh.Workspace(dev)
    vols = [synth.noise_volume(s, seed=i) for i, s in
            enumerate([(40, 41, 42), (3, 4, 5), (40, 41, 42), (20, 64, 64), (7, 7, 7)])]
    refs = [O.extract(v, 0.1, want=("vv",)) for v in vols]
    for rep in range(2):
        for v, ref in zip(vols, refs):
            r = mesh.extract(to_dev(v, dev), 0.1, want=("vv",), workspace=ws)
            assert r.count == ref["n"]
            assert r.vv.cpu().numpy().tobytes() == ref["vv"].view(np.int32).reshape(-1, 7).tobytes()


def test_marching_cubes_run_dropin(dev):
    vol, h = synth.sphere_volume(64)
    mc = mesh.MarchingCubes(dev)
    tri = torch.empty((200000, 7), dtype=torch.int32, device=dev)
    v2 = torch.empty((200000, 3), dtype=torch.int32, device=dev)
    view, tb, v2b, verv = mc.run(mesh.CuVoxel(to_dev(vol, dev), h), tri, v2, None)
    ref = O.extract(vol, h, want=("vv", "v2vox", "ver_voxel"))
    assert view.shape[0] == ref["n"] and view.data_ptr() == tri.data_ptr()
    assert view.cpu().numpy().tobytes() == ref["vv"].view(np.int32).reshape(-1, 7).tobytes()
    assert v2b[: ref["n"]].cpu().numpy().tobytes() == ref["v2vox"].tobytes()
    assert verv.cpu().numpy().tobytes() == ref["ver_voxel"].tobytes()


def test_full_size_sphere_512(dev):
    vol, h = synth.sphere_volume(512)
    n = check_against_oracle(vol, h, dev, want=("vv",))
    assert n > 2_000_000


@pytest.mark.parametrize("records", [0, 200])
def test_record_overflow_uses_staged_path(dev, records):
    # with no (or little) record scratch every (or most) tile re-reads its
    # voxel ranges from the volume; the output must not change
    from marching_cube_amd import _capi
    lib = _capi.load()
    for shape in [(20, 33, 47), (9, 64, 130)]:
        vol = synth.noise_volume(shape, seed=7, p_zero_w=0.1)
        Z, Y, X = shape
        ws = mesh.Workspace(dev, lib.mcx_extract_workspace_size_ex(X, Y, Z, records), grow=False)
        ref = O.extract(vol, 0.01, want=("vv", "v2vox", "ver_voxel"))
        v = to_dev(vol, dev)
        r = mesh.extract(v, 0.01, want=("vv", "v2vox", "ver_voxel"), workspace=ws)
        assert r.count == ref["n"]
        assert r.vv.cpu().numpy().tobytes() == ref["vv"].view(np.int32).reshape(-1, 7).tobytes()
        assert r.v2vox.cpu().numpy().tobytes() == ref["v2vox"].tobytes()
        assert r.ver_voxel.cpu().numpy().tobytes() == ref["ver_voxel"].tobytes()


def layered_volume(shape, seed):
    """Noise with surfaces parallel to the planes: layers of z where every
    cube is active (tiles of hundreds of records) between noisy slabs."""
    rng = np.random.default_rng(seed)
    Z, Y, X = shape
    t = np.where((np.arange(Z) % 5 < 2)[:, None, None], 400, -400) * np.ones((Z, Y, X), np.int64)
    t = t + rng.integers(-30, 30, shape)
    noisy = rng.random(shape) < 0.3
    t[noisy] = rng.integers(-2000, 2000, int(noisy.sum()))
    w = rng.integers(1, 100, shape)
    w[rng.random(shape) < 0.02] = 0
    return synth.pack_short2(t.astype(np.int16), w.astype(np.uint16))


@pytest.mark.parametrize("records", [300, 3000, None])
def test_large_tiles_and_mixed_staged_path(dev, records):
    # 128-bit count kernel: tiles of up to 511 records, so the
    # 64-record emission units start deep inside tiles; with a small record
    # scratch, staged-path tiles are interleaved with tiles that kept their
    # records
    lib = _capi.load()
    for shape in [(16, 12, 256), (11, 40, 96)]:
        print(f"[case] layered {shape} records={records}", flush=True)
        vol = layered_volume(shape, 11)
        Z, Y, X = shape
        ws = None
        if records is not None:
            ws = mesh.Workspace(dev, lib.mcx_extract_workspace_size_ex(X, Y, Z, records),
                                grow=False)
        ref = O.extract(vol, 0.01, want=("vv", "v2vox", "ver_voxel"))
        r = mesh.extract(to_dev(vol, dev), 0.01, want=("vv", "v2vox", "ver_voxel"), workspace=ws)
        assert r.count == ref["n"]
        assert r.vv.cpu().numpy().tobytes() == ref["vv"].view(np.int32).reshape(-1, 7).tobytes()
        assert r.v2vox.cpu().numpy().tobytes() == ref["v2vox"].tobytes()
        assert r.ver_voxel.cpu().numpy().tobytes() == ref["ver_voxel"].tobytes()


@pytest.mark.parametrize("shape", [(4, 6, 2600), (2050, 3, 5)])
def test_long_axis(dev, shape):
    # one axis above 2048 voxels, the others a few voxels (rows of one
    # voxel's tiles spanning several planes, or tiles of many rows)
    vol = synth.noise_volume(shape, seed=13, p_zero_w=0.05)
    check_against_oracle(vol, 0.0037, dev)


def test_normals_bit_exact(dev):
    for vol, h in [synth.sphere_volume(64), (synth.noise_volume((12, 13, 14), seed=3), 0.02)]:
        print(f"[case] normals {vol.shape}", flush=True)
        ref = O.extract(vol, h, want=("vv", "pos"))
        nref = O.normals(vol, h, ref["pos"])
        v = to_dev(vol, dev)
        r = mesh.extract(v, h, want=("vv",))
        n_vv = mesh.normals(v, h, r.vv)
        n_pos = mesh.normals(v, h, torch.from_numpy(ref["pos"]).to(dev))
        assert n_vv.cpu().numpy().tobytes() == nref.tobytes()
        assert n_pos.cpu().numpy().tobytes() == nref.tobytes()


def test_launch_shapes_do_not_change_output(dev):
    """Every tuning candidate gives the same bytes (vertex_voxel, int3 map,
    ver_voxel); the autotuner returns one of the candidates and applies it."""
    from marching_cube_amd import autotune
    print("[case] tunings x sphere 128 + noise", flush=True)
    vols = [synth.sphere_volume(128)]
    vols.append((synth.noise_volume((33, 40, 52), seed=5, p_zero_w=0.05), 0.02))
    saved = autotune.get()
    try:
        for f, h in vols:
            v = torch.from_numpy(f.view(np.int32)).to(dev)
            ref = O.extract(f, h, want=("vv", "v2vox", "ver_voxel"))
            for c, g in ((16, 4), (64, 8), (128, 16), (7, 1)):
                autotune.set(c, g)
                r = mesh.extract(v, h, want=("vv", "v2vox", "ver_voxel"))
                assert r.count == ref["n"], (c, g)
                assert r.vv.cpu().numpy().tobytes() == ref["vv"].view(np.int32).reshape(-1, 7).tobytes()
                assert r.v2vox.cpu().numpy().tobytes() == ref["v2vox"].tobytes()
                assert r.ver_voxel.cpu().numpy().tobytes() == ref["ver_voxel"].tobytes()
        best = autotune.tune_extract(torch.from_numpy(vols[0][0].view(np.int32)).to(dev),
                                     vols[0][1], columns=(32, 64), emit_groups=(4, 8), reps=2)
        assert (best["columns_per_cu"], best["emit_groups_per_cu"]) in {(32, 4), (32, 8), (64, 4), (64, 8)}
        assert autotune.get()["columns_per_cu"] == best["columns_per_cu"]
    finally:
        autotune.set(saved["columns_per_cu"], saved["emit_groups_per_cu"], saved["batch_slab"])


def test_workspace_shared_across_volume_sizes(dev):
    """One workspace, volumes of different sizes in turn (the scan's
    look-back descriptors sit at size-dependent places): every call matches
    the oracle, including a large volume after smaller ones."""
    print("[case] one workspace, sizes 300/64/200/300/150", flush=True)
    ws = mesh.Workspace(dev)
    for n in (300, 64, 200, 300, 150):
        f, h = synth.sphere_volume(n, radius=0.3 + 0.01 * (n % 7))
        v = torch.from_numpy(f.view(np.int32)).to(dev)
        ref = O.extract(f, h, want=("vv",), nthreads=8)
        cnt = mesh.extract_into(v, h, cap=0, workspace=ws)
        assert cnt == ref["n"], n
        vv = torch.empty((cnt, 7), dtype=torch.int32, device=dev)
        assert mesh.extract_into(v, h, vv=vv, cap=cnt, workspace=ws) == cnt
        assert vv.cpu().numpy().tobytes() == ref["vv"].view(np.int32).reshape(-1, 7).tobytes(), n
