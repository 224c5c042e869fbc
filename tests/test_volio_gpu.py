"""GPU `.vol` codec (mcx_rle_encode / mcx_rle_decode) against the CPU oracle:
* the dump bytes equal the oracle's for random, saturated, worst-case
  (every voxel its own literal run), long-literal-run (> 0x3fff) and fused
  volumes, odd slice sizes, slab boundaries and sub-ranges of slices;
* decoding equals the oracle's decode, for any subset and order of slices;
* file round trip (header + dump) and the region update/load semantics of
  cu_voxel_run_length_encode / _decode;
* malformed dumps are reported (MCX_ERR_FORMAT) and a small capacity gives
  MCX_ERR_CAPACITY with the required size."""
import ctypes

import numpy as np
import pytest
import torch

import test_rle_oracle as RO
from marching_cube_amd import _capi, fusion, volio
from marching_cube_amd.mesh import CuVoxel
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def random_volume(shape, seed, p_sat=0.6):
    rng = np.random.default_rng(seed)
    Z, Y, X = shape
    u = rng.random(shape)
    t = np.where(u < p_sat / 2, 32766, np.where(u < p_sat, -32766, rng.integers(-3, 4, shape)))
    # runs along x
    t[:, :, : X // 3] = t[:, :, :1]
    w = rng.integers(0, 300, shape)
    return (t.astype(np.int16).view(np.uint16).astype(np.uint32) |
            (w.astype(np.uint32) << np.uint32(16)))


def dev_vol(v, dev):
    return torch.from_numpy(np.ascontiguousarray(v).view(np.int32)).to(dev)


CASES = {
    "random_odd": lambda: random_volume((7, 13, 29), 1),
    "random_large": lambda: random_volume((9, 130, 150), 2),
    "saturated": lambda: np.full((5, 200, 300), RO.vox(32766, 0), np.uint32),
    "worst_case": lambda: (np.arange(4 * 33 * 47, dtype=np.uint32).reshape(4, 33, 47) % 1000) | (
        np.uint32(7) << np.uint32(16)),
    "long_literal": lambda: np.full((3, 170, 120), RO.vox(5, 9), np.uint32),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_encode_decode_bit_exact(dev, case):
    print(f"[case] rle {case}", flush=True)
    v = CASES[case]()
    Z, Y, X = v.shape
    ref = O.rle_encode_dump(v)
    got = volio.encode(dev_vol(v, dev))
    assert got == ref
    # slab boundaries inside the volume
    assert volio.encode(dev_vol(v, dev), slab_voxels=2 * X * Y + 1) == ref
    # a sub-range of slices
    if Z >= 3:
        assert volio.encode(dev_vol(v, dev), 1, Z - 2) == O.rle_encode_dump(v, 1, Z - 2)
    offs = volio.slice_offsets(ref, Z)
    dec = volio.decode(ref, X, Y, offs, dev).cpu().numpy().view(np.uint32)
    assert (dec == O.rle_decode_dump(ref, X, Y, 0, Z)).all()
    order = np.arange(Z)[::-1].copy()
    dec_r = volio.decode(ref, X, Y, offs[order], dev).cpu().numpy().view(np.uint32)
    assert (dec_r == dec[order]).all()


def test_fused_volume_round_trip(dev, tmp_path):
    print("[case] rle fused 96^3 volume file round trip", flush=True)
    vox = fusion.kinfu_volume(dim=96, frames=30, device=dev)
    v = vox.data.cpu().numpy().view(np.uint32)
    ref = O.rle_encode_dump(v)
    assert volio.encode(vox.data) == ref
    path = str(tmp_path / "fused.vol")
    volio.save(path, vox)
    with open(path, "rb") as fp:
        size, min_t, max_t, grid_s = volio.read_header(fp)
        assert fp.read() == ref
    assert size == (96, 96, 96) and np.float32(grid_s) == np.float32(vox.grid_s)
    back = volio.load(path, dev)
    canon = O.rle_decode_dump(ref, 96, 96, 0, 96)
    assert (back.data.cpu().numpy().view(np.uint32) == canon).all()
    # canonical volumes survive a second round trip unchanged
    volio.save(path, back)
    assert (volio.load(path, dev).data.cpu().numpy().view(np.uint32) == canon).all()


def test_region_update_and_load(dev, tmp_path):
    print("[case] rle region encode/decode", flush=True)
    g = random_volume((12, 40, 50), 5)
    path = str(tmp_path / "g.vol")
    gvox = CuVoxel(data=dev_vol(g, dev), grid_s=0.01, min_t=0.03, max_t=0.03)
    volio.run_length_encode(gvox, path)  # new file: the whole volume
    canon = O.rle_decode_dump(O.rle_encode_dump(g), 50, 40, 0, 12)
    loc = random_volume((6, 20, 30), 6)
    lvox = CuVoxel(data=dev_vol(loc, dev), grid_s=0.01)
    bl, bg, sz = (3, 2, 1), (10, 15, 4), (20, 12, 5)
    volio.run_length_encode(lvox, path, bl, bg, sz)
    expect = canon.copy()
    expect[4:9, 15:27, 10:30] = loc[1:6, 2:14, 3:23]
    expect = O.rle_decode_dump(O.rle_encode_dump(expect), 50, 40, 0, 12)
    got = volio.run_length_decode(None, path, device=dev)
    assert (got.data.cpu().numpy().view(np.uint32) == expect).all()
    # region load into a local volume
    dst = CuVoxel(data=torch.zeros((6, 20, 30), dtype=torch.int32, device=dev), grid_s=0.01)
    volio.run_length_decode(dst, path, (2, 1, 0), (5, 7, 3), (25, 18, 6))
    d = dst.data.cpu().numpy().view(np.uint32)
    assert (d[0:6, 1:19, 2:27] == expect[3:9, 7:25, 5:30]).all()
    assert (d[:, 0, :] == 0).all() and (d[:, :, :2] == 0).all()
    with pytest.raises(ValueError):
        volio.run_length_decode(dst, path, (0, 0, 0), (40, 0, 0), (20, 5, 5))
    with pytest.raises(ValueError):
        bad = CuVoxel(data=dst.data, grid_s=0.02)
        volio.run_length_encode(bad, path, (0, 0, 0), (0, 0, 0), (2, 2, 2))


def test_malformed_dumps_and_capacity(dev):
    print("[case] rle malformed dumps and capacity", flush=True)
    v = random_volume((3, 10, 12), 7)
    ref = O.rle_encode_dump(v)
    offs = volio.slice_offsets(ref, 3)
    # the first record's header of slice 1 rewritten as a run of length 0
    bad = bytearray(ref)
    o = int(offs[1]) + 4
    bad[o:o + 2] = (0x4000).to_bytes(2, "little")
    with pytest.raises(_capi.McxError) as e:
        volio.decode(bytes(bad), 12, 10, offs, dev)
    assert e.value.status == _capi.ERR_FORMAT
    # slice 2 claims one byte too many (its last record then needs bytes
    # beyond the dump, or bytes are left over)
    bad = bytearray(ref) + b"\x00\x00"
    o = int(offs[2])
    (b,) = np.frombuffer(bytes(bad[o:o + 4]), np.uint32)
    bad[o:o + 4] = int(b + 2).to_bytes(4, "little")
    with pytest.raises(_capi.McxError):
        volio.decode(bytes(bad), 12, 10, offs, dev)
    # an offset outside the dump
    with pytest.raises(_capi.McxError):
        volio.decode(ref, 12, 10, np.array([0, len(ref) + 10, 0]), dev)
    # the good dump still decodes
    good = volio.decode(ref, 12, 10, offs, dev).cpu().numpy().view(np.uint32)
    assert (good == O.rle_decode_dump(ref, 12, 10, 0, 3)).all()
    # capacity: nothing written, required size reported
    lib = _capi.load()
    dv = dev_vol(v, dev)
    ws = torch.empty(lib.mcx_rle_encode_workspace_size(12, 10, 3), dtype=torch.uint8, device=dev)
    out = torch.zeros(16, dtype=torch.uint8, device=dev)
    total = ctypes.c_int64(0)
    st = lib.mcx_rle_encode(dv.data_ptr(), 12, 10, 3, 0, 3, out.data_ptr(), 16, None,
                            ctypes.byref(total), ws.data_ptr(), ws.numel(),
                            torch.cuda.current_stream(dev).cuda_stream)
    assert st == _capi.ERR_CAPACITY and total.value == len(ref)
    assert (out.cpu().numpy() == 0).all()
