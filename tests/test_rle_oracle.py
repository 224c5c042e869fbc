"""Pins the CPU oracle of the `.vol` run-length codec
(cu_kinect_fusion_v1.cu:1362-1447 for the 4-byte voxel of structure.h:14-18):
* hand-assembled byte strings: +MAX / -MAX runs (2-byte records), literal
  runs (2-byte record + the run's first voxel), weights not compared;
* the 0x3fff run limit (16383 -> one record, 16384 -> two, 40000 -> three),
  for a saturated and for a literal run (the second record carries the
  voxel where it starts);
* round trip equals the canonical volume computed independently (every
  voxel replaced by its record's first voxel, saturated runs with weight
  0); re-encoding the decoded volume reproduces the bytes;
* a v = 0 record decodes as a literal; malformed slices are reported;
* the dump (per-slice uint32 byte count + records) and sub-ranges of it."""
import struct

import numpy as np
import pytest

from oracle import oracle as O

MAXT = 32766


def vox(t, w=0):
    return (np.uint32(np.uint16(np.int16(t))) | (np.uint32(w) << np.uint32(16)))


def voxels(pairs):
    return np.array([vox(t, w) for t, w in pairs], np.uint32)


def test_hand_assembled_slice():
    v = voxels([(MAXT, 0), (MAXT, 5), (MAXT, 9), (100, 3), (100, 4), (-MAXT, 1), (7, 2)])
    expect = bytes([0x03, 0x40,
                    0x02, 0xC0, 0x64, 0x00, 0x03, 0x00,
                    0x01, 0x80,
                    0x01, 0xC0, 0x07, 0x00, 0x02, 0x00])
    assert O.rle_encode_slice(v) == expect
    dec = O.rle_decode_slice(expect, 7)
    assert (dec == voxels([(MAXT, 0)] * 3 + [(100, 3)] * 2 + [(-MAXT, 0), (7, 2)])).all()


@pytest.mark.parametrize("n,hdrs", [(16383, [0xBFFF]), (16384, [0xBFFF, 0x8001]),
                                    (40000, [0xBFFF, 0xBFFF, 0x9C42])])
def test_run_limit_saturated(n, hdrs):
    v = np.full(n, vox(-MAXT, 3), np.uint32)
    assert O.rle_encode_slice(v) == struct.pack(f"<{len(hdrs)}H", *hdrs)


def test_run_limit_literal():
    n = 16384
    v = np.array([vox(5, i % 7) for i in range(n)], np.uint32)
    b = O.rle_encode_slice(v)
    assert b == struct.pack("<HhHHhH", 0xFFFF, 5, 0, 0xC001, 5, 16383 % 7)
    dec = O.rle_decode_slice(b, n)
    assert (dec[:16383] == vox(5, 0)).all() and dec[16383] == vox(5, 16383 % 7)


def canonical(v):
    """Each voxel replaced by the first voxel of its record (runs of equal
    tsdf split every 0x3fff voxels); saturated records carry weight 0."""
    t = (v & 0xFFFF).astype(np.uint16).view(np.int16)
    n = len(v)
    change = np.ones(n, bool)
    change[1:] = t[1:] != t[:-1]
    run_start = np.maximum.accumulate(np.where(change, np.arange(n), 0))
    rec_start = run_start + ((np.arange(n) - run_start) // 0x3fff) * 0x3fff
    out = v[rec_start].copy()
    sat = np.abs(t.astype(np.int32)) == MAXT
    out[sat] = (t[sat].view(np.uint16)).astype(np.uint32)
    nrec = int((rec_start == np.arange(n)).sum())
    nlit = int(((rec_start == np.arange(n)) & ~sat).sum())
    return out, 2 * nrec + 4 * nlit


def random_slice(n, seed):
    rng = np.random.default_rng(seed)
    kinds = rng.choice(4, size=n, p=[0.5, 0.2, 0.2, 0.1])
    t = np.where(kinds == 0, MAXT, np.where(kinds == 1, -MAXT,
                 rng.integers(-5, 6, n)))
    # long runs
    t[n // 3: n // 3 + 20000] = MAXT
    t[n // 2: n // 2 + 17000] = 4
    w = rng.integers(0, 200, n)
    return (t.astype(np.int16).view(np.uint16).astype(np.uint32) |
            (w.astype(np.uint32) << np.uint32(16)))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_round_trip_is_canonical_and_reencode_is_idempotent(seed):
    v = random_slice(60000, seed)
    b = O.rle_encode_slice(v)
    canon, nbytes = canonical(v)
    assert len(b) == nbytes
    dec = O.rle_decode_slice(b, len(v))
    assert (dec == canon).all()
    assert O.rle_encode_slice(dec) == b


def test_v0_literal_and_malformed():
    lit0 = struct.pack("<HhH", 0x0002, -9, 4)  # v = 0: literal as v = 3
    assert (O.rle_decode_slice(lit0, 2) == vox(-9, 4)).all()
    assert O.rle_decode_slice(struct.pack("<H", 0x4000), 1) is None          # n = 0
    assert O.rle_decode_slice(struct.pack("<HH", 0xC001, 7), 1) is None      # truncated
    assert O.rle_decode_slice(struct.pack("<H", 0x4003), 2) is None          # past the slice
    assert O.rle_decode_slice(struct.pack("<HH", 0x4001, 0x4001), 1) is None  # bytes left
    assert O.rle_decode_slice(struct.pack("<H", 0x4001), 2) is None          # short


def test_dump_and_subranges():
    Z, Y, X = 5, 7, 9
    rng = np.random.default_rng(3)
    vol = np.stack([random_slice(X * Y, s).reshape(Y, X) for s in range(Z)])
    d = O.rle_encode_dump(vol)
    off, sizes = 0, []
    for z in range(Z):
        (b,) = struct.unpack_from("<I", d, off)
        assert d[off + 4: off + 4 + b] == O.rle_encode_slice(vol[z])
        sizes.append(b)
        off += 4 + b
    assert off == len(d)
    full = O.rle_decode_dump(d, X, Y, 0, Z)
    for z in range(Z):
        assert (full[z].reshape(-1) == canonical(vol[z].reshape(-1))[0]).all()
    sub = O.rle_decode_dump(d, X, Y, 2, 2)
    assert (sub == full[2:4]).all()
    assert O.rle_decode_dump(d[:-1], X, Y, 0, Z) is None
    part = O.rle_encode_dump(vol, 1, 3)
    assert part == d[4 + sizes[0]: 4 + sizes[0] + 12 + sum(sizes[1:4])]
    del rng
