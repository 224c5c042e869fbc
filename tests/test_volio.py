"""Host side of the `.vol` files (no GPU): the 40-byte header
(cu_kinect_fusion_v1.cu:1303-1326) and the slice index of a dump."""
import io
import struct

import numpy as np
import pytest

import test_rle_oracle as RO
from marching_cube_amd import _capi, volio
from oracle import oracle as O


def test_header_layout_and_round_trip():
    f = io.BytesIO()
    volio.write_header(f, (512, 256, 128), 0.03, 0.04, 0.003)
    raw = f.getvalue()
    assert len(raw) == 40
    assert raw[:16] == b"VOLUME_TSDF_RGB\x00"
    assert struct.unpack_from("<3i3f", raw, 16)[:3] == (512, 256, 128)
    f.seek(0)
    size, min_t, max_t, grid_s = volio.read_header(f)
    assert size == (512, 256, 128)
    assert np.float32(min_t) == np.float32(0.03) and np.float32(grid_s) == np.float32(0.003)
    with pytest.raises(ValueError):
        volio.read_header(io.BytesIO(b"VOLUME_TSDF_RGX\x00" + raw[16:]))
    with pytest.raises(ValueError):
        volio.read_header(io.BytesIO(raw[:20]))


def test_slice_offsets():
    vol = np.stack([RO.random_slice(6 * 5, s).reshape(5, 6) for s in range(4)])
    d = O.rle_encode_dump(vol)
    offs = volio.slice_offsets(d, 4)
    assert offs[0] == 0
    for z in range(3):
        (b,) = struct.unpack_from("<I", d, int(offs[z]))
        assert offs[z + 1] == offs[z] + 4 + b
    with pytest.raises(_capi.McxError):
        volio.slice_offsets(d[:-1], 4)
    with pytest.raises(_capi.McxError):
        volio.slice_offsets(d, 5)
