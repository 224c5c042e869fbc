"""The C++ drop-in API (include/marching_cube_amd/*.h over libmcx) driven by
tests/cpp/test_dropin.cpp: MarchingCubes::run, cu_voxel_to_mesh (PLY, OBJ),
MyMarchingCube, cu_voxel_rlc_to_mesh, cu_voxel_new +
_cu_voxel_volumetric_integration, _cu_voxel_extract_depth_and_normal_map and
cu_voxel_run_length_encode / _decode, each checked against the oracle."""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

from marching_cube_amd import build, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def read_ply(path):
    data = open(path, "rb").read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    header = data[:end].decode()
    nv = int(re.search(r"element vertex (\d+)", header).group(1))
    nf = int(re.search(r"element face (\d+)", header).group(1))
    body = np.frombuffer(data[end:end + nv * 24], dtype=np.float32).reshape(nv, 6)
    faces = data[end + nv * 24:]
    assert len(faces) == nf * 13
    return body[:, :3], body[:, 3:], nf


def test_dropin_api(dev, tmp_path):
    if not os.path.exists(build.TEST_BIN):
        pytest.skip("tests/cpp/test_dropin not built")
    vol, h = synth.sphere_volume(64)
    vol.tofile(tmp_path / "vol.bin")
    # text volume for cu_voxel_rlc_to_mesh: a 24^3 sphere placed in the
    # reference's window x, y in [512, 512+dim), z in [0, dim)
    small, hs = synth.sphere_volume(24, radius=0.3)
    t, w = synth.unpack_short2(small)
    with open(tmp_path / "rlc.txt", "w") as f:
        for x in range(24):
            for y in range(24):
                for z in range(24):
                    f.write(f"{x + 512} {y + 512} {z} {int(t[z, y, x])} {int(w[z, y, x])}\n")
    print("[case] drop-in binary", flush=True)
    res = subprocess.run([build.TEST_BIN, str(tmp_path / "vol.bin"), "64", "64", "64", repr(h),
                          str(tmp_path), str(tmp_path / "rlc.txt"), "24"],
                         cwd=tmp_path, capture_output=True, text=True, timeout=240)
    print(res.stdout[-2000:], res.stderr[-2000:])
    assert res.returncode == 0
    out = dict(line.split(" ", 1) for line in res.stdout.splitlines() if line.count(" ") == 1)

    ref = O.extract(vol, h, want=("vv", "pos"))
    nref = O.normals(vol, h, ref["pos"])
    n = ref["n"]
    assert int(out["run_count"]) == n
    assert int(out["run_default_buffer"]) == 27_000_000
    got = np.fromfile(tmp_path / "run_vv.bin", dtype=np.uint8)
    assert got.tobytes() == ref["vv"].tobytes()

    for name in ("mesh.ply", "my.ply"):
        p, nn, nf = read_ply(tmp_path / name)
        assert nf == n // 3
        assert p.tobytes() == ref["pos"].tobytes()
        assert nn.tobytes() == nref.tobytes()
    assert int(out["my_count"]) == n

    # OBJ: vertices shared by their "%.4f" key, faces with a repeated vertex dropped
    keys = {}
    vid = []
    for q in ref["pos"]:
        k = "%.4f_%.4f_%.4f" % (q[0], q[1], q[2])
        vid.append(keys.setdefault(k, len(keys) + 1))
    vid = np.array(vid).reshape(-1, 3)
    keep = (vid[:, 0] != vid[:, 1]) & (vid[:, 0] != vid[:, 2]) & (vid[:, 1] != vid[:, 2])
    lines = open(tmp_path / "mesh.obj").read().splitlines()
    assert sum(1 for L in lines if L.startswith("v ")) == len(keys)
    assert sum(1 for L in lines if L.startswith("vn ")) == len(keys)
    faces = np.array([[int(a) for a in L.split()[1:]] for L in lines if L.startswith("f ")])
    assert (faces == vid[keep]).all()
    assert len(open(tmp_path / "qi_volume.txt").read().splitlines()) == len(keys) + 1

    # rlc: same mesh as extracting the small volume with grid 0.003
    pr, _, nfr = read_ply(tmp_path / "rlc.ply")
    rr = O.extract(small, 0.003, want=("pos",))
    assert pr.tobytes() == rr["pos"].tobytes() and nfr == rr["n"] // 3

    # integration of one 1.2 m float depth frame into a 16^3 volume with the
    # volume's two truncation distances (min_t 0.2 behind, max_t 0.1 in front)
    ref_vol = O.volume_init((16, 16, 16))
    pose = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0.8, 0.8, -0.5], np.float32)
    depth = np.full((1, 24, 32), 1.2, np.float32)
    O.integrate_ex(ref_vol, 0.1, 0.2, 0.1, 128, np.array([40, 40, 15.5, 11.5], np.float32),
                   pose[None], depth, 1.0)
    fused = np.fromfile(tmp_path / "fused.bin", dtype=np.uint32).reshape(16, 16, 16)
    t_ref, w_ref = synth.unpack_short2(ref_vol)
    assert (w_ref > 0).any()
    assert ((t_ref < -16383) & (w_ref > 0)).any()  # beyond max_t behind: kept by min_t
    assert fused.tobytes() == ref_vol.tobytes()

    # raycast of the fused volume from the integrating camera
    intr = np.array([40, 40, 15.5, 11.5], np.float32)
    rd, rn = O.raycast(ref_vol, 0.1, intr, pose, 32, 24, origin=0.5)
    gd = np.fromfile(tmp_path / "ray_depth.bin", dtype=np.float32).reshape(24, 32)
    gn = np.fromfile(tmp_path / "ray_normal.bin", dtype=np.float32).reshape(24, 32, 3)
    assert gd.tobytes() == rd.tobytes() and gn.tobytes() == rn.tobytes()
    assert (gd > 0).mean() > 0.5
    assert (np.fromfile(tmp_path / "ray_color.bin", dtype=np.uint8) == 255).all()

    # .vol files
    assert out["vol_size"] == "64x64x64"
    canon = O.rle_decode_dump(O.rle_encode_dump(vol), 64, 64, 0, 64)
    loaded = np.fromfile(tmp_path / "vol_loaded.bin", dtype=np.uint32).reshape(64, 64, 64)
    assert (loaded == canon).all()
    E = canon.copy()
    E[30:47, 10:32, 20:51] = vol.view(np.uint32)[7:24, 5:27, 3:34]
    raw = open(tmp_path / "sphere.vol", "rb").read()
    assert raw[:16] == b"VOLUME_TSDF_RGB\x00"
    assert struct.unpack_from("<3i3f", raw, 16)[:3] == (64, 64, 64)
    assert np.float32(struct.unpack_from("<3f", raw, 28)[1]) == np.float32(0.04)
    assert raw[40:] == O.rle_encode_dump(E)
    E = O.rle_decode_dump(raw[40:], 64, 64, 0, 64)
    reg = np.fromfile(tmp_path / "vol_region.bin", dtype=np.uint32).reshape(20, 30, 40)
    expect = np.zeros_like(reg)
    expect[3:18, 2:27, 1:36] = E[30:45, 20:45, 10:45]
    assert (reg == expect).all()
