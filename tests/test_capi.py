"""C-ABI surface checks that need no GPU: the library loads, exports every
function include/mcx.h declares, and rejects bad arguments before touching
the device."""
import ctypes
import subprocess

import pytest

from marching_cube_amd import _capi


def test_library_exports_every_declared_function():
    lib = _capi.load()
    declared = _capi.declared_functions()
    assert "mcx_extract" in declared
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding knows the signature of every declared function
    assert sorted(_capi.SIGNATURES) == declared


def test_exported_symbols_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for n in _capi.declared_functions():
        assert n in names, n


def test_version_and_strerror():
    lib = _capi.load()
    assert lib.mcx_version() >= 10000
    for st in range(8):
        assert lib.mcx_strerror(st)
    assert b"capacity" in lib.mcx_strerror(_capi.ERR_CAPACITY)


def test_argument_errors_without_device():
    lib = _capi.load()
    hc = ctypes.c_int64(0)
    # null volume, bad dims, bad kind, f32 with vertex_voxel output
    assert lib.mcx_extract(None, 0, 4, 4, 4, 1.0, None, None, None, None, 0, None,
                           ctypes.byref(hc), None, 0, None) == _capi.ERR_ARG
    fake = ctypes.c_void_p(0x1000)
    assert lib.mcx_extract(fake, 0, 0, 4, 4, 1.0, None, None, None, None, 0, None,
                           None, None, 0, None) == _capi.ERR_ARG
    assert lib.mcx_extract(fake, 7, 4, 4, 4, 1.0, None, None, None, None, 0, None,
                           None, None, 0, None) == _capi.ERR_ARG
    assert lib.mcx_extract(fake, 1, 4, 4, 4, 1.0, fake, None, None, None, 10, None,
                           None, None, 0, None) == _capi.ERR_ARG
    # size limits (checked before any device work): X*Y + X + 1024 < 2^30, Z <= 2^24
    assert lib.mcx_extract(fake, 0, 4, 4, (1 << 24) + 1, 1.0, None, None, None, None, 0, None,
                           None, None, 0, None) == _capi.ERR_UNSUPPORTED
    assert lib.mcx_extract(fake, 0, 1 << 15, 1 << 15, 4, 1.0, None, None, None, None, 0, None,
                           None, None, 0, None) == _capi.ERR_UNSUPPORTED
    assert lib.mcx_classify(None, 0, 4, 4, 4, None, None, None) == _capi.ERR_ARG
    # differentiable integration: the backward handles one pass of <= 512 frames;
    # the checks return before any device work
    intr = (ctypes.c_float * 4)(60.0, 60.0, 31.5, 23.5)
    K, W, H = 600, 64, 48
    wsb = lib.mcx_integrate_tsdf_backward_workspace_size(W, H, K)
    assert wsb >= K * W * H * 8
    assert lib.mcx_integrate_tsdf_backward(fake, None, 8, 8, 8, 0.1, 0.3, 128, intr, fake, fake,
                                           1.0, W, H, K, fake, fake, wsb, None) \
        == _capi.ERR_UNSUPPORTED
    assert lib.mcx_integrate_tsdf_backward(None, None, 8, 8, 8, 0.1, 0.3, 128, intr, fake, fake,
                                           1.0, W, H, 4, fake, fake, wsb, None) == _capi.ERR_ARG
    assert lib.mcx_integrate_tsdf_backward(fake, None, 8, 8, 8, 0.1, 0.3, 128, intr, fake, fake,
                                           1.0, W, H, 4, fake, fake, 16, None) \
        == _capi.ERR_WORKSPACE

    # raycast: null pointers, a too-small volume, h <= 0, a singular camera
    pose = (ctypes.c_float * 12)(1, 0, 0, 0, 1, 0, 0, 0, 1, 0.1, 0.1, 0.0)
    assert lib.mcx_raycast(None, 8, 8, 8, 0.1, 0.0, intr, pose, W, H, fake, None,
                           None) == _capi.ERR_ARG
    assert lib.mcx_raycast(fake, 1, 8, 8, 0.1, 0.0, intr, pose, W, H, fake, None,
                           None) == _capi.ERR_ARG
    assert lib.mcx_raycast(fake, 8, 8, 8, 0.0, 0.0, intr, pose, W, H, fake, None,
                           None) == _capi.ERR_ARG
    sing = (ctypes.c_float * 4)(0.0, 60.0, 31.5, 23.5)
    assert lib.mcx_raycast(fake, 8, 8, 8, 0.1, 0.0, sing, pose, W, H, fake, None,
                           None) == _capi.ERR_ARG

def test_workspace_size_grows_with_volume():
    lib = _capi.load()
    a = lib.mcx_extract_workspace_size(16, 16, 16)
    b = lib.mcx_extract_workspace_size(512, 512, 512)
    assert 256 < a < b
    # one u64 descriptor per tile and parity
    assert b >= 256 + 16 * (511 * 512 * 512 // 1024)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_capi, "_lib", None)
    monkeypatch.setattr(_capi, "LIB_PATH", str(tmp_path / "libmcx.so"))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _capi.load()


def test_tuning_set_get_and_bounds():
    lib = _capi.load()
    keys = (_capi.TUNE_COLUMNS_PER_CU, _capi.TUNE_EMIT_GROUPS_PER_CU, _capi.TUNE_BATCH_SLAB)
    saved = [lib.mcx_get_tuning(k) for k in keys]
    assert saved == [56, 16, 64]  # defaults
    try:
        assert lib.mcx_set_tuning(_capi.TUNE_COLUMNS_PER_CU, 32) == _capi.OK
        assert lib.mcx_get_tuning(_capi.TUNE_COLUMNS_PER_CU) == 32
        assert lib.mcx_set_tuning(_capi.TUNE_COLUMNS_PER_CU, 0) == _capi.ERR_ARG
        assert lib.mcx_set_tuning(_capi.TUNE_EMIT_GROUPS_PER_CU, 65) == _capi.ERR_ARG
        assert lib.mcx_set_tuning(99, 1) == _capi.ERR_ARG
        assert lib.mcx_get_tuning(99) == -1
    finally:
        for k, v in zip(keys, saved):
            lib.mcx_set_tuning(k, v)
