"""ctypes binding of libmcx.so (the C ABI declared in include/mcx.h).

The library is loaded from the package directory; a missing library raises
immediately -- there is no CPU fallback for any entry point.
"""
import ctypes
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libmcx.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "mcx.h")

OK = 0
ERR_ARG = 1
ERR_CAPACITY = 2
ERR_WORKSPACE = 3
ERR_HIP = 4
ERR_INTERNAL = 5
ERR_UNSUPPORTED = 6
ERR_FORMAT = 7

ELEM_SHORT2 = 0
ELEM_F32 = 1

DEPTH_U16 = 0
DEPTH_F32 = 1

TUNE_COLUMNS_PER_CU = 0
TUNE_EMIT_GROUPS_PER_CU = 1
TUNE_BATCH_SLAB = 2

_vp = ctypes.c_void_p
_i32 = ctypes.c_int
_i64 = ctypes.c_int64
_f32 = ctypes.c_float
_sz = ctypes.c_size_t
_u32 = ctypes.c_uint32

# name -> (restype, argtypes)
SIGNATURES = {
    "mcx_version": (_i32, []),
    "mcx_strerror": (ctypes.c_char_p, [_i32]),
    "mcx_get_tables": (_i32, [_vp, _vp]),
    "mcx_extract_workspace_size": (_sz, [_i32, _i32, _i32]),
    "mcx_extract_workspace_size_ex": (_sz, [_i32, _i32, _i32, _i64]),
    "mcx_workspace_init": (_i32, [_vp, _sz, _vp]),
    "mcx_workspace_status": (_i32, [_vp, _sz, _vp]),
    "mcx_extract": (_i32, [_vp, _i32, _i32, _i32, _i32, _f32, _vp, _vp, _vp, _vp, _i64, _vp,
                           _vp, _vp, _sz, _vp]),
    "mcx_classify": (_i32, [_vp, _i32, _i32, _i32, _i32, _vp, _vp, _vp]),
    "mcx_extract_static_limit": (_sz, [_i32, _i32, _i32]),
    "mcx_extract_static": (_i32, [_vp, _i32, _i32, _i32, _i32, _f32, _vp, _vp, _vp, _vp, _i64,
                                  _vp, _vp, _vp]),
    "mcx_batch_create": (_i32, [_vp, _i32, _vp, _vp, _vp, _i32]),
    "mcx_batch_destroy": (None, [_vp]),
    "mcx_batch_tiles": (_i64, [_vp]),
    "mcx_batch_workspace_size": (_sz, [_vp, _i64]),
    "mcx_extract_batched": (_i32, [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mcx_normals": (_i32, [_vp, _i32, _i32, _i32, _f32, _vp, _i64, _i64, _vp, _vp]),
    "mcx_extract_diff": (_i32, [_vp, _i32, _i32, _i32, _f32, _vp, _i64, _vp, _vp, _vp, _sz, _vp]),
    "mcx_backward_scratch_size": (_sz, [_i32, _i32, _i32, _sz]),
    "mcx_extract_backward": (_i32, [_vp, _i32, _i32, _i32, _f32, _vp, _vp, _vp, _sz, _vp, _sz,
                                    _vp]),
    "mcx_extract_backward_recompute": (_i32, [_vp, _i32, _i32, _i32, _f32, _vp, _vp, _vp, _sz,
                                              _vp, _sz, _vp]),
    "mcx_volume_init": (_i32, [_vp, _i32, _i32, _i32, _vp]),
    "mcx_set_tuning": (_i32, [_i32, _i32]),
    "mcx_extract_ex": (_i32, [_vp, _i32, _i32, _i32, _i32, _f32, _vp, _vp, _vp, _vp, _vp, _i64,
                              _vp, _vp, _vp, _sz, _vp]),
    "mcx_weld_workspace_size": (_sz, [_i64]),
    "mcx_weld": (_i32, [_vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mcx_get_tuning": (_i32, [_i32]),
    "mcx_tsdf_init": (_i32, [_vp, _vp, _i64, _f32, _vp]),
    "mcx_integrate_tsdf": (_i32, [_vp, _vp, _i32, _i32, _i32, _f32, _f32, _i32, _vp, _vp, _vp,
                                  _f32, _i32, _i32, _i32, _vp, _sz, _vp]),
    "mcx_integrate_tsdf_backward_workspace_size": (_sz, [_i32, _i32, _i32]),
    "mcx_integrate_tsdf_backward": (_i32, [_vp, _vp, _i32, _i32, _i32, _f32, _f32, _i32, _vp,
                                           _vp, _vp, _f32, _i32, _i32, _i32, _vp, _vp, _sz,
                                           _vp]),
    "mcx_integrate_workspace_size": (_sz, [_i32, _i32]),
    "mcx_integrate": (_i32, [_vp, _i32, _i32, _i32, _f32, _f32, _i32, _vp, _vp, _vp, _f32, _i32,
                             _i32, _i32, _vp, _sz, _vp]),
    "mcx_integrate_f32": (_i32, [_vp, _i32, _i32, _i32, _f32, _f32, _i32, _vp, _vp, _vp, _f32,
                                 _i32, _i32, _i32, _vp, _sz, _vp]),
    "mcx_raycast": (_i32, [_vp, _i32, _i32, _i32, _f32, _f32, _vp, _vp, _i32, _i32, _vp, _vp,
                           _vp]),
    "mcx_rle_encode_bound": (_sz, [_i32, _i32, _i32]),
    "mcx_rle_encode_workspace_size": (_sz, [_i32, _i32, _i32]),
    "mcx_rle_encode": (_i32, [_vp, _i32, _i32, _i32, _i32, _i32, _vp, _sz, _vp, _vp, _vp, _sz,
                              _vp]),
    "mcx_rle_decode_workspace_size": (_sz, [_i64, _i32, _i32, _i32]),
    "mcx_rle_decode": (_i32, [_vp, _i64, _vp, _i32, _i32, _i32, _vp, _vp, _sz, _vp]),
    "mcx_integrate_ex": (_i32, [_vp, _i32, _i32, _i32, _f32, _f32, _f32, _i32, _vp, _vp, _vp,
                                _i32, _f32, _i32, _i32, _i32, _vp, _sz, _vp]),
    "mcx_batch_voxels": (_i64, [_vp]),
    "mcx_integrate_tsdf_batched": (_i32, [_vp, _vp, _vp, _i32, _f32, _f32, _i32, _vp, _vp, _vp,
                                          _f32, _i32, _i32, _i32, _vp, _sz, _vp]),
    "mcx_extract_diff_batched": (_i32, [_vp, _vp, _i64, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mcx_batch_backward_scratch_size": (_sz, [_vp, _sz]),
    "mcx_extract_backward_batched": (_i32, [_vp, _vp, _vp, _vp, _vp, _sz, _vp, _sz, _vp]),
    "mcx_integrate_tsdf_backward_batched_workspace_size": (_sz, [_vp, _i32, _i32, _i32]),
    "mcx_integrate_tsdf_backward_batched": (_i32, [_vp, _vp, _vp, _vp, _f32, _i32, _vp, _vp,
                                                   _vp, _f32, _i32, _i32, _i32, _vp, _vp, _sz,
                                                   _vp]),
    "mcx_chamfer_workspace_size": (_sz, [_i64]),
    "mcx_chamfer_batched": (_i32, [_vp, _vp, _i32, _vp, _i32, _i64, _vp, _vp, _vp, _sz, _vp]),
    "mcx_synth_depth": (_i32, [_vp, _i32, _i32, _i32, _vp, _vp, _vp, _i32, _f32, _f32, _u32,
                               _f32, _vp]),
}


class McxError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = strerror(status)
        super().__init__(f"{what}: {msg} (status {status})" if what else msg)


_lib = None


def load():
    """Load libmcx.so and bind every symbol; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m marching_cube_amd.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def strerror(status):
    try:
        return load().mcx_strerror(status).decode()
    except Exception:  # library missing
        return f"status {status}"


def check(status, what=""):
    if status != OK:
        raise McxError(status, what)
    return status


def declared_functions(header=HEADER):
    """Names of the functions declared in include/mcx.h."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mcx_[a-z0-9_]+)\s*\(", text)))
