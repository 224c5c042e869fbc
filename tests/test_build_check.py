"""The build's device-code check (marching_cube_amd.build.check_no_packed_fp32)
on two libraries built from one small kernel whose float2 arithmetic hipcc -O3
turns into packed FP32 instructions: built with the library's flags the check
passes, built without -packed-fp32-ops it fails."""
import os
import subprocess

import pytest

from marching_cube_amd import build

SRC = r"""
#include <hip/hip_runtime.h>
__global__ void k(float2* a, const float2* b, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    float2 x = a[i], y = b[i];
    x.x = x.x * y.x + 1.f; x.y = x.y * y.y + 1.f;
    x.x += y.y; x.y += y.x;
    a[i] = x;
  }
}
"""


@pytest.mark.skipif(not os.path.exists(build.HIPCC), reason="hipcc missing")
def test_packed_fp32_check(tmp_path):
    src = tmp_path / "probe.hip"
    src.write_text(SRC)
    libs = {}
    for name, flags in (("guarded", build.HIP_FLAGS),
                        ("plain", ["--offload-arch=gfx950", "-O3", "-fPIC"])):
        out = str(tmp_path / f"lib{name}.so")
        subprocess.check_call([build.HIPCC, *flags, "-shared", "-o", out, str(src)],
                              stderr=subprocess.DEVNULL)
        libs[name] = out
    assert build.check_no_packed_fp32(libs["guarded"]) > 0
    with pytest.raises(RuntimeError, match="packed FP32"):
        build.check_no_packed_fp32(libs["plain"])
