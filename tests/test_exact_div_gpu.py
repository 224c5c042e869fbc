"""The exact fp32 forms the extraction and integration kernels use for short2
volumes (mcx_internal.h: tsdf_unit, div_weight_short2, tsdf_fraction) equal
IEEE division on their whole input domain, checked on the device by
tests/cpp/exact_div_check (65536 unpacks by 32766, 2^31 interpolation-weight
pairs, 65536 stored-tsdf fractions by 32767)."""
import os
import subprocess

import pytest

from marching_cube_amd import build

pytestmark = pytest.mark.gpu


def test_exact_forms_equal_ieee_division():
    if not os.path.exists(build.EXACT_DIV_BIN):
        pytest.skip("tests/cpp/exact_div_check not built")
    print("[case] exact division forms, all inputs", flush=True)
    res = subprocess.run([build.EXACT_DIV_BIN], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    f = res.stdout.split()
    vals = dict(zip(f[0::2], (int(x) for x in f[1::2])))
    assert vals["pairs"] == 1 << 31
    assert vals["unit_mismatch"] == 0
    assert vals["weight_mismatch"] == 0
    assert vals["fraction_mismatch"] == 0
