"""Host check of the scan kernel's split of the non-empty-tile list over its
workgroups (marching_cube_amd/csrc/mcx_scan_part.h, compiled with g++ from
the same header the kernel includes): every list entry is taken by exactly
one in-range slot, slots ascend within a workgroup, at most four per thread.

The same program counts the (list length, workgroups) pairs in which the
slot *indices* of two workgroups reach the list's last entry (a thread's
slots may run past its workgroup's range).  Publishing the list's totals from
"the thread whose slot index is nz - 1" therefore had two writers in those
pairs, and the emit kernel could read the smaller total and leave the last
entries' vertices unwritten: the list totals now come from the workgroup with
the last ticket.  The count must stay positive, or this check no longer
exercises the hazard."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_scan_list_partition(tmp_path):
    exe = str(tmp_path / "scan_part_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I",
                           os.path.join(ROOT, "marching_cube_amd", "csrc"), "-o", exe,
                           os.path.join(ROOT, "tests", "cpp", "scan_part_check.cpp")])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    tag, pairs, ambiguous = out.stdout.split()
    assert tag == "ok"
    assert int(pairs) > 30000
    assert int(ambiguous) > 0
