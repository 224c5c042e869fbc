"""Build libmcx.so (gfx950 HIP kernels + C ABI) in-tree.

    python -m marching_cube_amd.build           # or build_all() from Python

Every HIP source under csrc/ is compiled with hipcc for gfx950 only.  Flags:
  -Xclang -target-feature -Xclang -packed-fp32-ops  no packed-FP32 instructions
      (the device pass honours it; the host pass warns that it does not know
      the feature; check_no_packed_fp32 disassembles the built device code and
      fails the build on any v_pk_add/mul/fma_f32)
  -ffp-contract=off                       every fused multiply-add is explicit
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE fp32 division (as nvcc's default)
"""
import glob
import os
import re
import shutil
import subprocess
import sys
import tempfile

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libmcx.so")
DROPIN_LIB = os.path.join(PKG, "libmarching_cube_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Wall", "-Wno-unused-function",
]


LLVM_BIN = os.environ.get("LLVM_BIN", "/opt/rocm/llvm/bin")
PACKED_FP32 = re.compile(r"\bv_pk_(add|mul|fma)_f32\b")


def check_no_packed_fp32(lib):
    """Disassemble the gfx950 code objects of a library built here and fail
    on any packed FP32 add / multiply / fma (the build disables them with
    -packed-fp32-ops; this checks that the device compiler honoured it).
    Returns the number of instructions checked."""
    with tempfile.TemporaryDirectory() as tmp:
        copy = os.path.join(tmp, os.path.basename(lib))
        shutil.copy(lib, copy)
        # writes one file per bundle entry beside the copy
        subprocess.run([os.path.join(LLVM_BIN, "llvm-objdump"), "--offloading", copy],
                       check=True, capture_output=True, cwd=tmp)
        objs = [f for f in os.listdir(tmp) if "amdgcn" in f and "gfx950" in f]
        if not objs:
            raise RuntimeError(f"no gfx950 code object found in {lib}")
        ninsn = 0
        for f in sorted(objs):
            asm = subprocess.run([os.path.join(LLVM_BIN, "llvm-objdump"), "-d",
                                  os.path.join(tmp, f)], check=True, capture_output=True,
                                 text=True).stdout
            bad = [ln.strip() for ln in asm.splitlines() if PACKED_FP32.search(ln)]
            if bad:
                raise RuntimeError(f"{lib}: packed FP32 instructions in the device code: "
                                   f"{bad[:3]}")
            ninsn += sum(1 for ln in asm.splitlines() if ln.startswith("\t"))
    return ninsn


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_lib(jobs=8):
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for s in srcs:
        o = os.path.join(objdir, os.path.basename(s) + ".o")
        objs.append(o)
        cmd = [HIPCC, *HIP_FLAGS, "-I", INCLUDE, "-I", CSRC, "-c", s, "-o", o]
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise RuntimeError("hipcc failed")
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    n = check_no_packed_fp32(tmp)
    print(f"device code: {n} instructions, no packed FP32 add/mul/fma", flush=True)
    os.replace(tmp, LIB)
    return LIB


def build_dropin():
    """The C++ drop-in host API (MarchingCubes, cu_voxel_to_mesh, ...) above the C ABI."""
    srcs = sorted(glob.glob(os.path.join(CSRC, "dropin", "*.cpp")))
    if not srcs:
        return None
    tmp = DROPIN_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
          "-I", INCLUDE, "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
          "-o", tmp, *srcs, "-L", PKG, "-lmcx", "-Wl,-rpath,$ORIGIN",
          "-L/opt/rocm/lib", "-lamdhip64"])
    os.replace(tmp, DROPIN_LIB)
    return DROPIN_LIB


TEST_BIN = os.path.join(ROOT, "tests", "cpp", "test_dropin")
EXACT_DIV_BIN = os.path.join(ROOT, "tests", "cpp", "exact_div_check")


def build_tests():
    """C++ program that drives the drop-in API (run by tests/test_dropin_gpu.py)."""
    src = os.path.join(ROOT, "tests", "cpp", "test_dropin.cpp")
    if not os.path.exists(src):
        return None
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, "-I", "/opt/rocm/include",
          "-D__HIP_PLATFORM_AMD__", "-o", TEST_BIN, src, "-L", PKG, "-lmarching_cube_amd",
          "-lmcx", "-Wl,-rpath,$ORIGIN/../../marching_cube_amd", "-L/opt/rocm/lib",
          "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
    # device check of the exact fp32 forms (same flags as the library)
    chk = os.path.join(ROOT, "tests", "cpp", "exact_div_check.hip")
    _run([HIPCC, *HIP_FLAGS, "-I", INCLUDE, "-I", CSRC, "-o", EXACT_DIV_BIN, chk])
    return TEST_BIN


def build_all():
    build_lib()
    build_dropin()
    build_tests()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
