// Drives the C++ drop-in API (MarchingCubes, cu_voxel_to_mesh,
// cu_voxel_rlc_to_mesh, MyMarchingCube, cu_voxel_new +
// _cu_voxel_volumetric_integration) on inputs written by
// tests/test_dropin_gpu.py, which checks the files this program writes.
//   test_dropin <volume.bin> X Y Z h <outdir> <rlc.txt> <rlc_dim>
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "marching_cube_amd/MyMarchingCube.h"
#include "marching_cube_amd/cu_kinect_fusion.h"
#include "marching_cube_amd/cu_marching_cubes.h"

static void dump(const std::string& path, const void* p, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    std::perror(path.c_str());
    std::exit(2);
  }
  std::fwrite(p, 1, bytes, f);
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 9) {
    std::fprintf(stderr, "usage: %s volume.bin X Y Z h outdir rlc.txt rlc_dim\n", argv[0]);
    return 2;
  }
  const int X = std::atoi(argv[2]), Y = std::atoi(argv[3]), Z = std::atoi(argv[4]);
  const float h = (float)std::atof(argv[5]);
  const std::string out = argv[6];
  const size_t n = (size_t)X * Y * Z;
  std::vector<cu_voxel_t> host(n);
  FILE* f = std::fopen(argv[1], "rb");
  if (!f || std::fread(host.data(), sizeof(cu_voxel_t), n, f) != n) {
    std::fprintf(stderr, "cannot read %s\n", argv[1]);
    return 2;
  }
  std::fclose(f);

  // 1. MarchingCubes::run with caller-created buffers
  cu_array<cu_voxel_t> dvol;
  dvol.upload(host.data(), n);
  cu_voxel vol;
  vol.size = make_int3(X, Y, Z);
  vol.size_xy = X * Y;
  vol.min_t = vol.max_t = 0.0f;
  vol.grid_s = h;
  vol.data = dvol.ptr();
  MarchingCubes mc;
  cu_array<vertex_voxel> tri;
  cu_array<int3> v2vox;
  cu_array<short> ver_voxel;
  cu_array<vertex_voxel> view = mc.run(vol, tri, v2vox, ver_voxel);
  std::printf("run_count %zu\n", view.size());
  std::printf("run_default_buffer %zu\n", tri.size());
  std::vector<vertex_voxel> hv(view.size());
  view.download(hv.data());
  dump(out + "/run_vv.bin", hv.data(), hv.size() * sizeof(vertex_voxel));

  // 2. cu_voxel_to_mesh on the device volume (PLY and OBJ)
  cu_voxel_to_mesh(reinterpret_cast<short2*>(dvol.ptr()), X, Y, Z, h, (out + "/mesh.ply").c_str());
  cu_voxel_to_mesh(reinterpret_cast<short2*>(dvol.ptr()), X, Y, Z, h, (out + "/mesh.obj").c_str());

  // 3. MyMarchingCube on the host volume
  MyMarchingCube my(host.data(), X, Y, Z, h, (out + "/my.ply").c_str());
  std::printf("my_count %lld\n", my.vertices());

  // 4. cu_voxel_rlc_to_mesh on a text volume
  cu_voxel_rlc_to_mesh(argv[7], const_cast<char*>((out + "/rlc.ply").c_str()), std::atoi(argv[8]));

  // 5. cu_voxel_new + one integrated float depth frame: a camera at
  // (0.8, 0.8, -0.5) looking along +z at a wall 1.2 m away
  cu_voxel tv;
  tv.data = nullptr;
  cu_voxel_new(&tv, make_int3(16, 16, 16), 0.1f, 0.1f, 0.1f);
  float_camera_intrinsic ci{};
  ci.fc = make_float2(40.f, 40.f);
  ci.cc = make_float2(15.5f, 11.5f);
  float_camera_extrinsic ce{};
  ce.R.p[0] = make_float3(1, 0, 0);
  ce.R.p[1] = make_float3(0, 1, 0);
  ce.R.p[2] = make_float3(0, 0, 1);
  ce.t = make_float3(0.8f, 0.8f, -0.5f);
  std::vector<float> depth(32 * 24, 1.2f);
  cu_array<float> ddepth;
  ddepth.upload(depth.data(), depth.size());
  _cu_voxel_volumetric_integration(tv, ci, ce, make_int2(32, 24), ddepth.ptr(), nullptr, nullptr,
                                   false);
  std::vector<cu_voxel_t> fused(16 * 16 * 16);
  if (hipMemcpy(fused.data(), tv.data, fused.size() * sizeof(cu_voxel_t), hipMemcpyDeviceToHost) != hipSuccess) return 3;
  dump(out + "/fused.bin", fused.data(), fused.size() * sizeof(cu_voxel_t));
  cu_voxel_delete(&tv);
  std::printf("done\n");
  return 0;
}
