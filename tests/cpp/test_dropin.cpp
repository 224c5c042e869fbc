// Drives the C++ drop-in API (MarchingCubes, cu_voxel_to_mesh,
// cu_voxel_rlc_to_mesh, MyMarchingCube, cu_voxel_new +
// _cu_voxel_volumetric_integration, _cu_voxel_extract_depth_and_normal_map,
// cu_voxel_run_length_encode / _decode) on inputs written by
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
  // (0.8, 0.8, -0.5) looking along +z at a wall 1.2 m away; truncation
  // min_t 0.2 m behind the surface, max_t 0.1 m in front of it
  cu_voxel tv;
  tv.data = nullptr;
  cu_voxel_new(&tv, make_int3(16, 16, 16), 0.2f, 0.1f, 0.1f);
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

  // 6. ray-cast the fused volume from the same camera
  cu_array<float> rdepth(32 * 24);
  cu_array<float3> rnormal(32 * 24);
  cu_array<uchar3> rcolor(32 * 24);
  _cu_voxel_extract_depth_and_normal_map(tv, ci, ce, make_int2(32, 24), rdepth.ptr(),
                                         rnormal.ptr(), rcolor.ptr());
  std::vector<float> hd(32 * 24);
  std::vector<float3> hn(32 * 24);
  std::vector<uchar3> hc(32 * 24);
  rdepth.download(hd.data());
  rnormal.download(hn.data());
  rcolor.download(hc.data());
  dump(out + "/ray_depth.bin", hd.data(), hd.size() * sizeof(float));
  dump(out + "/ray_normal.bin", hn.data(), hn.size() * sizeof(float3));
  dump(out + "/ray_color.bin", hc.data(), hc.size() * sizeof(uchar3));
  cu_voxel_delete(&tv);

  // 7. .vol files: the sphere volume whole, loaded back, a region replaced
  // by the original voxels, and a region loaded into a zeroed volume
  const std::string vf = out + "/sphere.vol";
  vol.min_t = 0.03f;
  vol.max_t = 0.04f;
  cu_voxel_run_length_encode(&vol, const_cast<char*>(vf.c_str()), make_int3(0, 0, 0),
                             make_int3(0, 0, 0), make_int3(0, 0, 0));
  cu_voxel lv;
  lv.data = nullptr;
  cu_voxel_run_length_decode(&lv, const_cast<char*>(vf.c_str()), make_int3(0, 0, 0),
                             make_int3(0, 0, 0), make_int3(0, 0, 0));
  std::printf("vol_size %dx%dx%d\n", lv.size.x, lv.size.y, lv.size.z);
  std::vector<cu_voxel_t> hl(n);
  if (hipMemcpy(hl.data(), lv.data, n * sizeof(cu_voxel_t), hipMemcpyDeviceToHost) != hipSuccess) return 3;
  dump(out + "/vol_loaded.bin", hl.data(), n * sizeof(cu_voxel_t));
  cu_voxel_delete(&lv);
  cu_voxel_run_length_encode(&vol, const_cast<char*>(vf.c_str()), make_int3(3, 5, 7),
                             make_int3(20, 10, 30), make_int3(31, 22, 17));
  cu_voxel rv;
  rv.data = nullptr;
  cu_voxel_new(&rv, make_int3(40, 30, 20), 0.0f, 0.0f, h);
  if (hipMemset(rv.data, 0, (size_t)40 * 30 * 20 * sizeof(cu_voxel_t)) != hipSuccess) return 3;
  cu_voxel_run_length_decode(&rv, const_cast<char*>(vf.c_str()), make_int3(1, 2, 3),
                             make_int3(10, 20, 30), make_int3(35, 25, 15));
  std::vector<cu_voxel_t> hr(40 * 30 * 20);
  if (hipMemcpy(hr.data(), rv.data, hr.size() * sizeof(cu_voxel_t), hipMemcpyDeviceToHost) != hipSuccess) return 3;
  dump(out + "/vol_region.bin", hr.data(), hr.size() * sizeof(cu_voxel_t));
  cu_voxel_delete(&rv);
  std::printf("done\n");
  return 0;
}
