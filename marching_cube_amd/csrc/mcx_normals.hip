// mcx_normals.hip -- per-vertex surface normals for gfx950.
//
// Replaces d_voxel_extract_surface_normals / _d_voxel_extract_surface_normal
// (cu_kinect_fusion_v1.cu:3146-3238): p_f = ver_pos / grid_s (without the
// +0.5 shift, as the reference, :3241), p_i = (int)p_f, the 2x2x2 corner
// tsdf values at p_i; the normal is zero if any of them is +-32766
// (CU_VOXEL_MAX_TSDF); else the x/y/z differences of bilinear interpolations
// over the two opposite faces (:3201-3218), normalised when the norm exceeds
// FLT_EPSILON (float3_unit, cu_vector_functions.h:258-264).  Where p_i + 1
// leaves the volume (the reference reads out of bounds) the normal is zero.
// One thread per vertex; positions are read with a stride so vertex_voxel
// records (7 words) and packed float3 arrays (3 words) both work.
#include <algorithm>
#include <float.h>
#include <hip/hip_runtime.h>

#include "mcx_internal.h"

namespace mcx {
namespace {

__device__ __forceinline__ float bilinear(float v0, float v1, float v2, float v3, float x, float y) {
  const float a = 1.0f - x, b = 1.0f - y;
  return v0 * a * b + v1 * x * b + v2 * a * y + v3 * x * y;
}

__global__ __launch_bounds__(256) void normals_kernel(const uint32_t* __restrict__ vol, int X, int Y,
                                                      int Z, float h, const float* __restrict__ pos,
                                                      int64_t stride, int64_t n, float* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float pf[3], t[3];
  int pi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    pf[d] = pos[i * stride + d] / h;
    pi[d] = (int)pf[d];
    t[d] = pf[d] - (float)pi[d];
  }
  float nx = 0.f, ny = 0.f, nz = 0.f;
  const bool inb = pf[0] >= 0.f && pf[1] >= 0.f && pf[2] >= 0.f && pi[0] + 1 < X && pi[1] + 1 < Y &&
                   pi[2] + 1 < Z;
  if (inb) {
    const int64_t XY = (int64_t)X * Y;
    const int64_t b0 = ((int64_t)pi[2] * Y + pi[1]) * X + pi[0];
    const uint32_t w000 = vol[b0], w100 = vol[b0 + 1], w110 = vol[b0 + X + 1], w010 = vol[b0 + X];
    const uint32_t w001 = vol[b0 + XY], w101 = vol[b0 + XY + 1], w111 = vol[b0 + XY + X + 1],
                   w011 = vol[b0 + XY + X];
    const uint32_t all[8] = {w000, w100, w110, w010, w001, w101, w111, w011};
    bool bad = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int16_t v = (int16_t)(all[k] & 0xffffu);
      bad = bad || v == 32766 || v == -32766;
    }
    if (!bad) {
      auto f = [](uint32_t w) { return (float)(int16_t)(w & 0xffffu) / 32766.0f; };
      // reference slots: 0 V000, 1 V100, 2 V010, 3 V001, 4 V101, 5 V011, 6 V110, 7 V111
      const float s0 = f(w000), s1 = f(w100), s2 = f(w010), s3 = f(w001);
      const float s4 = f(w101), s5 = f(w011), s6 = f(w110), s7 = f(w111);
      nx = bilinear(s1, s6, s4, s7, t[1], t[2]) - bilinear(s0, s2, s3, s5, t[1], t[2]);
      ny = bilinear(s2, s6, s5, s7, t[0], t[2]) - bilinear(s0, s1, s3, s4, t[0], t[2]);
      nz = bilinear(s3, s4, s5, s7, t[0], t[1]) - bilinear(s0, s1, s2, s6, t[0], t[1]);
      const float nn = sqrtf(nx * nx + ny * ny + nz * nz);
      if (nn > FLT_EPSILON || nn < -FLT_EPSILON) {
        nx = nx / nn;
        ny = ny / nn;
        nz = nz / nn;
      }
    }
  }
  out[3 * i + 0] = nx;
  out[3 * i + 1] = ny;
  out[3 * i + 2] = nz;
}

}  // namespace
}  // namespace mcx

using namespace mcx;

extern "C" int mcx_normals(const void* d_vol, int X, int Y, int Z, float h, const float* d_pos,
                           int64_t stride_floats, int64_t n, float* d_normals, void* stream) {
  if (!d_vol || X < 1 || Y < 1 || Z < 1 || n < 0 || stride_floats < 3) return MCX_ERR_ARG;
  if (n == 0) return MCX_OK;
  if (!d_pos || !d_normals) return MCX_ERR_ARG;
  // a thread per vertex, in launches of at most kMaxLaunchBlocks workgroups
  const int64_t per = kMaxLaunchBlocks * 256;
  for (int64_t v0 = 0; v0 < n; v0 += per) {
    const int64_t nn = std::min(per, n - v0);
    hipLaunchKernelGGL(normals_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0,
                       as_stream(stream), static_cast<const uint32_t*>(d_vol), X, Y, Z, h,
                       d_pos + v0 * stride_floats, stride_floats, nn, d_normals + v0 * 3);
    MCX_HIP(hipGetLastError());
  }
  return MCX_OK;
}
