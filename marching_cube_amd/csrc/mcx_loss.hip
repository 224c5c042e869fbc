// mcx_loss.hip -- a Chamfer-style vertex loss over a batch of meshes (no
// counterpart in the reference; the loss of the differentiable fusion pass).
//
// Vertices of volume b are [vbase[b], vbase[b+1]) of one concatenated float3
// array; volume b has M target points.  Per vertex: its nearest target t*
// (squared distance from explicit differences, fma(dz, dz, fma(dy, dy,
// dx * dx)), the first index of the minimum), its loss |p - t*|^2 and its gradient
// 2 (p - t*).  Per volume: the sum of its vertices' losses in a fixed order
// (each thread of a workgroup sums a fixed stride, then a fixed tree), so
// the result does not depend on scheduling.
#include <hip/hip_runtime.h>

#include "mcx_internal.h"

namespace mcx {
namespace {

constexpr int kMaxTargets = 1024;  // targets staged in LDS per workgroup

// Volume of vertex v: the last b with vbase[b] <= v (vbase ascending).
__device__ __forceinline__ int volume_of_vertex(const int64_t* vbase, int nvol, int64_t v) {
  int lo = 0, hi = nvol - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (vbase[mid] <= v) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// One thread per vertex row (rows past the total or past cap: gradient 0).
// The targets of the volume of the workgroup's first vertex are staged in
// LDS; a vertex of another volume (a workgroup spanning a boundary) reads
// its targets from memory.
__global__ __launch_bounds__(256) void chamfer_vertex_kernel(const float* pos,
                                                             const int64_t* vbase, int nvol,
                                                             const float* targets, int M,
                                                             int64_t cap, float* gpos,
                                                             float* vloss) {
  __shared__ float s_t[kMaxTargets * 3];
  __shared__ int s_b0;
  const int64_t v0 = (int64_t)blockIdx.x * blockDim.x;
  const int64_t total = vbase[nvol];
  if (threadIdx.x == 0) s_b0 = v0 < total ? volume_of_vertex(vbase, nvol, v0) : 0;
  __syncthreads();
  const int b0 = s_b0;
  for (int i = threadIdx.x; i < 3 * M; i += blockDim.x) s_t[i] = targets[(int64_t)b0 * M * 3 + i];
  __syncthreads();
  const int64_t v = v0 + threadIdx.x;
  if (v >= cap) return;
  if (v >= total) {
    gpos[3 * v] = gpos[3 * v + 1] = gpos[3 * v + 2] = 0.0f;
    return;
  }
  const int b = v < vbase[b0 + 1] ? b0 : volume_of_vertex(vbase, nvol, v);
  const float* t = b == b0 ? s_t : targets + (int64_t)b * M * 3;
  const float px = pos[3 * v], py = pos[3 * v + 1], pz = pos[3 * v + 2];
  float best = __builtin_inff();
  int j = 0;
  for (int m = 0; m < M; ++m) {
    const float dx = px - t[3 * m], dy = py - t[3 * m + 1], dz = pz - t[3 * m + 2];
    const float d2 = __fmaf_rn(dz, dz, __fmaf_rn(dy, dy, dx * dx));
    if (d2 < best) {
      best = d2;
      j = m;
    }
  }
  gpos[3 * v] = 2.0f * (px - t[3 * j]);
  gpos[3 * v + 1] = 2.0f * (py - t[3 * j + 1]);
  gpos[3 * v + 2] = 2.0f * (pz - t[3 * j + 2]);
  vloss[v] = best;
}

// One workgroup per volume: the sum of its vertex losses (vertices below cap).
__global__ __launch_bounds__(256) void chamfer_sum_kernel(const float* vloss,
                                                          const int64_t* vbase, int64_t cap,
                                                          float* loss) {
  __shared__ float s_sum[256];
  const int b = blockIdx.x;
  const int64_t lo = vbase[b], hi = min(vbase[b + 1], cap);
  float acc = 0.0f;
  for (int64_t v = lo + threadIdx.x; v < hi; v += blockDim.x) acc += vloss[v];
  s_sum[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s_sum[threadIdx.x] += s_sum[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[b] = s_sum[0];
}

}  // namespace
}  // namespace mcx

using namespace mcx;

extern "C" size_t mcx_chamfer_workspace_size(int64_t cap) {
  return cap > 0 ? (size_t)cap * sizeof(float) : 0;
}

extern "C" int mcx_chamfer_batched(const float* d_pos, const int64_t* d_vbase, int nvol,
                                   const float* d_targets, int M, int64_t cap, float* d_loss,
                                   float* d_grad_pos, void* d_ws, size_t ws_bytes,
                                   void* stream) {
  if (!d_pos || !d_vbase || nvol < 1 || !d_targets || M < 1 || cap < 1 || !d_loss ||
      !d_grad_pos)
    return MCX_ERR_ARG;
  if (M > kMaxTargets) return MCX_ERR_UNSUPPORTED;
  if (!d_ws || ws_bytes < mcx_chamfer_workspace_size(cap)) return MCX_ERR_WORKSPACE;
  hipStream_t s = as_stream(stream);
  float* vloss = static_cast<float*>(d_ws);
  const int64_t blocks = (cap + 255) / 256;
  if (blocks > kMaxLaunchBlocks) return MCX_ERR_UNSUPPORTED;  // grid in work-items: 32 bits
  hipLaunchKernelGGL(chamfer_vertex_kernel, dim3((unsigned)blocks), dim3(256), 0, s, d_pos,
                     d_vbase, nvol, d_targets, M, cap, d_grad_pos, vloss);
  MCX_HIP(hipGetLastError());
  hipLaunchKernelGGL(chamfer_sum_kernel, dim3((unsigned)nvol), dim3(256), 0, s, vloss, d_vbase,
                     cap, d_loss);
  MCX_HIP(hipGetLastError());
  return MCX_OK;
}
