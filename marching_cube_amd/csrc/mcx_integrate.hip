// mcx_integrate.hip -- TSDF volume init and projective integration for gfx950.
//
// Replaces the reference's dead KinectFusion kernels
//   d_voxel_initialize / cu_voxel_new      cu_kinect_fusion_v1.cu:364-398
//   d_voxel_volumetric_depth_scaling       :416-425
//   d_voxel_volumetric_integration         :427-532
//   _cu_voxel_volumetric_integration       :534-549
// The reference gives each thread one (y, z) row and walks x inside the
// thread, so a wave strides the volume by 4*X bytes (uncoalesced), and it
// sweeps the whole volume once per frame.  Here a thread owns one voxel, a
// workgroup owns 256 consecutive voxels of a z-slab (x fastest, coalesced
// 1 KiB per workgroup), and K frames are fused per sweep: the voxel is read
// once, updated for every frame in order in registers, and written once, so a
// sweep costs 8 B/voxel plus the depth reads instead of 8 B/voxel/frame.
//
// Per voxel and frame (bit-for-bit the oracle's update_voxel):
//   p = ((i + 0.5) h) per axis; q = R^T (p - t)          (camera-to-world pose)
//   skip q.z < FLT_MIN; u = rint(fx q.x / q.z + cx), v likewise (round half even)
//   skip u,v outside [0, W-2) x [0, H-2)                 (reference :449)
//   zval = depth * scale * sqrt(((u-cx)/fx)^2 + ((v-cy)/fy)^2 + 1)  (ray distance)
//   skip zval <= FLT_MIN (dropout); sdf = zval - |q|; skip sdf < -trunc_neg
//   tsdf = sdf > 0 ? min(sdf / trunc, 1) : sdf / trunc_neg (:455-463, min_t
//   and max_t of the volume; one truncation gives clamp(sdf / trunc, -1, 1));
//   running average with integer weights capped at max_w; tsdf stored as
//   trunc(t * 32767) clamped to +-32767.
// A workgroup first marks the frames that none of its voxels can observe
// (frame_skip_flags) and passes over them; the per-voxel arithmetic and its
// results are unchanged.
#include <float.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "mcx_internal.h"

namespace mcx {
namespace {

constexpr int kFramesPerPass = 512;  // poses staged in LDS per launch

__global__ __launch_bounds__(256) void init_kernel(uint32_t* vol, int64_t n) {
  const uint32_t v = (uint32_t)(uint16_t)(int16_t)-32766;  // (-CU_VOXEL_MAX_TSDF, 0)
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) vol[i] = v;
}

// Per-pixel ray-length factor sqrt(((u-cx)/fx)^2 + ((v-cy)/fy)^2 + 1)
// (d_voxel_volumetric_depth_scaling, :416-425), shared by all frames.
__global__ __launch_bounds__(256) void rayfac_kernel(float* out, int W, int H, float fx, float fy,
                                                      float cx, float cy) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W * H) return;
  const int u = i % W, v = i / W;
  const float ax = ((float)u - cx) / fx, ay = ((float)v - cy) / fy;
  out[i] = sqrtf(__fmaf_rn(ay, ay, ax * ax) + 1.0f);
}

struct IntegrateParams {
  uint32_t* vol;
  int X, Y, Z;
  float h, trunc;        // trunc: the positive side (max_t)
  float trunc_neg;       // the negative side (min_t) of the packed integration
  int max_w;
  float fx, fy, cx, cy;
  const float* poses;      // [K][12] device
  const void* depth;       // [K][H][W] device, uint16 or float
  const float* rayfac;     // [H][W]
  float depth_scale;
  int W, H, K;
  FastDiv divX, divY;
  int step_x, step_y, step_z;  // 256 voxels as x, y, z steps (step_x < X, step_y < Y)
  int64_t first_block;   // workgroup index of this launch's blockIdx.x 0
};

// Depth value and ray factor of pixel pix (>= 0) of frame k, as buffer loads
// from uniform per-frame bases with 32-bit offsets (no 64-bit address
// arithmetic per voxel and frame).  Word 3 of a raw 32-bit-element buffer
// resource on gfx950.
constexpr int kRawBufferWord3 = 0x00020000;
template <typename DT>
__device__ __forceinline__ float frame_depth(const IntegrateParams& P, int k, uint32_t pix) {
  const int64_t wh = (int64_t)P.W * P.H;
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<DT*>(static_cast<const DT*>(P.depth) + (int64_t)k * wh), (short)0,
      (int)(wh * (int64_t)sizeof(DT)), kRawBufferWord3);
  if constexpr (sizeof(DT) == 2)
    return (float)(DT)__builtin_amdgcn_raw_buffer_load_b16(rs, pix * 2u, 0, 0);
  else
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, pix * 4u, 0, 0));
}
__device__ __forceinline__ float frame_rayfac(const IntegrateParams& P, uint32_t pix) {
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<float*>(P.rayfac), (short)0, (int)((int64_t)P.W * P.H * 4), kRawBufferWord3);
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, pix * 4u, 0, 0));
}

// Frames that no voxel of a workgroup can observe, as bits of s_skip (one per
// frame).  The workgroup's voxels, linear indices [i0, i1] of one X x Y x Z
// volume, lie in a box of voxel centres; frame k is skipped when the whole
// box is behind its camera, or in front of it (every corner deeper than 1 cm)
// and projecting past one edge of the pixel window: the image of a convex box
// in front of the camera is the convex hull of its corners' images.  Eight
// lanes take a frame, one corner each, with the voxels' own fp32 operations
// (a corner is a voxel centre) up to the quotients, which come from the
// hardware reciprocal (error within 2e-7 |u - cx| plus an ulp).  A voxel's
// fp32 q differs from the exact one by at most d = 4e-7 * (|dx| + |dy| +
// |dz|) per component (the subtraction and the fma chain, |R| <= 1), so its u
// differs from the exact projection by at most (fx d / qz_min) (1 + |u -
// cx|_max / fx), likewise v; the tests use four times that plus one pixel as
// the margin ("behind": every corner below -max(1e-4 m, 4 d)); the pixel
// covers the reciprocal's error wherever a test is close (|u|, |v| < 2^16:
// under 0.02 pixel), so every voxel of a skipped frame would have taken its
// q.z or window skip.  Called by every thread of the workgroup (barriers
// inside).
__device__ void frame_skip_flags(const float* s_pose, int K, const IntegrateParams& P,
                                 int64_t i0, int64_t i1, uint32_t* s_skip) {
  const int64_t XY = (int64_t)P.X * P.Y;
  const int64_t z0 = i0 / XY, z1 = i1 / XY;
  int64_t x0 = 0, x1 = P.X - 1, y0 = 0, y1 = P.Y - 1;
  if (z0 == z1) {
    const int64_t r0 = i0 - z0 * XY, r1 = i1 - z1 * XY;
    y0 = r0 / P.X;
    y1 = r1 / P.X;
    if (y0 == y1) {
      x0 = r0 - y0 * P.X;
      x1 = r1 - y1 * P.X;
    }
  }
  for (int w = threadIdx.x; w < (K + 31) / 32; w += blockDim.x) s_skip[w] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, c = lane & 7;
  const float px = ((float)((c & 1) ? x1 : x0) + 0.5f) * P.h;
  const float py = ((float)((c & 2) ? y1 : y0) + 0.5f) * P.h;
  const float pz = ((float)((c & 4) ? z1 : z0) + 0.5f) * P.h;
  const int nf = (int)(blockDim.x >> 3);  // frames per round
  for (int kb = (int)(threadIdx.x >> 6) * 8; kb < K; kb += nf) {
    const int k = kb + (lane >> 3);
    float qz = 0.0f, u = 0.0f, v = 0.0f, dist = 0.0f;
    if (k < K) {
      const float* R = s_pose + 12 * k;
      const float dx = px - R[9], dy = py - R[10], dz = pz - R[11];
      dist = fabsf(dx) + fabsf(dy) + fabsf(dz);
      const float qx = __fmaf_rn(R[6], dz, __fmaf_rn(R[3], dy, R[0] * dx));
      const float qy = __fmaf_rn(R[7], dz, __fmaf_rn(R[4], dy, R[1] * dx));
      qz = __fmaf_rn(R[8], dz, __fmaf_rn(R[5], dy, R[2] * dx));
      const float rz = __builtin_amdgcn_rcpf(qz);
      u = __fmaf_rn(P.fx, qx * rz, P.cx);
      v = __fmaf_rn(P.fy, qy * rz, P.cy);
    }
    float qzmin = qz, qzmax = qz, umin = u, umax = u, vmin = v, vmax = v, dmax = dist;
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) {
      dmax = fmaxf(dmax, __shfl_xor(dmax, o));
      qzmin = fminf(qzmin, __shfl_xor(qzmin, o));
      qzmax = fmaxf(qzmax, __shfl_xor(qzmax, o));
      umin = fminf(umin, __shfl_xor(umin, o));
      umax = fmaxf(umax, __shfl_xor(umax, o));
      vmin = fminf(vmin, __shfl_xor(vmin, o));
      vmax = fmaxf(vmax, __shfl_xor(vmax, o));
    }
    bool skip = false;
    if (k < K) {
      const float d = 4e-7f * dmax;  // bound on a voxel's fp32 error in q
      if (qzmax < -fmaxf(1e-4f, 4.0f * d)) {
        skip = true;
      } else if (qzmin > 0.01f) {
        // margins from reciprocals (relative error ~1e-7 in a bound taken
        // four times over)
        const float rq = __builtin_amdgcn_rcpf(qzmin);
        const float au = fmaxf(fabsf(umin - P.cx), fabsf(umax - P.cx)) * __builtin_amdgcn_rcpf(P.fx);
        const float av = fmaxf(fabsf(vmin - P.cy), fabsf(vmax - P.cy)) * __builtin_amdgcn_rcpf(P.fy);
        const float tu = 1.0f + 4.0f * (P.fx * d * rq) * (1.0f + au);
        const float tv = 1.0f + 4.0f * (P.fy * d * rq) * (1.0f + av);
        // a voxel is observed only at rint(u) in [0, W - 2), rint(v) in [0, H - 2)
        skip = umax < -0.5f - tu || umin > (float)P.W - 2.5f + tu || vmax < -0.5f - tv ||
               vmin > (float)P.H - 2.5f + tv;
      }
    }
    const uint64_t m = __ballot(skip && c == 0);  // bit 8 j: frame kb + j
    if (lane == 0 && m) {
      uint32_t bits = 0;
#pragma unroll
      for (int j = 0; j < 8; ++j) bits |= (uint32_t)((m >> (8 * j)) & 1u) << j;
      atomicOr(&s_skip[kb >> 5], bits << (kb & 31));
    }
  }
  __syncthreads();
}

// f(k) for every frame k < K not marked in s_skip, in order (a uniform
// 32-frame mask at a time: one LDS read per 32 frames).
template <typename F>
__device__ __forceinline__ void for_active_frames(const uint32_t* s_skip, int K, F&& f) {
  for (int k0 = 0; k0 < K; k0 += 32) {
    uint32_t act = ~s_skip[k0 >> 5];
    if (K - k0 < 32) act &= (1u << (K - k0)) - 1u;
    act = __builtin_amdgcn_readfirstlane(act);
    while (act) {
      const int k = k0 + __builtin_ctz(act);
      act &= act - 1u;
      f(k);
    }
  }
}

// Packed short2 update of voxel i by frames [0, K) (the oracle's
// update_voxel).  The voxel is read at its first update and written only if
// some frame updated it: outside every frustum and behind every surface it
// costs no memory traffic.
// One packed update (the oracle's update_voxel after its lookups): the new
// (tsdf, weight) word from the current one and the observation t.
// A voxel at +1 observed at +1 (free space seen again) stays at +1 with no
// division: told = 32767 / 32767 = 1, fma(1, wi, 1) = wi + 1 and
// (wi + 1) / (wi + 1) = 1 exactly (integers below 2^24), 1 * 32767 = 32767.
__device__ __forceinline__ uint32_t packed_update(uint32_t cur, float tsdf, int max_w) {
  const int16_t ti = (int16_t)(cur & 0xffffu);
  const int wi = (int)(cur >> 16);
  const int wnew = wi + 1 < max_w ? wi + 1 : max_w;
  if (ti == 32767 && tsdf == 1.0f) return 0x7fffu | ((uint32_t)wnew << 16);
  const float told = tsdf_fraction(ti);  // ti / 32767, correctly rounded
  const float tnew = __fmaf_rn(told, (float)wi, tsdf) / (float)(wi + 1);
  int fixedp = (int)(tnew * 32767.0f);
  fixedp = fixedp > 32767 ? 32767 : (fixedp < -32767 ? -32767 : fixedp);
  return (uint32_t)(uint16_t)(int16_t)fixedp | ((uint32_t)wnew << 16);
}

// The truncated observation of a signed distance: sdf / max_t in front of the
// surface (at most 1), sdf / min_t behind it (at least -1 once sdf >= -min_t
// was tested): one division by the selected distance.  At sdf >= max_t (> 0,
// checked by every entry point) the quotient is at least 1 before rounding,
// so at least 1 after it (rounding is monotonic): 1 without the division.
__device__ __forceinline__ float truncated(float sdf, float trunc_pos, float trunc_neg) {
  if (sdf >= trunc_pos) return 1.0f;
  return fminf(sdf / (sdf > 0.0f ? trunc_pos : trunc_neg), 1.0f);
}

// Coordinates of voxel i of an X x Y x Z volume (x fastest).  Below 2^32
// voxels: two 32-bit magic-number divisions (i / X, then the row / Y), exact
// for every 32-bit numerator; larger volumes divide in 64 bits.
__device__ __forceinline__ void voxel_coords(const IntegrateParams& P, int64_t i, int* x, int* y,
                                             int* z) {
  if (i < ((int64_t)1 << 32)) {
    const uint32_t row = P.divX.div((uint32_t)i);
    const uint32_t zz = P.divY.div(row);
    *x = (int)((uint32_t)i - row * (uint32_t)P.X);
    *y = (int)(row - zz * (uint32_t)P.Y);
    *z = (int)zz;
  } else {
    const uint32_t xy = (uint32_t)((int64_t)P.X * P.Y);
    const int zz = (int)(i / xy);
    const uint32_t rem = (uint32_t)(i - (int64_t)zz * xy);
    const uint32_t yy = P.divX.div(rem);
    *x = (int)(rem - yy * (uint32_t)P.X);
    *y = (int)yy;
    *z = zz;
  }
}

// Packed short2 update of voxel i by frames [0, K) (the oracle's
// update_voxel).  The voxel is read at its first update and written only if
// some frame updated it: outside every frustum and behind every surface it
// costs no memory traffic.
template <typename DT>
__device__ __forceinline__ void integrate_voxel(const IntegrateParams& P, const float* s_pose,
                                                const uint32_t* s_skip, float px, float py,
                                                float pz, int64_t i) {
  uint32_t cur = 0;
  bool have = false;
  for_active_frames(s_skip, P.K, [&](int k) {
    const float* R = s_pose + 12 * k;
    const float dx = px - R[9], dy = py - R[10], dz = pz - R[11];
    const float qx = __fmaf_rn(R[6], dz, __fmaf_rn(R[3], dy, R[0] * dx));
    const float qy = __fmaf_rn(R[7], dz, __fmaf_rn(R[4], dy, R[1] * dx));
    const float qz = __fmaf_rn(R[8], dz, __fmaf_rn(R[5], dy, R[2] * dx));
    if (qz < FLT_MIN) return;
    const float u = __fmaf_rn(P.fx, qx / qz, P.cx);
    const float v = __fmaf_rn(P.fy, qy / qz, P.cy);
    if (!(fabsf(u) < 1e9f) || !(fabsf(v) < 1e9f)) return;
    const int ui = (int)rintf(u), vi = (int)rintf(v);
    if (ui < 0 || vi < 0 || ui >= P.W - 2 || vi >= P.H - 2) return;
    const uint32_t pix = (uint32_t)(vi * P.W + ui);
    const float zval = (frame_depth<DT>(P, k, pix) * P.depth_scale) * frame_rayfac(P, pix);
    if (!(zval > FLT_MIN)) return;
    const float qn = sqrtf(__fmaf_rn(qz, qz, __fmaf_rn(qy, qy, qx * qx)));
    const float sdf = zval - qn;
    if (!(sdf >= -P.trunc_neg)) return;
    const float tsdf = truncated(sdf, P.trunc, P.trunc_neg);
    if (!have) {
      cur = P.vol[i];
      have = true;
    }
    cur = packed_update(cur, tsdf, P.max_w);
  });
  if (have) P.vol[i] = cur;
}

// Launches size the shared memory to K frames: poses (12 floats each), then
// one skip bit per frame.
__host__ __device__ constexpr size_t frames_lds_bytes(int K) {
  return (size_t)K * 12 * sizeof(float) + (size_t)((K + 31) / 32) * sizeof(uint32_t);
}

template <typename DT>
__global__ __launch_bounds__(256) void integrate_kernel(IntegrateParams P) {
  extern __shared__ float s_frames[];
  float* const s_pose = s_frames;
  uint32_t* const s_skip = reinterpret_cast<uint32_t*>(s_frames + 12 * P.K);
  for (int i = threadIdx.x; i < P.K * 12; i += blockDim.x) s_pose[i] = P.poses[i];
  __syncthreads();
  const int64_t n = (int64_t)P.X * P.Y * P.Z;
  const int64_t ib = (P.first_block + blockIdx.x) * (int64_t)blockDim.x;
  frame_skip_flags(s_pose, P.K, P, ib, min(ib + (int64_t)blockDim.x, n) - 1, s_skip);
  const int64_t i = ib + threadIdx.x;
  if (i >= n) return;
  int x, y, z;
  voxel_coords(P, i, &x, &y, &z);
  const float px = ((float)x + 0.5f) * P.h;
  const float py = ((float)y + 0.5f) * P.h;
  const float pz = ((float)z + 0.5f) * P.h;
  integrate_voxel<DT>(P, s_pose, s_skip, px, py, pz, i);
}

// One frame (the reference's per-frame call): eight voxels a thread (voxel
// ib + t + 256 j of the workgroup's 2048, coalesced).  The thread's first
// voxel has its coordinates from two divisions; each next one lies 256
// voxels further, (step_x, step_y, step_z) with carries.  The projections
// come first, then the depth, ray-factor and voxel loads of all eight are
// issued together (at clamped, valid addresses), then the updates: the
// operations and results of integrate_voxel, with the memory latency
// overlapped.  The voxels are addressed from the workgroup's first one
// (32-bit offsets).  Only voxels the frame updates are written.
// 0.422 ms per 640x480 frame into 512^3 against 0.435 ms with four voxels a
// thread, each by two divisions (four with steps: 0.445; eight by divisions:
// 0.434; sixteen: 0.422 with steps, 0.444 by divisions).
constexpr int kVoxelsPerThread1 = 8;
template <typename DT>
__global__ __launch_bounds__(256) void integrate1_kernel(IntegrateParams P) {
  constexpr int NV = kVoxelsPerThread1;
  __shared__ float s_pose[12];
  __shared__ uint32_t s_skip[1];
  if (threadIdx.x < 12) s_pose[threadIdx.x] = P.poses[threadIdx.x];
  __syncthreads();
  const int64_t n = (int64_t)P.X * P.Y * P.Z;
  const int64_t ib = (P.first_block + blockIdx.x) * 256 * NV;
  const int nloc = (int)min((int64_t)256 * NV, n - ib);  // voxels of the workgroup
  frame_skip_flags(s_pose, 1, P, ib, ib + nloc - 1, s_skip);
  if (s_skip[0] & 1u) return;  // no voxel of the workgroup can see the frame
  const float* R = s_pose;
  uint32_t* const vb = P.vol + ib;  // voxel ib + l is vb[l]
  const int t = (int)threadIdx.x;
  int x, y, z;
  voxel_coords(P, ib + min(t, nloc - 1), &x, &y, &z);
  int pix[NV];
  float qn[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    if (j > 0) {  // x + step_x < 2X, y + step_y + 1 < 2Y
      x += P.step_x;
      y += P.step_y;
      z += P.step_z;
      if (x >= P.X) {
        x -= P.X;
        ++y;
      }
      if (y >= P.Y) {
        y -= P.Y;
        ++z;
      }
    }
    pix[j] = -1;
    qn[j] = 0.0f;
    if (t + 256 * j >= nloc) continue;
    const float px = ((float)x + 0.5f) * P.h;
    const float py = ((float)y + 0.5f) * P.h;
    const float pz = ((float)z + 0.5f) * P.h;
    const float dx = px - R[9], dy = py - R[10], dz = pz - R[11];
    const float qx = __fmaf_rn(R[6], dz, __fmaf_rn(R[3], dy, R[0] * dx));
    const float qy = __fmaf_rn(R[7], dz, __fmaf_rn(R[4], dy, R[1] * dx));
    const float qz = __fmaf_rn(R[8], dz, __fmaf_rn(R[5], dy, R[2] * dx));
    if (qz < FLT_MIN) continue;
    const float u = __fmaf_rn(P.fx, qx / qz, P.cx);
    const float v = __fmaf_rn(P.fy, qy / qz, P.cy);
    if (!(fabsf(u) < 1e9f) || !(fabsf(v) < 1e9f)) continue;
    const int ui = (int)rintf(u), vv = (int)rintf(v);
    if (ui < 0 || vv < 0 || ui >= P.W - 2 || vv >= P.H - 2) continue;
    pix[j] = vv * P.W + ui;
    qn[j] = sqrtf(__fmaf_rn(qz, qz, __fmaf_rn(qy, qy, qx * qx)));
  }
  DT dm[NV];
  float rf[NV];
  uint32_t cur[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const int q = pix[j] >= 0 ? pix[j] : 0;
    dm[j] = static_cast<const DT*>(P.depth)[(uint32_t)q];
    rf[j] = P.rayfac[(uint32_t)q];
    cur[j] = vb[min(t + 256 * j, nloc - 1)];
  }
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    if (pix[j] < 0) continue;
    const float zval = ((float)dm[j] * P.depth_scale) * rf[j];
    if (!(zval > FLT_MIN)) continue;
    const float sdf = zval - qn[j];
    if (!(sdf >= -P.trunc_neg)) continue;
    vb[t + 256 * j] = packed_update(cur[j], truncated(sdf, P.trunc, P.trunc_neg), P.max_w);
  }
}


// ---- fp32 TSDF volumes (differentiable form; no counterpart in the reference)
// tsdf (fp32) and weight (int32) in separate [z][y][x] arrays; the same
// observation as above, stored unpacked:
//   f <- fma(f, w, t) / (w + 1),  w <- min(w + 1, max_w),  t = clamp(sdf / trunc, -1, 1).
// Observation k of voxel (px, py, pz): pixel, t and dt/d(depth value) (zero
// where t is clamped).  Returns false where the voxel is not updated.
template <typename DT>
__device__ __forceinline__ bool observe(const IntegrateParams& P, const float* R, int k,
                                        float px, float py, float pz, int* pix_out, float* t_out,
                                        float* dtdd_out) {
  const float dx = px - R[9], dy = py - R[10], dz = pz - R[11];
  const float qx = __fmaf_rn(R[6], dz, __fmaf_rn(R[3], dy, R[0] * dx));
  const float qy = __fmaf_rn(R[7], dz, __fmaf_rn(R[4], dy, R[1] * dx));
  const float qz = __fmaf_rn(R[8], dz, __fmaf_rn(R[5], dy, R[2] * dx));
  if (qz < FLT_MIN) return false;
  const float u = __fmaf_rn(P.fx, qx / qz, P.cx);
  const float v = __fmaf_rn(P.fy, qy / qz, P.cy);
  if (!(fabsf(u) < 1e9f) || !(fabsf(v) < 1e9f)) return false;
  const int ui = (int)rintf(u), vi = (int)rintf(v);
  if (ui < 0 || vi < 0 || ui >= P.W - 2 || vi >= P.H - 2) return false;
  const int pix = vi * P.W + ui;
  const float rf = frame_rayfac(P, (uint32_t)pix);
  const float d = frame_depth<DT>(P, k, (uint32_t)pix) * P.depth_scale;
  const float zval = d * rf;
  if (!(zval > FLT_MIN)) return false;
  const float qn = sqrtf(__fmaf_rn(qz, qz, __fmaf_rn(qy, qy, qx * qx)));
  const float sdf = zval - qn;
  if (!(sdf >= -P.trunc)) return false;
  *pix_out = pix;
  if (sdf >= P.trunc) {
    // sdf / trunc >= 1 before rounding, so after it: clamped to 1, no gradient
    *t_out = 1.0f;
    *dtdd_out = 0.0f;
    return true;
  }
  const float ts = sdf / P.trunc;
  *t_out = fminf(fmaxf(ts, -1.0f), 1.0f);
  *dtdd_out = (ts > -1.0f && ts < 1.0f) ? P.depth_scale * rf / P.trunc : 0.0f;
  return true;
}

// fp32 update of one voxel by frames [0, K).  With tsdf / wvol, (f, w) is
// read from them at the first update (none: no memory traffic) and *changed
// says whether any frame updated it; without, (f, w) holds the start value.
__device__ __forceinline__ void integrate_tsdf_voxel(const IntegrateParams& P, const float* s_pose,
                                                     const uint32_t* s_skip, float px, float py,
                                                     float pz, float& f, int& w,
                                                     const float* tsdf, const int32_t* wvol,
                                                     bool* changed) {
  bool have = tsdf == nullptr;
  for_active_frames(s_skip, P.K, [&](int k) {
    int pix;
    float t, dtdd;
    if (!observe<float>(P, s_pose + 12 * k, k, px, py, pz, &pix, &t, &dtdd)) return;
    if (!have) {
      f = *tsdf;
      w = *wvol;
      have = true;
    }
    *changed = true;
    // f = 1 observed at 1 stays 1: fma(1, w, 1) = w + 1 = (float)(w + 1)
    // exactly for 0 <= w < 2^24 (a weight read from the caller's volume may
    // lie outside), and the quotient of equal values is 1
    if (!(f == 1.0f && t == 1.0f && (uint32_t)w < (1u << 24)))
      f = __fmaf_rn(f, (float)w, t) / (float)(w + 1);
    w = w + 1 < P.max_w ? w + 1 : P.max_w;
  });
}

__device__ __forceinline__ void voxel_centre(const IntegrateParams& P, int64_t i, float* px,
                                             float* py, float* pz) {
  int x, y, z;
  voxel_coords(P, i, &x, &y, &z);
  *px = ((float)x + 0.5f) * P.h;
  *py = ((float)y + 0.5f) * P.h;
  *pz = ((float)z + 0.5f) * P.h;
}

__global__ __launch_bounds__(256) void integrate_tsdf_kernel(IntegrateParams P, float* tsdf,
                                                             int32_t* wvol) {
  extern __shared__ float s_frames[];
  float* const s_pose = s_frames;
  uint32_t* const s_skip = reinterpret_cast<uint32_t*>(s_frames + 12 * P.K);
  for (int i = threadIdx.x; i < P.K * 12; i += blockDim.x) s_pose[i] = P.poses[i];
  __syncthreads();
  const int64_t n = (int64_t)P.X * P.Y * P.Z;
  const int64_t ib = (P.first_block + blockIdx.x) * (int64_t)blockDim.x;
  frame_skip_flags(s_pose, P.K, P, ib, min(ib + (int64_t)blockDim.x, n) - 1, s_skip);
  const int64_t i = ib + threadIdx.x;
  if (i >= n) return;
  float px, py, pz;
  voxel_centre(P, i, &px, &py, &pz);
  float f = 0.0f;
  int w = 0;
  bool changed = false;
  integrate_tsdf_voxel(P, s_pose, s_skip, px, py, pz, f, w, tsdf + i, wvol + i, &changed);
  if (changed) {
    tsdf[i] = f;
    wvol[i] = w;
  }
}

// Deterministic max of |g| over n floats: per-block max, then an integer
// atomicMax on the bit pattern (order-preserving for non-negative floats).
__global__ __launch_bounds__(256) void maxabs_kernel(const float* g, int64_t n, uint32_t* out) {
  __shared__ uint32_t s_m[4];
  uint32_t m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    m = max(m, __float_as_uint(g[i]) & 0x7fffffffu);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o));
  if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(out, max(max(s_m[0], s_m[1]), max(s_m[2], s_m[3])));
}

// Fixed-point scale of the depth-gradient accumulators: contributions are
// bounded by max|g| * max(dt/dd); at most 2^20 of them reach one pixel, so
// scale * bound * 2^20 <= 2^62 keeps every int64 sum exact.
__device__ __forceinline__ double grad_scale(uint32_t maxg_bits, float max_dtdd) {
  const double bound = (double)__uint_as_float(maxg_bits) * (double)max_dtdd;
  return bound > 0.0 ? 4398046511104.0 / bound : 1.0;  // 2^42 / bound
}

// Backward of the fp32 update of one voxel for the K frames of one call:
// recompute its observations; observation i (weight before it wb_i =
// min(w0 + i, max_w)) has df/dt_i = 1/(wb_i + 1) * prod over later
// observations j of wb_j / (wb_j + 1) (closed form: the uncapped part
// telescopes, the capped part is a power of max_w / (max_w + 1)); the
// contribution g * df/dt_i * dt_i/dd goes to the observed pixel of acc (K
// frames of W x H) as an int64 fixed-point atomic add (integer sums do not
// depend on the order).
__device__ __forceinline__ void backward_voxel(const IntegrateParams& P, const float* s_pose,
                                               const uint32_t* s_skip, float px, float py,
                                               float pz, float gi, int w0, double scale,
                                               unsigned long long* acc) {
  const int M = P.max_w;
  int nobs = 0;
  for_active_frames(s_skip, P.K, [&](int k) {
    int pix;
    float t, dtdd;
    if (observe<float>(P, s_pose + 12 * k, k, px, py, pz, &pix, &t, &dtdd)) ++nobs;
  });
  const int jc = M - w0 > 0 ? M - w0 : 0;  // first observation that sees the cap
  const double cap_a = (double)M / (double)(M + 1);
  int o = 0;
  for_active_frames(s_skip, P.K, [&](int k) {
    int pix;
    float t, dtdd;
    if (!observe<float>(P, s_pose + 12 * k, k, px, py, pz, &pix, &t, &dtdd)) return;
    const int wb = w0 + o < M ? w0 + o : M;
    double c = 1.0 / (double)(wb + 1);
    const int last_unc = (nobs - 1 < jc - 1) ? nobs - 1 : jc - 1;  // last uncapped index
    if (o + 1 <= last_unc) c *= (double)(w0 + o + 1) / (double)(w0 + last_unc + 1);
    const int first_cap = o + 1 > jc ? o + 1 : jc;
    const int ncap = nobs - first_cap;
    if (ncap > 0) {
      // cap_a^ncap by squaring (a library pow inlined here costs the kernel
      // registers on every path, capped or not)
      double pw = 1.0, base = cap_a;
      for (int e = ncap; e; e >>= 1) {
        if (e & 1) pw *= base;
        base *= base;
      }
      c *= pw;
    }
    const double contrib = (double)gi * c * (double)dtdd;
    if (contrib != 0.0) {
      const long long q = llrint(contrib * scale);
      atomicAdd(acc + (int64_t)k * P.W * P.H + pix, (unsigned long long)q);
    }
    ++o;
  });
}

__global__ __launch_bounds__(256) void integrate_tsdf_backward_kernel(
    IntegrateParams P, const float* __restrict__ g, const int32_t* __restrict__ w0vol,
    const uint32_t* maxg, float max_dtdd, unsigned long long* acc) {
  extern __shared__ float s_frames[];
  float* const s_pose = s_frames;
  uint32_t* const s_skip = reinterpret_cast<uint32_t*>(s_frames + 12 * P.K);
  const int64_t n = (int64_t)P.X * P.Y * P.Z;
  const int64_t ib = (P.first_block + blockIdx.x) * (int64_t)blockDim.x;
  const int64_t i = ib + threadIdx.x;
  // most blocks carry no gradient: they leave after reading it, before the
  // poses and the frame tests
  const float gi = i < n ? g[i] : 0.0f;
  if (!__syncthreads_or(gi != 0.0f)) return;
  for (int k = threadIdx.x; k < P.K * 12; k += blockDim.x) s_pose[k] = P.poses[k];
  __syncthreads();
  frame_skip_flags(s_pose, P.K, P, ib, min(ib + (int64_t)blockDim.x, n) - 1, s_skip);
  if (i >= n || gi == 0.0f) return;
  float px, py, pz;
  voxel_centre(P, i, &px, &py, &pz);
  backward_voxel(P, s_pose, s_skip, px, py, pz, gi, w0vol ? w0vol[i] : 0,
                 grad_scale(*maxg, max_dtdd), acc);
}

// ---- batches of fp32 volumes (the per-volume layout of a batch plan:
// concatenated per-voxel arrays, K frames of W x H and K poses per volume,
// one set of intrinsics); one workgroup = 256 voxels of one volume.
struct BatchIntegrateParams {
  IntegrateParams base;  // intrinsics, truncation, frames; depth / poses of volume 0
  const BatchVol* bvol;
  int nvol;
  int from_init;         // start every voxel from (init_value, 0) without reading it
  float init_value;
  int64_t first_block;   // workgroup index of this launch's blockIdx.x 0
};

// The volume of this workgroup's block (binary search over the block ranges).
__device__ __forceinline__ int batch_block_volume(const BatchIntegrateParams& B) {
  return batch_volume_of<&BatchVol::blk_begin>(B.bvol, B.nvol, B.first_block + blockIdx.x);
}

// This workgroup's volume vi: its launch parameters and its poses in LDS.
__device__ __forceinline__ int batch_setup(const BatchIntegrateParams& B, int vi, IntegrateParams& P,
                                           float* s_pose, int64_t* vox_begin, int64_t* i0) {
  const int64_t blk = B.first_block + blockIdx.x;
  const BatchVol& V = B.bvol[vi];
  P = B.base;
  P.X = V.X;
  P.Y = V.Y;
  P.Z = V.Z;
  P.h = V.h;
  P.divX = V.divX;
  P.divY = V.divY;
  const int64_t frames = (int64_t)P.K * P.W * P.H;
  P.depth = static_cast<const float*>(B.base.depth) + (int64_t)vi * frames;
  const float* poses = B.base.poses + (int64_t)vi * P.K * 12;
  for (int i = threadIdx.x; i < P.K * 12; i += blockDim.x) s_pose[i] = poses[i];
  *vox_begin = V.vox_begin;
  *i0 = (blk - V.blk_begin) * 256;
  __syncthreads();
  return vi;
}

__global__ __launch_bounds__(256) void integrate_tsdf_batched_kernel(BatchIntegrateParams B,
                                                                     float* tsdf, int32_t* wvol) {
  extern __shared__ float s_frames[];
  float* const s_pose = s_frames;
  uint32_t* const s_skip = reinterpret_cast<uint32_t*>(s_frames + 12 * B.base.K);
  IntegrateParams P;
  int64_t vb, ib;
  batch_setup(B, batch_block_volume(B), P, s_pose, &vb, &ib);
  const int64_t n = (int64_t)P.X * P.Y * P.Z;
  frame_skip_flags(s_pose, P.K, P, ib, min(ib + 256, n) - 1, s_skip);
  const int64_t i = ib + threadIdx.x;
  if (i >= n) return;
  float px, py, pz;
  voxel_centre(P, i, &px, &py, &pz);
  float f = B.init_value;
  int w = 0;
  bool changed = false;
  if (B.from_init) {
    // every voxel is written: the start value where no frame observes it
    integrate_tsdf_voxel(P, s_pose, s_skip, px, py, pz, f, w, nullptr, nullptr, &changed);
    tsdf[vb + i] = f;
    wvol[vb + i] = w;
  } else {
    integrate_tsdf_voxel(P, s_pose, s_skip, px, py, pz, f, w, tsdf + vb + i, wvol + vb + i,
                         &changed);
    if (changed) {
      tsdf[vb + i] = f;
      wvol[vb + i] = w;
    }
  }
}

__global__ __launch_bounds__(256) void integrate_tsdf_backward_batched_kernel(
    BatchIntegrateParams B, const float* __restrict__ g, const int32_t* __restrict__ w0vol,
    const uint32_t* maxg, float max_dtdd, unsigned long long* acc) {
  extern __shared__ float s_frames[];
  float* const s_pose = s_frames;
  uint32_t* const s_skip = reinterpret_cast<uint32_t*>(s_frames + 12 * B.base.K);
  // most blocks carry no gradient: they leave after reading it, before the
  // poses and the frame tests
  const int vi = batch_block_volume(B);
  float gi = 0.0f;
  {
    const BatchVol& V = B.bvol[vi];
    const int64_t i = (B.first_block + blockIdx.x - V.blk_begin) * 256 + threadIdx.x;
    if (i < (int64_t)V.X * V.Y * V.Z) gi = g[V.vox_begin + i];
  }
  if (!__syncthreads_or(gi != 0.0f)) return;
  IntegrateParams P;
  int64_t vb, ib;
  batch_setup(B, vi, P, s_pose, &vb, &ib);
  const int64_t n = (int64_t)P.X * P.Y * P.Z;
  frame_skip_flags(s_pose, P.K, P, ib, min(ib + 256, n) - 1, s_skip);
  const int64_t i = ib + threadIdx.x;
  if (i >= n || gi == 0.0f) return;
  float px, py, pz;
  voxel_centre(P, i, &px, &py, &pz);
  backward_voxel(P, s_pose, s_skip, px, py, pz, gi, w0vol ? w0vol[vb + i] : 0,
                 grad_scale(*maxg, max_dtdd), acc + (int64_t)vi * P.K * P.W * P.H);
}

__global__ __launch_bounds__(256) void fixed_to_float_kernel(const unsigned long long* acc,
                                                             int64_t n, const uint32_t* maxg,
                                                             float max_dtdd, float* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double scale = grad_scale(*maxg, max_dtdd);
  out[i] = (float)((double)(long long)acc[i] / scale);
}

__global__ __launch_bounds__(256) void fill_f32_i32_kernel(float* f, int32_t* w, int64_t n,
                                                           float fv, int32_t wv) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (f) f[i] = fv;
    if (w) w[i] = wv;
  }
}

// ---- synthetic depth frames: ray-cast spheres and axis-aligned boxes -------
__device__ __forceinline__ uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}
__device__ __forceinline__ float u01(uint32_t h) { return ((h >> 8) + 0.5f) * (1.0f / 16777216.0f); }

__global__ __launch_bounds__(256) void synth_kernel(uint16_t* depth, int W, int H, int K, float fx,
                                                     float fy, float cx, float cy,
                                                     const float* poses, const float* objs,
                                                     int nobj, float sigma_m, float dropout,
                                                     uint32_t seed, float depth_unit) {
  const int64_t n = (int64_t)W * H * K;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = (int)(i / ((int64_t)W * H));
  const int pix = (int)(i - (int64_t)k * W * H);
  const int u = pix % W, v = pix / W;
  const float* R = poses + 12 * k;
  const float a = ((float)u - cx) / fx, b = ((float)v - cy) / fy;
  // world direction of the pixel ray with unit camera-z component
  const float dx = R[0] * a + R[1] * b + R[2];
  const float dy = R[3] * a + R[4] * b + R[5];
  const float dz = R[6] * a + R[7] * b + R[8];
  const float ox = R[9], oy = R[10], oz = R[11];
  float best = 1e30f;
  for (int j = 0; j < nobj; ++j) {
    const float* o = objs + 8 * j;
    if (o[0] < 0.5f) {  // sphere
      const float lx = ox - o[1], ly = oy - o[2], lz = oz - o[3];
      const float A = dx * dx + dy * dy + dz * dz;
      const float B = lx * dx + ly * dy + lz * dz;
      const float C = lx * lx + ly * ly + lz * lz - o[4] * o[4];
      const float disc = B * B - A * C;
      if (disc >= 0.f) {
        const float sq = sqrtf(disc);
        float s = (-B - sq) / A;
        if (s <= 1e-4f) s = (-B + sq) / A;
        if (s > 1e-4f && s < best) best = s;
      }
    } else {  // axis-aligned box
      float tmin = -1e30f, tmax = 1e30f;
      const float org[3] = {ox, oy, oz}, dir[3] = {dx, dy, dz};
      bool miss = false;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float lo = o[1 + c] - o[4 + c], hi = o[1 + c] + o[4 + c];
        if (fabsf(dir[c]) < 1e-12f) {
          if (org[c] < lo || org[c] > hi) miss = true;
        } else {
          float t0 = (lo - org[c]) / dir[c], t1 = (hi - org[c]) / dir[c];
          if (t0 > t1) {
            const float tt = t0;
            t0 = t1;
            t1 = tt;
          }
          tmin = fmaxf(tmin, t0);
          tmax = fminf(tmax, t1);
        }
      }
      if (!miss && tmax >= tmin) {
        const float s = tmin > 1e-4f ? tmin : tmax;
        if (s > 1e-4f && s < best) best = s;
      }
    }
  }
  uint16_t out = 0;
  if (best < 1e29f) {
    const uint32_t h0 = hash32(seed ^ hash32((uint32_t)i * 2u + 1u));
    const uint32_t h1 = hash32(h0 + 0x9e3779b9u);
    const uint32_t h2 = hash32(h1 + 0x9e3779b9u);
    const float r = sqrtf(-2.0f * logf(u01(h0)));
    const float gauss = r * cosf(6.28318530718f * u01(h1));
    const float mm = rintf((best + sigma_m * gauss) / depth_unit);
    if (u01(h2) >= dropout && mm >= 1.f) out = (uint16_t)fminf(mm, 65535.f);
  }
  depth[i] = out;
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace
}  // namespace mcx

using namespace mcx;

extern "C" int mcx_volume_init(void* d_vol, int X, int Y, int Z, void* stream) {
  if (!d_vol || X < 1 || Y < 1 || Z < 1) return MCX_ERR_ARG;
  const int64_t n = (int64_t)X * Y * Z;
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  hipLaunchKernelGGL(init_kernel, dim3((unsigned)b), dim3(256), 0, as_stream(stream),
                     static_cast<uint32_t*>(d_vol), n);
  MCX_HIP(hipGetLastError());
  return MCX_OK;
}

extern "C" size_t mcx_integrate_workspace_size(int W, int H) {
  return (size_t)(W > 0 ? W : 0) * (size_t)(H > 0 ? H : 0) * sizeof(float);
}

// Checks shared by the integration entry points; fills the launch parameters
// (all but vol, K, poses and depth) and computes the per-pixel ray factors.
int prepare_integrate(IntegrateParams& P, int X, int Y, int Z, float h, float trunc, int max_w,
                      const float* intr, float depth_scale, int W, int H, int K, float* rayfac,
                      hipStream_t s) {
  if (!intr || X < 1 || Y < 1 || Z < 1 || W < 3 || H < 3 || K < 0) return MCX_ERR_ARG;
  if (!(h > 0.f) || !(trunc > 0.f) || max_w < 1 || max_w > 65535) return MCX_ERR_ARG;
  if ((int64_t)X * Y > 0xffffffffll) return MCX_ERR_UNSUPPORTED;
  // frames are read through buffer resources of at most 2^31 - 1 bytes
  if ((int64_t)W * H * 4 >= ((int64_t)1 << 31)) return MCX_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(rayfac_kernel, dim3(blocks_for((int64_t)W * H)), dim3(256), 0, s, rayfac, W,
                     H, intr[0], intr[1], intr[2], intr[3]);
  MCX_HIP(hipGetLastError());
  P.vol = nullptr;
  P.X = X;
  P.Y = Y;
  P.Z = Z;
  P.h = h;
  P.trunc = trunc;
  P.trunc_neg = trunc;
  P.max_w = max_w;
  P.fx = intr[0];
  P.fy = intr[1];
  P.cx = intr[2];
  P.cy = intr[3];
  P.rayfac = rayfac;
  P.depth_scale = depth_scale;
  P.W = W;
  P.H = H;
  P.divX = FastDiv((uint32_t)X);
  P.divY = FastDiv((uint32_t)Y);
  {
    const int64_t XY = (int64_t)X * Y, r = 256 % XY;
    P.step_z = (int)(256 / XY);
    P.step_y = (int)(r / X);
    P.step_x = (int)(r % X);
  }
  P.K = 0;
  P.poses = nullptr;
  P.depth = nullptr;
  return MCX_OK;
}

template <typename DT>
int integrate_impl(void* d_vol, int X, int Y, int Z, float h, float trunc_neg, float trunc,
                   int max_w, const float* intr, const float* d_poses, const DT* d_depth,
                   float depth_scale, int W, int H, int K, void* d_ws, size_t ws_bytes,
                   void* stream) {
  if (!d_vol || !(trunc_neg > 0.f)) return MCX_ERR_ARG;
  if (K > 0 && (!d_poses || !d_depth)) return MCX_ERR_ARG;
  if (!d_ws || ws_bytes < mcx_integrate_workspace_size(W, H)) return MCX_ERR_WORKSPACE;
  if (K == 0) return MCX_OK;
  hipStream_t s = as_stream(stream);
  IntegrateParams P;
  int rc = prepare_integrate(P, X, Y, Z, h, trunc, max_w, intr, depth_scale, W, H, K,
                             static_cast<float*>(d_ws), s);
  if (rc) return rc;
  P.vol = static_cast<uint32_t*>(d_vol);
  P.trunc_neg = trunc_neg;
  const int64_t n = (int64_t)X * Y * Z;
  for (int k0 = 0; k0 < K; k0 += kFramesPerPass) {
    P.K = K - k0 < kFramesPerPass ? K - k0 : kFramesPerPass;
    P.poses = d_poses + (int64_t)12 * k0;
    P.depth = d_depth + (int64_t)k0 * W * H;
    const int64_t per = P.K == 1 ? 256 * kVoxelsPerThread1 : 256;
    const int64_t nb = (n + per - 1) / per;
    for (int64_t b0 = 0; b0 < nb; b0 += kMaxLaunchBlocks) {
      P.first_block = b0;
      const dim3 grid((unsigned)std::min(kMaxLaunchBlocks, nb - b0));
      if (P.K == 1)
        hipLaunchKernelGGL(integrate1_kernel<DT>, grid, dim3(256), 0, s, P);
      else
        hipLaunchKernelGGL(integrate_kernel<DT>, grid, dim3(256), frames_lds_bytes(P.K), s, P);
      MCX_HIP(hipGetLastError());
    }
  }
  return MCX_OK;
}

// Largest dt/d(depth value) of any pixel in the window (the ray factor grows
// towards the corners), with a margin: the bound of the fixed-point scale.
float max_dtdd(const float* intr, float depth_scale, float trunc, int W, int H) {
  double m = 0.0;
  const int us[2] = {0, W - 3}, vs[2] = {0, H - 3};
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      const double ax = ((double)us[a] - intr[2]) / intr[0], ay = ((double)vs[b] - intr[3]) / intr[1];
      m = std::max(m, std::sqrt(ax * ax + ay * ay + 1.0));
    }
  return (float)(1.01 * m * (double)depth_scale / (double)trunc);
}

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

extern "C" int mcx_integrate(void* d_vol, int X, int Y, int Z, float h, float trunc, int max_w,
                             const float* intr, const float* d_poses, const uint16_t* d_depth,
                             float depth_scale, int W, int H, int K, void* d_ws, size_t ws_bytes,
                             void* stream) {
  return integrate_impl<uint16_t>(d_vol, X, Y, Z, h, trunc, trunc, max_w, intr, d_poses, d_depth,
                                  depth_scale, W, H, K, d_ws, ws_bytes, stream);
}

extern "C" int mcx_integrate_f32(void* d_vol, int X, int Y, int Z, float h, float trunc,
                                 int max_w, const float* intr, const float* d_poses,
                                 const float* d_depth, float depth_scale, int W, int H, int K,
                                 void* d_ws, size_t ws_bytes, void* stream) {
  return integrate_impl<float>(d_vol, X, Y, Z, h, trunc, trunc, max_w, intr, d_poses, d_depth,
                               depth_scale, W, H, K, d_ws, ws_bytes, stream);
}

extern "C" int mcx_integrate_ex(void* d_vol, int X, int Y, int Z, float h, float trunc_neg,
                                float trunc_pos, int max_w, const float* intr,
                                const float* d_poses, const void* d_depth, int depth_kind,
                                float depth_scale, int W, int H, int K, void* d_ws,
                                size_t ws_bytes, void* stream) {
  if (depth_kind == MCX_DEPTH_U16)
    return integrate_impl<uint16_t>(d_vol, X, Y, Z, h, trunc_neg, trunc_pos, max_w, intr,
                                    d_poses, static_cast<const uint16_t*>(d_depth), depth_scale,
                                    W, H, K, d_ws, ws_bytes, stream);
  if (depth_kind == MCX_DEPTH_F32)
    return integrate_impl<float>(d_vol, X, Y, Z, h, trunc_neg, trunc_pos, max_w, intr, d_poses,
                                 static_cast<const float*>(d_depth), depth_scale, W, H, K, d_ws,
                                 ws_bytes, stream);
  return MCX_ERR_ARG;
}

extern "C" int mcx_synth_depth(uint16_t* d_depth, int W, int H, int K, const float* intr,
                               const float* d_poses, const float* d_objs, int nobj,
                               float noise_sigma_m, float dropout, uint32_t seed,
                               float depth_unit_m, void* stream) {
  if (!d_depth || !intr || !d_poses || (nobj > 0 && !d_objs) || W < 1 || H < 1 || K < 1 ||
      !(depth_unit_m > 0.f))
    return MCX_ERR_ARG;
  const int64_t n = (int64_t)W * H * K;
  hipLaunchKernelGGL(synth_kernel, dim3(blocks_for(n)), dim3(256), 0, as_stream(stream), d_depth,
                     W, H, K, intr[0], intr[1], intr[2], intr[3], d_poses, d_objs, nobj,
                     noise_sigma_m, dropout, seed, depth_unit_m);
  MCX_HIP(hipGetLastError());
  return MCX_OK;
}

// ---- fp32 TSDF volumes: differentiable integration (no counterpart in the
//      reference; the forward is the same observation, stored unpacked)
extern "C" int mcx_tsdf_init(float* d_tsdf, int32_t* d_weight, int64_t n, float tsdf_value,
                             void* stream) {
  if (n < 0 || (n > 0 && !d_tsdf && !d_weight)) return MCX_ERR_ARG;
  if (n == 0) return MCX_OK;
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  hipLaunchKernelGGL(fill_f32_i32_kernel, dim3((unsigned)b), dim3(256), 0, as_stream(stream),
                     d_tsdf, d_weight, n, tsdf_value, 0);
  MCX_HIP(hipGetLastError());
  return MCX_OK;
}

extern "C" int mcx_integrate_tsdf(float* d_tsdf, int32_t* d_weight, int X, int Y, int Z, float h,
                                  float trunc, int max_w, const float* intr,
                                  const float* d_poses, const float* d_depth, float depth_scale,
                                  int W, int H, int K, void* d_ws, size_t ws_bytes,
                                  void* stream) {
  if (!d_tsdf || !d_weight) return MCX_ERR_ARG;
  if (K > 0 && (!d_poses || !d_depth)) return MCX_ERR_ARG;
  if (!d_ws || ws_bytes < mcx_integrate_workspace_size(W, H)) return MCX_ERR_WORKSPACE;
  if (K == 0) return MCX_OK;
  hipStream_t s = as_stream(stream);
  IntegrateParams P;
  int rc = prepare_integrate(P, X, Y, Z, h, trunc, max_w, intr, depth_scale, W, H, K,
                             static_cast<float*>(d_ws), s);
  if (rc) return rc;
  const int64_t n = (int64_t)X * Y * Z;
  for (int k0 = 0; k0 < K; k0 += kFramesPerPass) {
    P.K = K - k0 < kFramesPerPass ? K - k0 : kFramesPerPass;
    P.poses = d_poses + (int64_t)12 * k0;
    P.depth = d_depth + (int64_t)k0 * W * H;
    const int64_t nb = (n + 255) / 256;
    for (int64_t b0 = 0; b0 < nb; b0 += kMaxLaunchBlocks) {
      P.first_block = b0;
      hipLaunchKernelGGL(integrate_tsdf_kernel, dim3((unsigned)std::min(kMaxLaunchBlocks, nb - b0)),
                         dim3(256), frames_lds_bytes(P.K), s, P, d_tsdf, d_weight);
      MCX_HIP(hipGetLastError());
    }
  }
  return MCX_OK;
}

extern "C" size_t mcx_integrate_tsdf_backward_workspace_size(int W, int H, int K) {
  if (W < 1 || H < 1 || K < 0) return 0;
  return al256((size_t)W * H * sizeof(float)) + 256 + (size_t)K * W * H * 8;
}

extern "C" int mcx_integrate_tsdf_backward(const float* d_grad_tsdf, const int32_t* d_weight0,
                                           int X, int Y, int Z, float h, float trunc, int max_w,
                                           const float* intr, const float* d_poses,
                                           const float* d_depth, float depth_scale, int W,
                                           int H, int K, float* d_grad_depth, void* d_ws,
                                           size_t ws_bytes, void* stream) {
  if (!d_grad_tsdf) return MCX_ERR_ARG;
  if (K > 0 && (!d_poses || !d_depth || !d_grad_depth)) return MCX_ERR_ARG;
  if (!d_ws || ws_bytes < mcx_integrate_tsdf_backward_workspace_size(W, H, K))
    return MCX_ERR_WORKSPACE;
  if (K == 0) return MCX_OK;
  if (K > kFramesPerPass) return MCX_ERR_UNSUPPORTED;  // one pass: weights known per voxel
  hipStream_t s = as_stream(stream);
  char* b = static_cast<char*>(d_ws);
  float* rayfac = reinterpret_cast<float*>(b);
  uint32_t* maxg = reinterpret_cast<uint32_t*>(b + al256((size_t)W * H * sizeof(float)));
  unsigned long long* acc = reinterpret_cast<unsigned long long*>(
      b + al256((size_t)W * H * sizeof(float)) + 256);
  IntegrateParams P;
  int rc = prepare_integrate(P, X, Y, Z, h, trunc, max_w, intr, depth_scale, W, H, K, rayfac, s);
  if (rc) return rc;
  const int64_t n = (int64_t)X * Y * Z;
  const int64_t npix = (int64_t)K * W * H;
  rc = fill_bytes_zero(maxg, 256 + (size_t)npix * 8, s);  // max and accumulators
  if (rc) return rc;
  int64_t mb = (n + 255) / 256;
  if (mb > 4096) mb = 4096;
  hipLaunchKernelGGL(maxabs_kernel, dim3((unsigned)mb), dim3(256), 0, s, d_grad_tsdf, n, maxg);
  MCX_HIP(hipGetLastError());
  const float mdt = max_dtdd(intr, depth_scale, trunc, W, H);
  P.K = K;
  P.poses = d_poses;
  P.depth = d_depth;
  const int64_t nb = (n + 255) / 256;
  for (int64_t b0 = 0; b0 < nb; b0 += kMaxLaunchBlocks) {
    P.first_block = b0;
    hipLaunchKernelGGL(integrate_tsdf_backward_kernel,
                       dim3((unsigned)std::min(kMaxLaunchBlocks, nb - b0)), dim3(256),
                       frames_lds_bytes(K), s, P, d_grad_tsdf, d_weight0, maxg, mdt, acc);
    MCX_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(fixed_to_float_kernel, dim3(blocks_for(npix)), dim3(256), 0, s, acc, npix,
                     maxg, mdt, d_grad_depth);
  MCX_HIP(hipGetLastError());
  return MCX_OK;
}

// ---- batched fp32 fusion (no counterpart in the reference) ----------------
extern "C" int mcx_integrate_tsdf_batched(const mcx_batch* plan, float* d_tsdf, int32_t* d_weight,
                                          int from_init, float init_value, float trunc,
                                          int max_w, const float* intr, const float* d_poses,
                                          const float* d_depth, float depth_scale, int W, int H,
                                          int K, void* d_ws, size_t ws_bytes, void* stream) {
  if (!plan || !d_tsdf || !d_weight) return MCX_ERR_ARG;
  if (plan->elem != MCX_ELEM_F32) return MCX_ERR_ARG;
  if (K > 0 && (!d_poses || !d_depth)) return MCX_ERR_ARG;
  if (K > kFramesPerPass) return MCX_ERR_UNSUPPORTED;
  if (!d_ws || ws_bytes < mcx_integrate_workspace_size(W, H)) return MCX_ERR_WORKSPACE;
  hipStream_t s = as_stream(stream);
  if (K == 0) {
    if (!from_init) return MCX_OK;
    hipLaunchKernelGGL(fill_f32_i32_kernel, dim3(blocks_for(plan->voxels) < 65536u
                                                      ? blocks_for(plan->voxels) : 65536u),
                       dim3(256), 0, s, d_tsdf, d_weight, plan->voxels, init_value, 0);
    MCX_HIP(hipGetLastError());
    return MCX_OK;
  }
  if (plan->blocks >= ((int64_t)1 << 31)) return MCX_ERR_UNSUPPORTED;
  BatchIntegrateParams B;
  int rc = prepare_integrate(B.base, 1, 1, 1, 1.0f, trunc, max_w, intr, depth_scale, W, H, K,
                             static_cast<float*>(d_ws), s);
  if (rc) return rc;
  B.base.K = K;
  B.base.poses = d_poses;
  B.base.depth = d_depth;
  B.bvol = plan->d_bvol;
  B.nvol = plan->nvol;
  B.from_init = from_init ? 1 : 0;
  B.init_value = init_value;
  for (int64_t b0 = 0; b0 < plan->blocks; b0 += kMaxLaunchBlocks) {
    B.first_block = b0;
    hipLaunchKernelGGL(integrate_tsdf_batched_kernel,
                       dim3((unsigned)std::min(kMaxLaunchBlocks, plan->blocks - b0)), dim3(256),
                       frames_lds_bytes(K), s, B, d_tsdf, d_weight);
    MCX_HIP(hipGetLastError());
  }
  return MCX_OK;
}

extern "C" size_t mcx_integrate_tsdf_backward_batched_workspace_size(const mcx_batch* plan, int W,
                                                                     int H, int K) {
  if (!plan || W < 1 || H < 1 || K < 0) return 0;
  return al256((size_t)W * H * sizeof(float)) + 256 + (size_t)plan->nvol * K * W * H * 8;
}

extern "C" int mcx_integrate_tsdf_backward_batched(const mcx_batch* plan,
                                                   const float* d_grad_tsdf,
                                                   const int32_t* d_weight0,
                                                   const uint32_t* d_maxabs, float trunc,
                                                   int max_w, const float* intr,
                                                   const float* d_poses, const float* d_depth,
                                                   float depth_scale, int W, int H, int K,
                                                   float* d_grad_depth, void* d_ws,
                                                   size_t ws_bytes, void* stream) {
  if (!plan || !d_grad_tsdf) return MCX_ERR_ARG;
  if (plan->elem != MCX_ELEM_F32) return MCX_ERR_ARG;
  if (K > 0 && (!d_poses || !d_depth || !d_grad_depth)) return MCX_ERR_ARG;
  if (!d_ws || ws_bytes < mcx_integrate_tsdf_backward_batched_workspace_size(plan, W, H, K))
    return MCX_ERR_WORKSPACE;
  if (K == 0) return MCX_OK;
  if (K > kFramesPerPass) return MCX_ERR_UNSUPPORTED;
  if (plan->blocks >= ((int64_t)1 << 31)) return MCX_ERR_UNSUPPORTED;
  hipStream_t s = as_stream(stream);
  char* b = static_cast<char*>(d_ws);
  float* rayfac = reinterpret_cast<float*>(b);
  uint32_t* maxg = reinterpret_cast<uint32_t*>(b + al256((size_t)W * H * sizeof(float)));
  unsigned long long* acc = reinterpret_cast<unsigned long long*>(
      b + al256((size_t)W * H * sizeof(float)) + 256);
  BatchIntegrateParams B;
  int rc = prepare_integrate(B.base, 1, 1, 1, 1.0f, trunc, max_w, intr, depth_scale, W, H, K,
                             rayfac, s);
  if (rc) return rc;
  const int64_t npix = (int64_t)plan->nvol * K * W * H;
  if (d_maxabs) {
    rc = fill_bytes_zero(acc, (size_t)npix * 8, s);  // the accumulators
    if (rc) return rc;
  } else {
    rc = fill_bytes_zero(maxg, 256 + (size_t)npix * 8, s);  // max and accumulators
    if (rc) return rc;
    int64_t mb = (plan->voxels + 255) / 256;
    if (mb > 4096) mb = 4096;
    hipLaunchKernelGGL(maxabs_kernel, dim3((unsigned)mb), dim3(256), 0, s, d_grad_tsdf,
                       plan->voxels, maxg);
    MCX_HIP(hipGetLastError());
  }
  const uint32_t* mg = d_maxabs ? d_maxabs : maxg;
  const float mdt = max_dtdd(intr, depth_scale, trunc, W, H);
  B.base.K = K;
  B.base.poses = d_poses;
  B.base.depth = d_depth;
  B.bvol = plan->d_bvol;
  B.nvol = plan->nvol;
  B.from_init = 0;
  B.init_value = 0.0f;
  for (int64_t b0 = 0; b0 < plan->blocks; b0 += kMaxLaunchBlocks) {
    B.first_block = b0;
    hipLaunchKernelGGL(integrate_tsdf_backward_batched_kernel,
                       dim3((unsigned)std::min(kMaxLaunchBlocks, plan->blocks - b0)), dim3(256),
                       frames_lds_bytes(K), s, B, d_grad_tsdf, d_weight0, mg, mdt, acc);
    MCX_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(fixed_to_float_kernel, dim3(blocks_for(npix)), dim3(256), 0, s, acc, npix,
                     mg, mdt, d_grad_depth);
  MCX_HIP(hipGetLastError());
  return MCX_OK;
}
