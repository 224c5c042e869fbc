// mcx_raycast.hip -- depth and normal maps ray-marched through a packed short2
// TSDF volume (gfx950).
//
// Replaces _cu_voxel_extract_depth_and_normal_map (cu_kinect_fusion_v1.cu:972-986)
// and its kernel d_voxel_extract_depth_and_normal_map (:786-970) with the
// samplers d_voxel_extract_trilinear_interpolated_sdf (:551-659) and
// _d_voxel_extract_surface_normal (:697-776).  Per pixel:
//   ray   = unit(IM (x, y, 1)), IM = (K R^T)^-1 (host, cofactors / det),
//           zero components replaced by 1e-15;
//   t range from the six planes at 1 and size-1 (t_min = 1 when one plane is
//           crossed); march t += 1, or 0.25 inside the band |sdf| < 1 (one
//           step back on entering it), samples outside [0, size-1) skipped;
//   hit   sdf1 in [0, 1) then sdf in (-1, 0): t2 = t1 - dt sdf1 / (sdf - sdf1),
//           z-depth |ray t2| / |(K^-1 (x, y, 1))| * h, normal R^T unit(grad);
//   back face  sdf1 in (-1, 0] then sdf in [0, 1): the ray ends, no hit.
// The trilinear sample returns +-1 at the first corner (V000 V100 V010 V001
// V101 V011 V110 V111) holding +-32766; the normal is 0 if any corner does.
// Two guards the reference lacks end every ray: t past the last parameter at
// which the ray can lie in the volume grown by one voxel (all later samples
// would be skipped), and 2^20 steps.
//
// Layout: one thread per pixel; each wave covers an 8 x 8 pixel tile so the
// 64 rays of a wave march through neighbouring voxels (shared cache lines);
// workgroups of 16 x 16 pixels.  Each sample reads its eight corner words as
// four x-neighbour pairs (one 64-bit load each).  Bit-exact with the CPU oracle: every fused
// multiply-add is explicit and the build disables contraction.
#include <hip/hip_runtime.h>

#include <float.h>

#include "mcx_internal.h"

namespace mcx {
namespace {

constexpr int kMaxTsdf = 32766;
constexpr int kRayMaxSteps = 1 << 20;
constexpr int kTileW = 16, kTileH = 16;

struct RaycastArgs {
  const int16_t* vol16;  // tsdf halves: word i's tsdf at vol16[2 i]
  float* depth;
  float* normal;  // 3 floats per pixel, may be null
  int X, Y, Z, W, H;
  float h;
  float im[9];  // (K R^T)^-1, row-major
  float R[9];   // camera-to-world rotation, row-major
  float cc[3];  // camera centre in grid units
  float fx, fy, cx, cy;
};

__device__ __forceinline__ float lerpf(float a, float b, float t) { return fmaf(t, b - a, a); }

// s / 32766 correctly rounded for every 16-bit s: one multiply by the
// rounded reciprocal and one fused correction step (checked exhaustively
// against IEEE division over [-32767, 32767]).
__device__ __forceinline__ float unpack_tsdf(int s) {
  constexpr float r = 1.0f / 32766.0f;
  const float x = (float)s;
  const float q = x * r;
  const float e = fmaf(-q, 32766.0f, x);
  return fmaf(e, r, q);
}

// two consecutive voxel words (4-byte aligned, which the vector memory path
// accepts for a 64-bit load)
typedef uint32_t u32x2a4 __attribute__((ext_vector_type(2), aligned(4), may_alias));

// Corner tsdf values in the reference's order V000 V100 V010 V001 V101 V011
// V110 V111 for the cell whose lower corner is word `base`: the words of
// (x, x + 1) in the rows (y, z), (y + 1, z), (y, z + 1), (y + 1, z + 1).
__device__ __forceinline__ void load_corners(const int16_t* __restrict__ v16, int64_t base,
                                             int64_t X, int64_t XY, int* s) {
  const uint32_t* w = reinterpret_cast<const uint32_t*>(v16) + base;
  auto pair = [](const uint32_t* q) { return *reinterpret_cast<const u32x2a4*>(q); };
  const u32x2a4 a = pair(w);           // V000 V100
  const u32x2a4 b = pair(w + X);       // V010 V110
  const u32x2a4 c = pair(w + XY);      // V001 V101
  const u32x2a4 d = pair(w + XY + X);  // V011 V111
  s[0] = (int16_t)a[0];
  s[1] = (int16_t)a[1];
  s[2] = (int16_t)b[0];
  s[3] = (int16_t)c[0];
  s[4] = (int16_t)c[1];
  s[5] = (int16_t)d[0];
  s[6] = (int16_t)b[1];
  s[7] = (int16_t)d[1];
}

__device__ __forceinline__ float sample_sdf(const RaycastArgs& A, int64_t XY, float px, float py,
                                            float pz) {
  const int ix = (int)px, iy = (int)py, iz = (int)pz;
  const float tx = px - (float)ix, ty = py - (float)iy, tz = pz - (float)iz;
  int s[8];
  load_corners(A.vol16, (int64_t)iz * XY + (int64_t)iy * A.X + ix, A.X, XY, s);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (s[k] == kMaxTsdf) return 1.0f;
    if (s[k] == -kMaxTsdf) return -1.0f;
  }
  float v[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = unpack_tsdf(s[k]);
  const float c00 = lerpf(v[0], v[1], tx);
  const float c10 = lerpf(v[2], v[6], tx);
  const float c01 = lerpf(v[3], v[4], tx);
  const float c11 = lerpf(v[5], v[7], tx);
  return lerpf(lerpf(c00, c10, ty), lerpf(c01, c11, ty), tz);
}

__device__ void surface_normal(const RaycastArgs& A, int64_t XY, float px, float py, float pz,
                               float* n) {
  const int ix = min(max((int)px, 0), A.X - 2);
  const int iy = min(max((int)py, 0), A.Y - 2);
  const int iz = min(max((int)pz, 0), A.Z - 2);
  const float tx = px - (float)ix, ty = py - (float)iy, tz = pz - (float)iz;
  int s[8];
  load_corners(A.vol16, (int64_t)iz * XY + (int64_t)iy * A.X + ix, A.X, XY, s);
  n[0] = n[1] = n[2] = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if (s[k] == kMaxTsdf || s[k] == -kMaxTsdf) return;
  float v[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = unpack_tsdf(s[k]);
  const float x0 = lerpf(lerpf(v[0], v[2], ty), lerpf(v[3], v[5], ty), tz);
  const float x1 = lerpf(lerpf(v[1], v[6], ty), lerpf(v[4], v[7], ty), tz);
  const float y0 = lerpf(lerpf(v[0], v[1], tx), lerpf(v[3], v[4], tx), tz);
  const float y1 = lerpf(lerpf(v[2], v[6], tx), lerpf(v[5], v[7], tx), tz);
  const float z0 = lerpf(lerpf(v[0], v[1], tx), lerpf(v[2], v[6], tx), ty);
  const float z1 = lerpf(lerpf(v[3], v[4], tx), lerpf(v[5], v[7], tx), ty);
  float gx = x1 - x0, gy = y1 - y0, gz = z1 - z0;
  const float norm = sqrtf(fmaf(gz, gz, fmaf(gy, gy, gx * gx)));
  if (norm > FLT_EPSILON || norm < -FLT_EPSILON) {
    gx = gx / norm;
    gy = gy / norm;
    gz = gz / norm;
  }
  n[0] = gx;
  n[1] = gy;
  n[2] = gz;
}

__global__ __launch_bounds__(256) void raycast_kernel(RaycastArgs A) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int x = blockIdx.x * kTileW + (w & 1) * 8 + (lane & 7);
  const int y = blockIdx.y * kTileH + (w >> 1) * 8 + (lane >> 3);
  if (x >= A.W || y >= A.H) return;
  const int64_t pix = (int64_t)y * A.W + x;
  float out_d = 0.0f, out_n[3] = {0.0f, 0.0f, 0.0f};
  const float xf = (float)x, yf = (float)y;
  float r[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    r[i] = fmaf(A.im[3 * i + 1], yf, fmaf(A.im[3 * i + 0], xf, A.im[3 * i + 2]));
  const float rn = sqrtf(fmaf(r[2], r[2], fmaf(r[1], r[1], r[0] * r[0])));
  if (rn > FLT_EPSILON || rn < -FLT_EPSILON) {
#pragma unroll
    for (int i = 0; i < 3; ++i) r[i] = r[i] / rn;
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (r[i] == 0.0f) r[i] = 1e-15f;
  const float mx[3] = {(float)A.X - 1.0f, (float)A.Y - 1.0f, (float)A.Z - 1.0f};
  const float margin = 1.0f;
  float t_min = FLT_MAX, t_max = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const float t = ((side ? mx[a] : margin) - A.cc[a]) / r[a];
      if (t > 0.0f) {
        const float u = fmaf(r[b], t, A.cc[b]), v = fmaf(r[c], t, A.cc[c]);
        if (u > margin && u < mx[b] && v > margin && v < mx[c]) {
          t_min = fminf(t_min, t);
          t_max = fmaxf(t_max, t);
        }
      }
    }
  }
  if (!(t_min == FLT_MAX && t_max == 0.0f)) {
    const float dt_max = 1.0f, dt_min = 0.25f;
    if (t_min == t_max) t_min = dt_max * margin;
    float t_far = FLT_MAX;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float e = ((r[a] > 0.0f ? mx[a] + 1.0f : -1.0f) - A.cc[a]) / r[a];
      t_far = fminf(t_far, e);
    }
    t_far = t_far + fabsf(t_far) * 0x1p-20f + 1.0f;
    const int64_t XY = (int64_t)A.X * A.Y;
    float dt = dt_max, t1 = 0.0f, sdf1 = 1.0f;
    bool fine = false;
    int steps = 0;
    for (float t = t_min; t < t_max; t += dt) {
      if (t > t_far || ++steps > kRayMaxSteps) break;
      const float px = fmaf(r[0], t, A.cc[0]), py = fmaf(r[1], t, A.cc[1]),
                  pz = fmaf(r[2], t, A.cc[2]);
      if (px < 0.0f || py < 0.0f || pz < 0.0f || px >= mx[0] || py >= mx[1] || pz >= mx[2])
        continue;
      const float sdf = sample_sdf(A, XY, px, py, pz);
      if (sdf < 0.0f && sdf > -1.0f && sdf1 >= 0.0f && sdf1 < 1.0f) {
        const float t2 = t1 - dt * sdf1 / (sdf - sdf1);
        const float nx = r[0] * t2, ny = r[1] * t2, nz = r[2] * t2;
        const float ax = (xf - A.cx) / A.fx, ay = (yf - A.cy) / A.fy;
        const float len = sqrtf(fmaf(nz, nz, fmaf(ny, ny, nx * nx)));
        out_d = len / sqrtf(fmaf(ay, ay, ax * ax) + 1.0f) * A.h;
        float n[3];
        surface_normal(A, XY, nx + A.cc[0], ny + A.cc[1], nz + A.cc[2], n);
        out_n[0] = fmaf(A.R[6], n[2], fmaf(A.R[3], n[1], A.R[0] * n[0]));
        out_n[1] = fmaf(A.R[7], n[2], fmaf(A.R[4], n[1], A.R[1] * n[0]));
        out_n[2] = fmaf(A.R[8], n[2], fmaf(A.R[5], n[1], A.R[2] * n[0]));
        break;
      }
      if (sdf1 <= 0.0f && sdf1 > -1.0f && sdf >= 0.0f && sdf < 1.0f) break;  // back face
      if (sdf > -1.0f && sdf < 1.0f) {
        if (!fine) {
          t -= dt;
          fine = true;
          dt = dt_min;
          continue;
        }
      } else if (fine) {
        fine = false;
        dt = dt_max;
      }
      t1 = t;
      sdf1 = sdf;
    }
  }
  A.depth[pix] = out_d;
  if (A.normal) {
    A.normal[3 * pix + 0] = out_n[0];
    A.normal[3 * pix + 1] = out_n[1];
    A.normal[3 * pix + 2] = out_n[2];
  }
}

// IM = (K R^T)^-1 by cofactors over the determinant (float33_inv,
// cu_vector_functions.h:555-571); false when |det| < FLT_EPSILON.
bool camera_inverse(const float* intr, const float* R, float* im) {
  const float fx = intr[0], fy = intr[1], cx = intr[2], cy = intr[3];
  float M[3][3];
  for (int j = 0; j < 3; ++j) {
    M[0][j] = fx * R[3 * j + 0] + cx * R[3 * j + 2];
    M[1][j] = fy * R[3 * j + 1] + cy * R[3 * j + 2];
    M[2][j] = R[3 * j + 2];
  }
  float C[3][3];
  C[0][0] = M[1][1] * M[2][2] - M[1][2] * M[2][1];
  C[0][1] = M[1][2] * M[2][0] - M[1][0] * M[2][2];
  C[0][2] = M[1][0] * M[2][1] - M[1][1] * M[2][0];
  C[1][0] = M[0][2] * M[2][1] - M[0][1] * M[2][2];
  C[1][1] = M[0][0] * M[2][2] - M[0][2] * M[2][0];
  C[1][2] = M[0][1] * M[2][0] - M[0][0] * M[2][1];
  C[2][0] = M[0][1] * M[1][2] - M[0][2] * M[1][1];
  C[2][1] = M[0][2] * M[1][0] - M[0][0] * M[1][2];
  C[2][2] = M[0][0] * M[1][1] - M[0][1] * M[1][0];
  const float det = M[0][0] * C[0][0] + M[0][1] * C[0][1] + M[0][2] * C[0][2];
  if (det > -FLT_EPSILON && det < FLT_EPSILON) return false;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) im[3 * i + j] = C[j][i] / det;
  return true;
}

}  // namespace
}  // namespace mcx

using namespace mcx;

extern "C" int mcx_raycast(const void* d_vol, int X, int Y, int Z, float h, float origin,
                           const float* intr, const float* pose, int W, int H, float* d_depth,
                           float* d_normals, void* stream) {
  if (!d_vol || !intr || !pose || !d_depth) return MCX_ERR_ARG;
  if (X < 2 || Y < 2 || Z < 2 || W < 1 || H < 1 || !(h > 0.0f)) return MCX_ERR_ARG;
  if ((int64_t)W * H >= ((int64_t)1 << 40)) return MCX_ERR_ARG;
  RaycastArgs A;
  if (!camera_inverse(intr, pose, A.im)) return MCX_ERR_ARG;
  A.vol16 = static_cast<const int16_t*>(d_vol);
  A.depth = d_depth;
  A.normal = d_normals;
  A.X = X;
  A.Y = Y;
  A.Z = Z;
  A.W = W;
  A.H = H;
  A.h = h;
  for (int i = 0; i < 9; ++i) A.R[i] = pose[i];
  for (int a = 0; a < 3; ++a) A.cc[a] = pose[9 + a] / h - origin;
  A.fx = intr[0];
  A.fy = intr[1];
  A.cx = intr[2];
  A.cy = intr[3];
  const dim3 grid((unsigned)((W + kTileW - 1) / kTileW), (unsigned)((H + kTileH - 1) / kTileH));
  if (grid.y > 65535u) return MCX_ERR_ARG;
  hipLaunchKernelGGL(raycast_kernel, grid, dim3(256), 0, as_stream(stream), A);
  MCX_HIP(hipGetLastError());
  return MCX_OK;
}
