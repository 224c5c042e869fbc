/*
 * raycast_oracle.c -- CPU oracle for rendering a depth and normal map from a
 * packed short2 TSDF volume by ray marching (TEST INFRASTRUCTURE; never
 * linked by the product).
 *
 * Follows the reference's dead KinectFusion raycast:
 *   _cu_voxel_extract_depth_and_normal_map   cu_kinect_fusion_v1.cu:972-986
 *       M = K R^T, IM = M^-1 by cofactors over det (float33_inv,
 *       cu_vector_functions.h:555-571), not launched when |det| < FLT_EPSILON;
 *       camera centre in grid units t / grid_s
 *   d_voxel_extract_depth_and_normal_map     :786-970
 *       ray = unit(IM (x, y, 1)), zero components -> 1e-15 (:813-822);
 *       entry/exit t from the six planes at margin 1 and size-1 (:826-904),
 *       t_min = 1 when only one plane is crossed (:909);
 *       march with dt 1, 0.25 inside the band (:911-969), skipping samples
 *       outside [0, size-1); zero crossing sdf1 in [0,1), sdf in (-1,0)
 *       -> t2 = t1 - dt sdf1 / (sdf - sdf1) (:927-936), z-depth
 *       |ray t2| / |(K^-1 (x, y, 1))| * grid_s (:937-939), normal R^T n
 *       (:941-942); back face sdf1 in (-1,0], sdf in [0,1) ends the ray
 *       (:948-950)
 *   d_voxel_extract_trilinear_interpolated_sdf  :551-659
 *       corners V000 V100 V010 V001 V101 V011 V110 V111, the first that is
 *       +-32766 returns +-1; values tsdf / 32766 (cu_kinect_fusion_v1.h:253)
 *   _d_voxel_extract_surface_normal           :697-776
 *       0 if any corner is +-32766; else per axis the difference of the two
 *       bilinear face values, normalised when the norm exceeds FLT_EPSILON
 *       (float3_unit, cu_vector_functions.h:258-264)
 * Kept the reference's grid convention (voxel i at grid coordinate i); the
 * caller's `origin` shifts it: voxel i sits at world (i + origin) grid_s, so
 * origin 0 is the reference and 0.5 matches voxel-centre volumes.
 * Interpolation is written as nested lerps fmaf(t, b - a, a) (the same
 * polynomial as the reference's weighted sums).  Colour is not produced (the
 * volume carries none).
 * Two guards the reference lacks, so that every ray ends: the march stops
 * once t passes the last parameter at which the ray can be inside the
 * volume grown by one voxel (every later sample would be skipped), and
 * after 2^20 steps.
 *
 * vol: X*Y*Z uint32 words, tsdf in the low 16 bits (signed).  intr: {fx, fy,
 * cx, cy}.  pose: {R row-major 3x3, t} camera-to-world in metres.  Outputs
 * depth[H*W] (0 = no hit) and normals[H*W*3] (0 = no hit), row-major by
 * pixel (y, x); rows split over nthreads OpenMP threads.  Returns 0, or 1
 * when M is singular (outputs untouched).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>

#include "mc_oracle.h"

#define ORC_MAX_TSDF 32766
#define ORC_RAY_MAX_STEPS (1 << 20)

int orc_raycast_setup(const float* intr, const float* pose, float* im) {
  const float fx = intr[0], fy = intr[1], cx = intr[2], cy = intr[3];
  const float* R = pose;
  float M[3][3];
  for (int j = 0; j < 3; ++j) {
    /* M = K R^T, K = [[fx, 0, cx], [0, fy, cy], [0, 0, 1]] */
    M[0][j] = fx * R[3 * j + 0] + cx * R[3 * j + 2];
    M[1][j] = fy * R[3 * j + 1] + cy * R[3 * j + 2];
    M[2][j] = R[3 * j + 2];
  }
  /* cofactors C[i][j] */
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
  if (det > -FLT_EPSILON && det < FLT_EPSILON) return 1;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) im[3 * i + j] = C[j][i] / det; /* adjugate / det */
  return 0;
}

static inline float lerpf_(float a, float b, float t) { return fmaf(t, b - a, a); }

static inline int16_t tsdf_at(const uint32_t* vol, int64_t i) {
  return (int16_t)(vol[i] & 0xffffu);
}

/* corner offsets in the reference's order V000 V100 V010 V001 V101 V011 V110 V111 */
static inline void corner_index(int64_t base, int64_t X, int64_t XY, int64_t* o) {
  o[0] = base;
  o[1] = base + 1;
  o[2] = base + X;
  o[3] = base + XY;
  o[4] = base + XY + 1;
  o[5] = base + XY + X;
  o[6] = base + X + 1;
  o[7] = base + XY + X + 1;
}

static float sample_sdf(const uint32_t* vol, int X, int Y, float px, float py, float pz) {
  const int ix = (int)px, iy = (int)py, iz = (int)pz;
  const float tx = px - (float)ix, ty = py - (float)iy, tz = pz - (float)iz;
  const int64_t XY = (int64_t)X * Y;
  int64_t o[8];
  corner_index((int64_t)iz * XY + (int64_t)iy * X + ix, X, XY, o);
  float v[8];
  for (int k = 0; k < 8; ++k) {
    const int s = tsdf_at(vol, o[k]);
    if (s == ORC_MAX_TSDF) return 1.0f;
    if (s == -ORC_MAX_TSDF) return -1.0f;
    v[k] = (float)s / (float)ORC_MAX_TSDF;
  }
  const float c00 = lerpf_(v[0], v[1], tx); /* y0 z0 */
  const float c10 = lerpf_(v[2], v[6], tx); /* y1 z0 */
  const float c01 = lerpf_(v[3], v[4], tx); /* y0 z1 */
  const float c11 = lerpf_(v[5], v[7], tx); /* y1 z1 */
  const float c0 = lerpf_(c00, c10, ty);
  const float c1 = lerpf_(c01, c11, ty);
  return lerpf_(c0, c1, tz);
}

static void surface_normal(const uint32_t* vol, int X, int Y, int Z, float px, float py,
                           float pz, float* n) {
  int ix = (int)px, iy = (int)py, iz = (int)pz;
  if (ix < 0) ix = 0;
  if (iy < 0) iy = 0;
  if (iz < 0) iz = 0;
  if (ix > X - 2) ix = X - 2;
  if (iy > Y - 2) iy = Y - 2;
  if (iz > Z - 2) iz = Z - 2;
  const float tx = px - (float)ix, ty = py - (float)iy, tz = pz - (float)iz;
  const int64_t XY = (int64_t)X * Y;
  int64_t o[8];
  corner_index((int64_t)iz * XY + (int64_t)iy * X + ix, X, XY, o);
  float v[8];
  n[0] = n[1] = n[2] = 0.0f;
  for (int k = 0; k < 8; ++k) {
    const int s = tsdf_at(vol, o[k]);
    if (s == ORC_MAX_TSDF || s == -ORC_MAX_TSDF) return;
    v[k] = (float)s / (float)ORC_MAX_TSDF;
  }
  /* bilinear face values: f(a00, a10, a01, a11; s, t) = lerp(lerp(a00,a10,s), lerp(a01,a11,s), t) */
  const float x0 = lerpf_(lerpf_(v[0], v[2], ty), lerpf_(v[3], v[5], ty), tz);
  const float x1 = lerpf_(lerpf_(v[1], v[6], ty), lerpf_(v[4], v[7], ty), tz);
  const float y0 = lerpf_(lerpf_(v[0], v[1], tx), lerpf_(v[3], v[4], tx), tz);
  const float y1 = lerpf_(lerpf_(v[2], v[6], tx), lerpf_(v[5], v[7], tx), tz);
  const float z0 = lerpf_(lerpf_(v[0], v[1], tx), lerpf_(v[2], v[6], tx), ty);
  const float z1 = lerpf_(lerpf_(v[3], v[4], tx), lerpf_(v[5], v[7], tx), ty);
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

static void cast_one(const uint32_t* vol, int X, int Y, int Z, float h, const float* intr,
                     const float* pose, const float* im, const float* cc, int x, int y,
                     float* depth, float* normal) {
  depth[0] = 0.0f;
  normal[0] = normal[1] = normal[2] = 0.0f;
  const float xf = (float)x, yf = (float)y;
  float r[3];
  for (int i = 0; i < 3; ++i) r[i] = fmaf(im[3 * i + 1], yf, fmaf(im[3 * i + 0], xf, im[3 * i + 2]));
  const float rn = sqrtf(fmaf(r[2], r[2], fmaf(r[1], r[1], r[0] * r[0])));
  if (rn > FLT_EPSILON || rn < -FLT_EPSILON)
    for (int i = 0; i < 3; ++i) r[i] = r[i] / rn;
  for (int i = 0; i < 3; ++i)
    if (r[i] == 0.0f) r[i] = 1e-15f;
  const float mx[3] = {(float)X - 1.0f, (float)Y - 1.0f, (float)Z - 1.0f};
  const float margin = 1.0f;
  float t_min = FLT_MAX, t_max = 0.0f;
  for (int a = 0; a < 3; ++a) {
    const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
    for (int side = 0; side < 2; ++side) {
      const float t = ((side ? mx[a] : margin) - cc[a]) / r[a];
      if (t > 0.0f) {
        const float u = fmaf(r[b], t, cc[b]), w = fmaf(r[c], t, cc[c]);
        if (u > margin && u < mx[b] && w > margin && w < mx[c]) {
          if (t < t_min) t_min = t;
          if (t > t_max) t_max = t;
        }
      }
    }
  }
  if (t_min == FLT_MAX && t_max == 0.0f) return;
  const float dt_max = 1.0f, dt_min = 0.25f;
  if (t_min == t_max) t_min = dt_max * margin;
  /* last t at which the ray can lie in [-1, size]^3 */
  float t_far = FLT_MAX;
  for (int a = 0; a < 3; ++a) {
    const float e = ((r[a] > 0.0f ? mx[a] + 1.0f : -1.0f) - cc[a]) / r[a];
    if (e < t_far) t_far = e;
  }
  t_far = t_far + fabsf(t_far) * 0x1p-20f + 1.0f;
  float dt = dt_max, t1 = 0.0f, sdf1 = 1.0f;
  int fine = 0, steps = 0;
  for (float t = t_min; t < t_max; t += dt) {
    if (t > t_far || ++steps > ORC_RAY_MAX_STEPS) break;
    const float px = fmaf(r[0], t, cc[0]), py = fmaf(r[1], t, cc[1]), pz = fmaf(r[2], t, cc[2]);
    if (px < 0.0f || py < 0.0f || pz < 0.0f || px >= mx[0] || py >= mx[1] || pz >= mx[2])
      continue;
    const float sdf = sample_sdf(vol, X, Y, px, py, pz);
    if (sdf < 0.0f && sdf > -1.0f && sdf1 >= 0.0f && sdf1 < 1.0f) {
      const float t2 = t1 - dt * sdf1 / (sdf - sdf1);
      const float nx = r[0] * t2, ny = r[1] * t2, nz = r[2] * t2;
      const float ax = (xf - intr[2]) / intr[0], ay = (yf - intr[3]) / intr[1];
      const float len = sqrtf(fmaf(nz, nz, fmaf(ny, ny, nx * nx)));
      depth[0] = len / sqrtf(fmaf(ay, ay, ax * ax) + 1.0f) * h;
      float n[3];
      surface_normal(vol, X, Y, Z, nx + cc[0], ny + cc[1], nz + cc[2], n);
      /* R^T n */
      normal[0] = fmaf(pose[6], n[2], fmaf(pose[3], n[1], pose[0] * n[0]));
      normal[1] = fmaf(pose[7], n[2], fmaf(pose[4], n[1], pose[1] * n[0]));
      normal[2] = fmaf(pose[8], n[2], fmaf(pose[5], n[1], pose[2] * n[0]));
      return;
    }
    if (sdf1 <= 0.0f && sdf1 > -1.0f && sdf >= 0.0f && sdf < 1.0f) return; /* back face */
    if (sdf > -1.0f && sdf < 1.0f) {
      if (!fine) {
        t -= dt;
        fine = 1;
        dt = dt_min;
        continue;
      }
    } else if (fine) {
      fine = 0;
      dt = dt_max;
    }
    t1 = t;
    sdf1 = sdf;
  }
}

int orc_raycast(const uint32_t* vol, int X, int Y, int Z, float h, float origin,
                const float* intr, const float* pose, int W, int H, float* depth,
                float* normals, int nthreads) {
  float im[9];
  if (orc_raycast_setup(intr, pose, im)) return 1;
  float cc[3];
  for (int a = 0; a < 3; ++a) cc[a] = pose[9 + a] / h - origin;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1)
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const int64_t p = (int64_t)y * W + x;
      cast_one(vol, X, Y, Z, h, intr, pose, im, cc, x, y, depth + p, normals + 3 * p);
    }
  return 0;
}
