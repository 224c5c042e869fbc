/*
 * integrate_oracle.c -- CPU oracle for TSDF volume init and projective
 * integration (TEST INFRASTRUCTURE; never linked by the product).
 *
 * Follows the reference's dead KinectFusion kernels
 *   d_voxel_initialize            cu_kinect_fusion_v1.cu:364-379
 *   d_voxel_volumetric_depth_scaling  :416-425
 *   d_voxel_volumetric_integration    :427-532
 * with the packing the benchmark configuration fixes (PCL-KinFu style):
 *   tsdf_new = (tsdf_old * w_old + clamp(sdf / trunc, -1, 1)) / (w_old + 1),
 *   w_new    = min(w_old + 1, max_w), tsdf stored as trunc-toward-zero of
 *   tsdf_new * 32767 clamped to +-32767.
 * Kept from the reference: ray-distance sdf (depth scaled by
 * sqrt(|K^-1 [u v]|^2 + 1) minus |q|, :420-424, :456-458), nearest-pixel lookup
 * with round-half-even and the [0, W-2) x [0, H-2) pixel window (:447-449),
 * the q.z < FLT_MIN and depth <= FLT_MIN skips (:446, :455), the two
 * truncation distances (:455-463) and init to (-32766, 0) (:375-376).
 * Changed on purpose: voxel centres at (i + 0.5) h so that integration and
 * extraction (getNodeCoo) agree; integer weights.
 *
 * Every float operation is written out (fmaf where a fused multiply-add is
 * meant) so the GPU kernel can reproduce the sequence bit for bit.
 *
 * intr = {fx, fy, cx, cy}; poses = nframes x 12 floats {R row-major 3x3, t},
 * camera-to-world (p_world = R q + t), so q = R^T (p - t)
 * (float_camera_extrinsic_backward, cu_geometry_functions.h:125-129).
 */
#include <float.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#include "mc_oracle.h"

void orc_volume_init(uint32_t* vol, int64_t n) {
  /* tsdf = -CU_VOXEL_MAX_TSDF (-32766), w = 0 */
  uint32_t v = (uint32_t)(uint16_t)(int16_t)-32766;
  for (int64_t i = 0; i < n; ++i) vol[i] = v;
}

/* One observation of a packed voxel.  depth_f32 selects float frames (the
 * reference's float *d_depth_ptr, :534) over uint16 ones.  Truncation as the
 * reference (:455-463, the volume's min_t and max_t): skip sdf < -trunc_neg,
 * t = sdf > 0 ? min(sdf / trunc_pos, 1) : sdf / trunc_neg; with equal
 * distances this is clamp(sdf / trunc, -1, 1) (sdf >= -trunc gives
 * sdf / trunc >= -1 exactly). */
static inline uint32_t update_voxel(uint32_t cur, float px, float py, float pz,
                                    float trunc_neg, float trunc_pos, int max_w,
                                    const float* intr, const float* P, const void* depth,
                                    int depth_f32, float depth_scale, int W, int H) {
  float fx = intr[0], fy = intr[1], cx = intr[2], cy = intr[3];
  float dx = px - P[9], dy = py - P[10], dz = pz - P[11];
  /* q = R^T d */
  float qx = fmaf(P[6], dz, fmaf(P[3], dy, P[0] * dx));
  float qy = fmaf(P[7], dz, fmaf(P[4], dy, P[1] * dx));
  float qz = fmaf(P[8], dz, fmaf(P[5], dy, P[2] * dx));
  if (qz < FLT_MIN) return cur;
  float u = fmaf(fx, qx / qz, cx);
  float v = fmaf(fy, qy / qz, cy);
  if (!(fabsf(u) < 1e9f) || !(fabsf(v) < 1e9f)) return cur;
  int ui = (int)rintf(u), vi = (int)rintf(v);
  if (ui < 0 || vi < 0 || ui >= W - 2 || vi >= H - 2) return cur;
  int64_t pix = (int64_t)vi * W + ui;
  float dm = depth_f32 ? ((const float*)depth)[pix] : (float)((const uint16_t*)depth)[pix];
  float d = dm * depth_scale;
  float ax = ((float)ui - cx) / fx, ay = ((float)vi - cy) / fy;
  float zval = d * sqrtf(fmaf(ay, ay, ax * ax) + 1.0f);
  if (!(zval > FLT_MIN)) return cur;
  float qn = sqrtf(fmaf(qz, qz, fmaf(qy, qy, qx * qx)));
  float sdf = zval - qn;
  if (!(sdf >= -trunc_neg)) return cur;
  float tsdf;
  if (sdf > 0.0f) {
    tsdf = sdf / trunc_pos;
    if (tsdf > 1.0f) tsdf = 1.0f;
  } else {
    tsdf = sdf / trunc_neg;
  }
  int16_t ti = (int16_t)(cur & 0xffffu);
  int wi = (int)(cur >> 16);
  float told = (float)ti / 32767.0f;
  float tnew = fmaf(told, (float)wi, tsdf) / (float)(wi + 1);
  int wnew = wi + 1 < max_w ? wi + 1 : max_w;
  float sc = tnew * 32767.0f;
  int fixedp = (int)sc; /* truncation toward zero */
  if (fixedp > 32767) fixedp = 32767;
  if (fixedp < -32767) fixedp = -32767;
  return (uint32_t)(uint16_t)(int16_t)fixedp | ((uint32_t)wnew << 16);
}

void orc_integrate_ex(uint32_t* vol, int X, int Y, int Z, float h, float trunc_neg,
                      float trunc_pos, int max_w, const float* intr, const float* poses,
                      const void* depth, int depth_f32, float depth_scale, int W, int H,
                      int nframes) {
  const size_t esize = depth_f32 ? sizeof(float) : sizeof(uint16_t);
#pragma omp parallel for schedule(static)
  for (int z = 0; z < Z; ++z)
    for (int y = 0; y < Y; ++y)
      for (int x = 0; x < X; ++x) {
        int64_t i = ((int64_t)z * Y + y) * X + x;
        float px = ((float)x + 0.5f) * h, py = ((float)y + 0.5f) * h,
              pz = ((float)z + 0.5f) * h;
        uint32_t cur = vol[i];
        for (int k = 0; k < nframes; ++k)
          cur = update_voxel(cur, px, py, pz, trunc_neg, trunc_pos, max_w, intr, poses + 12 * k,
                             (const char*)depth + (size_t)k * W * H * esize, depth_f32,
                             depth_scale, W, H);
        vol[i] = cur;
      }
}

void orc_integrate(uint32_t* vol, int X, int Y, int Z, float h, float trunc,
                   int max_w, const float* intr, const float* poses,
                   const uint16_t* depth, float depth_scale, int W, int H,
                   int nframes) {
  orc_integrate_ex(vol, X, Y, Z, h, trunc, trunc, max_w, intr, poses, depth, 0, depth_scale, W,
                   H, nframes);
}

/* ---- fp32 TSDF volumes (differentiable form): the same observation stored
 * unpacked, f <- fma(f, w, t) / (w + 1), w <- min(w + 1, max_w); float depth
 * frames.  The pixel lookup and every float operation follow update_voxel. */
static inline int observe_f32(float px, float py, float pz, float trunc, const float* intr,
                              const float* P, const float* depth, float depth_scale, int W,
                              int H, float* t_out) {
  float fx = intr[0], fy = intr[1], cx = intr[2], cy = intr[3];
  float dx = px - P[9], dy = py - P[10], dz = pz - P[11];
  float qx = fmaf(P[6], dz, fmaf(P[3], dy, P[0] * dx));
  float qy = fmaf(P[7], dz, fmaf(P[4], dy, P[1] * dx));
  float qz = fmaf(P[8], dz, fmaf(P[5], dy, P[2] * dx));
  if (qz < FLT_MIN) return 0;
  float u = fmaf(fx, qx / qz, cx);
  float v = fmaf(fy, qy / qz, cy);
  if (!(fabsf(u) < 1e9f) || !(fabsf(v) < 1e9f)) return 0;
  int ui = (int)rintf(u), vi = (int)rintf(v);
  if (ui < 0 || vi < 0 || ui >= W - 2 || vi >= H - 2) return 0;
  float d = depth[(int64_t)vi * W + ui] * depth_scale;
  float ax = ((float)ui - cx) / fx, ay = ((float)vi - cy) / fy;
  float zval = d * sqrtf(fmaf(ay, ay, ax * ax) + 1.0f);
  if (!(zval > FLT_MIN)) return 0;
  float qn = sqrtf(fmaf(qz, qz, fmaf(qy, qy, qx * qx)));
  float sdf = zval - qn;
  if (!(sdf >= -trunc)) return 0;
  float t = sdf / trunc;
  if (t > 1.0f) t = 1.0f;
  if (t < -1.0f) t = -1.0f;
  *t_out = t;
  return 1;
}

void orc_integrate_tsdf(float* tsdf, int32_t* w, int X, int Y, int Z, float h, float trunc,
                        int max_w, const float* intr, const float* poses, const float* depth,
                        float depth_scale, int W, int H, int nframes) {
#pragma omp parallel for schedule(static)
  for (int z = 0; z < Z; ++z)
    for (int y = 0; y < Y; ++y)
      for (int x = 0; x < X; ++x) {
        int64_t i = ((int64_t)z * Y + y) * X + x;
        float px = ((float)x + 0.5f) * h, py = ((float)y + 0.5f) * h,
              pz = ((float)z + 0.5f) * h;
        float f = tsdf[i];
        int wi = w[i];
        for (int k = 0; k < nframes; ++k) {
          float t;
          if (!observe_f32(px, py, pz, trunc, intr, poses + 12 * k,
                           depth + (int64_t)k * W * H, depth_scale, W, H, &t))
            continue;
          f = fmaf(f, (float)wi, t) / (float)(wi + 1);
          wi = wi + 1 < max_w ? wi + 1 : max_w;
        }
        tsdf[i] = f;
        w[i] = wi;
      }
}

/* FP64 form for finite differences with respect to the depth values: the
 * pixel of each observation is chosen exactly as above (in fp32, from the
 * pose: independent of the depth), everything that depends on the depth is
 * evaluated in double. */
void orc_integrate_tsdf_f64(double* tsdf, int32_t* w, int X, int Y, int Z, float h, float trunc,
                            int max_w, const float* intr, const float* poses,
                            const double* depth, double depth_scale, int W, int H, int nframes) {
  const float fx = intr[0], fy = intr[1], cx = intr[2], cy = intr[3];
#pragma omp parallel for schedule(static)
  for (int z = 0; z < Z; ++z)
    for (int y = 0; y < Y; ++y)
      for (int x = 0; x < X; ++x) {
        int64_t i = ((int64_t)z * Y + y) * X + x;
        float px = ((float)x + 0.5f) * h, py = ((float)y + 0.5f) * h,
              pz = ((float)z + 0.5f) * h;
        double f = tsdf[i];
        int wi = w[i];
        for (int k = 0; k < nframes; ++k) {
          const float* P = poses + 12 * k;
          float dx = px - P[9], dy = py - P[10], dz = pz - P[11];
          float qx = fmaf(P[6], dz, fmaf(P[3], dy, P[0] * dx));
          float qy = fmaf(P[7], dz, fmaf(P[4], dy, P[1] * dx));
          float qz = fmaf(P[8], dz, fmaf(P[5], dy, P[2] * dx));
          if (qz < FLT_MIN) continue;
          float u = fmaf(fx, qx / qz, cx);
          float v = fmaf(fy, qy / qz, cy);
          if (!(fabsf(u) < 1e9f) || !(fabsf(v) < 1e9f)) continue;
          int ui = (int)rintf(u), vi = (int)rintf(v);
          if (ui < 0 || vi < 0 || ui >= W - 2 || vi >= H - 2) continue;
          double d = depth[(int64_t)k * W * H + (int64_t)vi * W + ui] * depth_scale;
          double ax = ((double)ui - cx) / fx, ay = ((double)vi - cy) / fy;
          double zval = d * sqrt(ay * ay + ax * ax + 1.0);
          if (!(zval > (double)FLT_MIN)) continue;
          double qn = sqrt((double)qz * qz + (double)qy * qy + (double)qx * qx);
          double sdf = zval - qn;
          if (!(sdf >= -(double)trunc)) continue;
          double t = sdf / trunc;
          if (t > 1.0) t = 1.0;
          if (t < -1.0) t = -1.0;
          f = (f * wi + t) / (double)(wi + 1);
          wi = wi + 1 < max_w ? wi + 1 : max_w;
        }
        tsdf[i] = f;
        w[i] = wi;
      }
}
