/*
 * cu_kinect_fusion.h -- TSDF volume, integration and mesh entry points of the
 * reference's cu_kinect_fusion_v1.h, backed by libmcx's gfx950 kernels.
 *
 *   cu_voxel_new / cu_voxel_delete          (:187-188, dead in the reference)
 *   _cu_voxel_volumetric_integration        (:215, dead in the reference);
 *        color-less, weights +1 per observation capped at CU_VOXEL_MAX_W_INT
 *   cu_voxel_to_mesh(short2*, ...)          (:245) device volume -> PLY/OBJ
 *   cu_voxel_rlc_to_mesh(char*, char*, int) (:243) text volume -> PLY/OBJ
 * Camera model as cu_geometry_functions.h:96-129: intrinsic focal lengths and
 * principal point; extrinsic R (columns p[0..2]) and t with
 * q = R^T (p - t) mapping a world point into the camera.
 */
#ifndef MARCHING_CUBE_AMD_CU_KINECT_FUSION_H
#define MARCHING_CUBE_AMD_CU_KINECT_FUSION_H

#include "structure.h"

#define CU_VOXEL_MAX_W 0xffff
#define CU_VOXEL_MAX_TSDF 0x7ffe
#define CU_VOXEL_MAX_W_INT 128  // integer weight cap of the integration

typedef struct _float_camera_intrinsic {
  float2 fc;     // focal lengths
  float2 cc;     // principal point
  float dc[5];   // distortion (unused: frames are undistorted)
} float_camera_intrinsic;

typedef struct _float33 {
  float3 p[3];   // columns
} float33;

typedef struct _float_camera_extrinsic {
  float33 R;
  float3 t;
} float_camera_extrinsic;

void cu_voxel_new(cu_voxel* v, int3 size, float min_t, float max_t, float grid_s);
void cu_voxel_delete(cu_voxel* v);

// Fuse one float depth frame (metres, device pointer, size.x * size.y) taken
// with (ci, ce) into v.  d_normal_ptr, d_color_ptr and is_dynamic are
// accepted for signature compatibility and ignored.
void _cu_voxel_volumetric_integration(cu_voxel v, float_camera_intrinsic ci,
                                      float_camera_extrinsic ce, int2 size, float* d_depth_ptr,
                                      float3* d_normal_ptr, uchar3* d_color_ptr, bool is_dynamic);

void cu_voxel_to_mesh(short2* _local_volume_data, unsigned int _size_x, unsigned int _size_y,
                      unsigned int _size_z, float _grid_s, const char* file_output);
void cu_voxel_rlc_to_mesh(char* file_input, char* file_output, int volume_dim);

#endif
