/*
 * cu_kinect_fusion.h -- TSDF volume, integration and mesh entry points of the
 * reference's cu_kinect_fusion_v1.h, backed by libmcx's gfx950 kernels.
 *
 *   cu_voxel_new / cu_voxel_delete          (:187-188, dead in the reference)
 *   _cu_voxel_volumetric_integration        (:215, dead in the reference);
 *        color-less, weights +1 per observation capped at CU_VOXEL_MAX_W_INT
 *   _cu_voxel_extract_depth_and_normal_map  (:221, dead in the reference);
 *        depth + camera-frame normal maps; no colour in the volume, so the
 *        colour map is filled with the reference's default (255, 255, 255)
 *   cu_voxel_run_length_write_header / _read_header / _encode / _decode
 *        (cu_kinect_fusion_v1.cu:1303-1600, dead in the reference) the
 *        `.vol` file format for the 4-byte voxel
 *   cu_voxel_to_mesh(short2*, ...)          (:245) device volume -> PLY/OBJ
 *   cu_voxel_rlc_to_mesh(char*, char*, int) (:243) text volume -> PLY/OBJ
 * Camera model as cu_geometry_functions.h:96-129: intrinsic focal lengths and
 * principal point; extrinsic R (columns p[0..2]) and t with
 * q = R^T (p - t) mapping a world point into the camera.
 */
#ifndef MARCHING_CUBE_AMD_CU_KINECT_FUSION_H
#define MARCHING_CUBE_AMD_CU_KINECT_FUSION_H

#include <cstdio>

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

// Ray-cast a depth map (metres, 0 = no surface) and camera-frame normals of
// v seen from (ci, ce) into device arrays of size.x * size.y; c_ptr (may be
// null) is filled with (255, 255, 255).  Voxel i sits at (i + 0.5) grid_s,
// as _cu_voxel_volumetric_integration places it.  Nothing is written when
// K R^T is singular (as the reference, which skips the launch).
void _cu_voxel_extract_depth_and_normal_map(cu_voxel v, float_camera_intrinsic ci,
                                            float_camera_extrinsic ce, int2 size, float* d_ptr,
                                            float3* n_ptr, uchar3* c_ptr);

// `.vol` files: a 40-byte header then per z slice a uint32 byte count and
// the slice's run-length records (GPU codec of libmcx).
void cu_voxel_run_length_write_header(cu_voxel* v, FILE* fp);
void cu_voxel_run_length_read_header(cu_voxel* v, FILE* fp);
// A missing file receives v_local whole; in an existing one the region
// [begin_global, begin_global + size) is replaced by v_local's region at
// begin_local (same grid_s, inside both volumes).
void cu_voxel_run_length_encode(cu_voxel* v_local, char* file_name, int3 begin_local,
                                int3 begin_global, int3 size);
// v_local->data == nullptr: read the whole file into a new device volume
// (v_local's fields set from the header); else copy the file's region
// [begin_global, +size) into v_local at begin_local.
void cu_voxel_run_length_decode(cu_voxel* v_local, char* file_name, int3 begin_local,
                                int3 begin_global, int3 size);

void cu_voxel_to_mesh(short2* _local_volume_data, unsigned int _size_x, unsigned int _size_y,
                      unsigned int _size_z, float _grid_s, const char* file_output);
void cu_voxel_rlc_to_mesh(char* file_input, char* file_output, int volume_dim);

#endif
