/*
 * structure.h -- voxel, volume and vertex types of the TSDF -> mesh path.
 * Same layouts as the reference's structure.h:14-39 so existing callers
 * compile against this header unchanged:
 *   cu_voxel_t   {short tsdf; ushort w}        4 bytes, same bits as short2
 *   cu_voxel     {int3 size; int size_xy; float min_t, max_t, grid_s; data}
 *   vertex_voxel {float3 ver_pos; int3 vox_pos; short tsdf; short w}  28 bytes
 * The color field the reference comments out (:17) is absent here too.
 */
#ifndef MARCHING_CUBE_AMD_STRUCTURE_H
#define MARCHING_CUBE_AMD_STRUCTURE_H

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

typedef unsigned char uchar;
typedef unsigned short ushort;

typedef struct _CU_VOXEL_T {
  short tsdf;  // normalized tsdf (f = tsdf / 32766)
  ushort w;    // weight
} cu_voxel_t;

typedef struct _CU_VOXEL {
  int3 size;
  int size_xy;          // size.x * size.y
  float min_t;          // truncation behind the surface, metres
  float max_t;          // truncation in front of the surface, metres
  float grid_s;         // voxel size, metres
  cu_voxel_t* data;     // device pointer, [z][y][x]
} cu_voxel;

typedef struct _VERTEX_VOXEL {
  float3 ver_pos;
  int3 vox_pos;
  short tsdf;
  short w;
} vertex_voxel;

static_assert(sizeof(cu_voxel_t) == 4, "cu_voxel_t is 4 bytes");
static_assert(sizeof(vertex_voxel) == 28, "vertex_voxel is 28 bytes");

#define cu_voxel_access_z(v, _z) ((v)->data + (_z) * (v)->size_xy)
#define cu_voxel_access_yz(v, _y, _z) (cu_voxel_access_z(v, _z) + (_y) * (v)->size.x)
#define cu_voxel_access_xyz(v, _x, _y, _z) (cu_voxel_access_yz(v, _y, _z) + (_x))

#endif
