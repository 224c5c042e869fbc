/*
 * cu_marching_cubes.h -- class MarchingCubes, the reference's host API
 * (cu_marching_cubes_v1.h:28-65), backed by libmcx's gfx950 kernels.
 *
 * run() extracts the iso-0 surface of a short2 TSDF volume into
 * triangles_buffer (3 consecutive vertex_voxel records per triangle, in
 * canonical order) and fills vertexindex_to_voxel and ver_voxel; empty
 * buffers are created (DEFAULT_TRIANGLES_BUFFER_SIZE, X*Y*Z); the return value
 * is a non-owning view of the first total-vertex records of triangles_buffer,
 * empty for an empty mesh.  If the buffers are too small the reference
 * silently drops cubes; here the call writes the first complete triangles
 * that fit, prints an error and returns them.
 */
#ifndef MARCHING_CUBE_AMD_CU_MARCHING_CUBES_H
#define MARCHING_CUBE_AMD_CU_MARCHING_CUBES_H

#include "cu_memory.hpp"
#include "structure.h"

class MarchingCubes {
 public:
  enum { POINTS_PER_TRIANGLE = 3, DEFAULT_TRIANGLES_BUFFER_SIZE = 9 * 1000 * 1000 * POINTS_PER_TRIANGLE };

  MarchingCubes();
  ~MarchingCubes();

  cu_array<vertex_voxel> run(cu_voxel& tsdf, cu_array<vertex_voxel>& triangles_buffer,
                             cu_array<int3>& vertexindex_to_voxel, cu_array<short>& ver_voxel);

  // Stream for the kernels (default: the null stream).
  void set_stream(hipStream_t s) { stream_ = s; }
  // Vertices of the last run() before clamping to the buffer capacity.
  long long last_total_vertices() const { return last_total_; }

 private:
  cu_array<unsigned char> workspace_;
  hipStream_t stream_ = nullptr;
  long long last_total_ = 0;
};

#endif
