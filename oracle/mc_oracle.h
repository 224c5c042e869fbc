/*
 * mc_oracle.h -- CPU oracle for the TSDF -> mesh path (TEST INFRASTRUCTURE).
 *
 * This is the checker, not the product: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  It restates, in plain C, the
 * algorithm of the reference's GPU marching cubes (cu_marching_cubes_v1.cu)
 * and the live normals kernel (cu_kinect_fusion_v1.cu:3146-3242), with each
 * function citing the reference lines it follows.  It also serves as the
 * CPU implementation behind the reference's declared-but-unimplemented
 * MyMarchingCube API (MyMarchingCube.h:4-10), timed as the CPU baseline.
 *
 * Output order is canonical: cubes in ascending linear voxel id
 * (z*X*Y + y*X + x), and within a cube the reference's triangle slot order with
 * its reversed winding (out[off+i+2]=triTable[i], +1=[i+1], +0=[i+2];
 * cu_marching_cubes_v1.cu:319-332).
 *
 * Built with -ffp-contract=off; where the reference's nvcc build contracts
 * a*b+c into a fused multiply-add (vertex_interp, :256-264) the oracle calls
 * fmaf explicitly, so the fp32 sequence is fixed and written down.
 */
#ifndef MC_ORACLE_H
#define MC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* voxel element kinds (same values as the product's MCX_ELEM_*) */
#define ORC_ELEM_SHORT2 0 /* cu_voxel_t {int16 tsdf; uint16 w} (structure.h:14-18) */
#define ORC_ELEM_F32 1    /* plain fp32 field, no weights */
#define ORC_ELEM_F64 2    /* plain FP64 field (finite-difference checks only) */

/* vertex_voxel (structure.h:29-35): 28 bytes */
typedef struct {
  float x, y, z;
  int32_t vx, vy, vz;
  int16_t tsdf;
  int16_t w;
} orc_vertex_voxel;

/* Table accessors (test support). */
const int* orc_tables_tri(void);        /* [256*16] */
const int* orc_tables_num_verts(void);  /* [256] */
const int* orc_tables_edge(void);       /* [256] */

/* Cube index of cube (x,y,z); cu_marching_cubes_v1.cu:70-94.  Fills f[8]
 * with the unpacked corner values read before any early exit. */
int orc_cube_index(const void* vol, int elem, int X, int Y, int Z, int x, int y,
                   int z, float f[8]);

/* Dense per-voxel classification: ci[i], nv[i] for the cube whose min corner
 * is voxel i; 0 where the cube does not exist (x=X-1, y=Y-1 or z=Z-1).
 * cu_marching_cubes_v1.cu:120-131.  Returns the number of active cubes. */
int64_t orc_classify(const void* vol, int elem, int X, int Y, int Z, uint8_t* ci,
                     uint8_t* nv);

/* fp32 extraction with the reference formulas (:228-367).  Writes records
 * with index < cap into any non-NULL output; ver_voxel (N int16) is zeroed
 * then scattered as at :354-358.  Returns the total vertex count.
 * nthreads > 1 parallelises over z planes with OpenMP (same output). */
int64_t orc_extract(const void* vol, int elem, int X, int Y, int Z, float h,
                    orc_vertex_voxel* vv, float* pos, int32_t* v2vox,
                    int16_t* ver_voxel, int64_t cap, int nthreads);

/* Same topology and order, positions evaluated in FP64 (pos: 3 doubles per
 * vertex).  Also returns, per vertex, the generating cube id and edge id. */
int64_t orc_extract_f64(const void* vol, int elem, int X, int Y, int Z, double h,
                        double* pos, int64_t* cube, int8_t* edge, int64_t cap);

/* Per-vertex normals (cu_kinect_fusion_v1.cu:3146-3235) for short2 volumes. */
void orc_normals(const void* vol, int X, int Y, int Z, float h, const float* pos,
                 int64_t stride_floats, int64_t n, float* normals);

/* Exact FP64 gradient of sum(g . p) over the extracted vertices with respect
 * to the fp32 field (elem F32), for the extraction above: for each vertex on
 * edge (a->b), p = p_a + t (p_b - p_a), t = -f_a / (f_b - f_a + 1e-15).
 * grad (N doubles) is overwritten. */
void orc_extract_backward_f64(const float* vol, int X, int Y, int Z, double h,
                              const double* gpos, int64_t nverts, double* grad);

/* TSDF volume init + projective integration (see integrate_oracle.c). */
void orc_volume_init(uint32_t* vol, int64_t n);
void orc_integrate_tsdf(float* tsdf, int32_t* w, int X, int Y, int Z, float h, float trunc,
                        int max_w, const float* intr, const float* poses, const float* depth,
                        float depth_scale, int W, int H, int nframes);
void orc_integrate_tsdf_f64(double* tsdf, int32_t* w, int X, int Y, int Z, float h, float trunc,
                            int max_w, const float* intr, const float* poses,
                            const double* depth, double depth_scale, int W, int H, int nframes);
void orc_integrate_ex(uint32_t* vol, int X, int Y, int Z, float h, float trunc_neg,
                      float trunc_pos, int max_w, const float* intr, const float* poses,
                      const void* depth, int depth_f32, float depth_scale, int W, int H,
                      int nframes);
void orc_integrate(uint32_t* vol, int X, int Y, int Z, float h, float trunc,
                   int max_w, const float* intr, const float* poses,
                   const uint16_t* depth, float depth_scale, int W, int H,
                   int nframes);

/* Ray-marched depth and normal maps (see raycast_oracle.c). */
int orc_raycast_setup(const float* intr, const float* pose, float* im);
int orc_raycast(const uint32_t* vol, int X, int Y, int Z, float h, float origin,
                const float* intr, const float* pose, int W, int H, float* depth,
                float* normals, int nthreads);

/* Run-length `.vol` codec (see rle_oracle.c). */
int64_t orc_rle_encode_slice(const uint32_t* vox, int64_t n, uint8_t* out);
int orc_rle_decode_slice(const uint8_t* in, int64_t nbytes, uint32_t* vox, int64_t n);
int64_t orc_rle_encode_dump(const uint32_t* vol, int X, int Y, int z0, int nz, uint8_t* out);
int orc_rle_decode_dump(const uint8_t* dump, int64_t nbytes, int X, int Y, int z0, int nz,
                        uint32_t* out);

#ifdef __cplusplus
}
#endif
#endif
