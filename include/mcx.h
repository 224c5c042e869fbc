/*
 * mcx.h -- C ABI of the gfx950 TSDF -> mesh library (libmcx.so).
 *
 * Plain C: pointers, sizes and status codes only.  Device pointers are HIP
 * device addresses; `stream` is a hipStream_t passed as void* (NULL = the
 * null stream).  Every call is stream-ordered and re-entrant: all mutable
 * state lives in caller-owned workspace memory, never in device globals
 * (unlike global_count/output_count/blocks_done in the reference,
 * cu_marching_cubes_v1.cu:49-51).  Calls that share one workspace must be
 * ordered on one stream.
 *
 * Status codes replace the reference's abort-on-error (checkCudaErrors,
 * cu_device.h:45-151) and its silent overflow clamp (:158-161, :176).
 */
#ifndef MCX_H
#define MCX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define MCX_OK 0
#define MCX_ERR_ARG 1         /* bad argument (null pointer, size, kind) */
#define MCX_ERR_CAPACITY 2    /* output capacity too small; *count holds the
                                 required number of vertices */
#define MCX_ERR_WORKSPACE 3   /* workspace too small or not initialised */
#define MCX_ERR_HIP 4         /* a HIP runtime call failed */
#define MCX_ERR_INTERNAL 5    /* device-side consistency check tripped */
#define MCX_ERR_UNSUPPORTED 6 /* combination not supported (e.g. too wide) */
#define MCX_ERR_FORMAT 7      /* malformed input data (e.g. a .vol dump) */

/* ---- voxel kinds ------------------------------------------------------- */
/* cu_voxel_t / short2 {int16 tsdf; uint16 w}, [z][y][x], x fastest
 * (structure.h:14-18, 37-39); f = tsdf / 32766 (cu_kinect_fusion_v1.h:253-256),
 * a cube with any zero weight is empty (cu_marching_cubes_v1.cu:70-94). */
#define MCX_ELEM_SHORT2 0
/* plain fp32 field, same layout, no weights (differentiable form). */
#define MCX_ELEM_F32 1

/* vertex_voxel (structure.h:29-35), 28 bytes: ver_pos, vox_pos, tsdf, w. */
typedef struct {
  float x, y, z;
  int32_t vx, vy, vz;
  int16_t tsdf;
  int16_t w;
} mcx_vertex_voxel;

/* ---- library ----------------------------------------------------------- */
int mcx_version(void);
const char* mcx_strerror(int status);
/* Host copy of the packed case table used by the kernels (host-only call):
 * case_table[c] = 15 x 4-bit edge ids | (#triangles << 60); num_verts[c] =
 * 3 * #triangles.  Re-encodes cu_marching_cubes_v1.cu:509-1028. */
int mcx_get_tables(uint64_t* case_table, uint8_t* num_verts);

/* ---- workspace --------------------------------------------------------- */
/* Bytes of workspace that mcx_extract needs for an X*Y*Z volume: the tile
 * tables plus scratch for one 36-byte record per active cube, sized for one
 * active cube in 16.  _ex sizes the record scratch explicitly; tiles whose
 * records do not fit are re-read from the volume instead (same output). */
size_t mcx_extract_workspace_size(int X, int Y, int Z);
size_t mcx_extract_workspace_size_ex(int X, int Y, int Z, int64_t records);
/* Zero a workspace once before its first use (stream-ordered fill kernel). */
int mcx_workspace_init(void* d_ws, size_t ws_bytes, void* stream);
/* Device-side status of the extraction calls made on a workspace since the
 * last report (no counterpart in the reference, whose kernels cannot fail
 * this way).  The scan's look-back waits are bounded; a wait that reaches its
 * bound (a broken invariant, not expected in normal operation) marks the
 * workspace and makes the call publish -1 as its vertex count (*d_count, and
 * every d_vbase entry of a batched call).  Synchronises the stream, clears
 * the mark and returns MCX_ERR_INTERNAL if it was set, else MCX_OK.  Calls
 * made with h_count report (and clear) the mark themselves.  Meant for
 * callers that pass no h_count: asynchronous loops and captured graphs. */
int mcx_workspace_status(void* d_ws, size_t ws_bytes, void* stream);

/* ---- extraction (replaces MarchingCubes::run, cu_marching_cubes_v1.cu:417-467:
 *      getOccupiedVoxels :185-206, computeOffsetsAndTotalVertexes :208-225,
 *      generateTriangles :372-396; three kernels: count, scan, emit) -------
 * Outputs (each optional, NULL = not written):
 *   d_vv        vertex_voxel records (SHORT2 only)
 *   d_pos       float3 positions
 *   d_v2vox     int3 vertex->voxel map (vertexindex_to_voxel)
 *   d_ver_voxel int16 per voxel, zeroed then scattered (SHORT2 only)
 * Vertices are in canonical order (ascending cube linear id, then the
 * reference's reversed-winding slot order); entries with index >= cap are
 * not written.  d_count (device int64, optional) receives the total.
 * h_count (host, optional): the call synchronises the stream, stores the
 * total and returns MCX_ERR_CAPACITY if it exceeds cap.  Sizes: X*Y + X +
 * 1024 < 2^30 and Z <= 2^24, else MCX_ERR_UNSUPPORTED (also for batched
 * plans, per volume). */
int mcx_extract(const void* d_vol, int elem, int X, int Y, int Z, float h,
                mcx_vertex_voxel* d_vv, float* d_pos, int32_t* d_v2vox,
                int16_t* d_ver_voxel, int64_t cap, int64_t* d_count,
                int64_t* h_count, void* d_ws, size_t ws_bytes, void* stream);

/* mcx_extract with one more optional output: d_edge_key (int64 per vertex) =
 * the grid edge the vertex lies on, as (lower corner voxel linear id) * 3 +
 * axis (0 x, 1 y, 2 z); the up to 4 cubes around an edge give it the same
 * key (input of mcx_weld). */
int mcx_extract_ex(const void* d_vol, int elem, int X, int Y, int Z, float h,
                   mcx_vertex_voxel* d_vv, float* d_pos, int32_t* d_v2vox,
                   int16_t* d_ver_voxel, int64_t* d_edge_key, int64_t cap, int64_t* d_count,
                   int64_t* h_count, void* d_ws, size_t ws_bytes, void* stream);

/* ---- extraction without allocation (no counterpart in the reference) -----
 * For callers that cannot allocate device memory: code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc.
 * mcx_extract_static = mcx_extract without a workspace argument: the
 * library's own device array of 64 MiB per GPU (MCX_STATIC_SCRATCH_BYTES,
 * loaded with the library's code, held until the process ends) is the
 * workspace.  The call allocates nothing and clears the array's counters
 * with a fill kernel.  mcx_extract_static_limit(X, Y, Z) is 0 for volumes
 * whose tile tables leave no room for 65,536 records in the array (those
 * calls return MCX_ERR_UNSUPPORTED; 512^3 fits), else the array's size;
 * tiles whose records do not fit the rest are re-read from the volume (same
 * output, slower).  Every call on a GPU shares that GPU's array and leaves
 * its contents behind: calls that overlap on one GPU (other streams, other
 * threads) race, calls ordered on one stream do not.  A device-side error
 * shows as -1 in *d_count (MCX_ERR_INTERNAL when h_count is passed). */
size_t mcx_extract_static_limit(int X, int Y, int Z);
int mcx_extract_static(const void* d_vol, int elem, int X, int Y, int Z, float h,
                       mcx_vertex_voxel* d_vv, float* d_pos, int32_t* d_v2vox,
                       int16_t* d_ver_voxel, int64_t cap, int64_t* d_count, int64_t* h_count,
                       void* stream);

/* Per-cube classification (the per-cube part of OccupiedVoxels,
 * cu_marching_cubes_v1.cu:120-131): d_ci / d_nv (X*Y*Z bytes each, indexed by
 * the cube's min-corner voxel; 0 for cubes that do not exist). */
int mcx_classify(const void* d_vol, int elem, int X, int Y, int Z,
                 uint8_t* d_ci, uint8_t* d_nv, void* stream);

/* ---- batched extraction (no counterpart in the reference) --------------
 * A plan describes nvol volumes (dims_xyz: X,Y,Z per volume; h: grid size per
 * volume; d_vols: device pointer per volume; all of one elem kind).  It is
 * built once (device allocation + upload) and reused.  mcx_extract_batched
 * extracts every volume in one launch sequence (count, scan, emit) into one
 * concatenated output: volume b's vertices are [d_vbase[b], d_vbase[b+1]),
 * d_vbase has nvol+1 entries, vox_pos is relative to the own volume.  It
 * only launches kernels (no allocation, no memset), so it can be captured in
 * a hipGraph.  ver_voxel is not produced in batched form. */
typedef struct mcx_batch mcx_batch;
int mcx_batch_create(mcx_batch** out, int nvol, const int* dims_xyz, const float* h,
                     const void* const* d_vols, int elem);
void mcx_batch_destroy(mcx_batch* plan);
int64_t mcx_batch_tiles(const mcx_batch* plan);
/* records < 0: default record scratch (one per 16 cubes) */
size_t mcx_batch_workspace_size(const mcx_batch* plan, int64_t records);
int mcx_extract_batched(const mcx_batch* plan, mcx_vertex_voxel* d_vv, float* d_pos,
                        int32_t* d_v2vox, int64_t cap, int64_t* d_vbase, int64_t* d_count,
                        int64_t* h_count, void* d_ws, size_t ws_bytes, void* stream);

/* Voxels of all volumes of a plan (the length of a batch's concatenated
 * per-voxel arrays: volume b's voxels follow those of volumes 0..b-1). */
int64_t mcx_batch_voxels(const mcx_batch* plan);

/* ---- batched differentiable fusion (no counterpart in the reference) ----
 * A plan of fp32 volumes (elem MCX_ELEM_F32) whose fp32 TSDF lives in one
 * concatenated array d_tsdf (volume b at d_tsdf + the voxels of volumes
 * 0..b-1; d_vols[b] of mcx_batch_create must point there), with int32
 * weights d_weight in the same layout.  Every volume is fused from K depth
 * frames of W x H of its own (d_depth [nvol][K][H][W] floats, metres =
 * value * depth_scale; d_poses [nvol][K][12]) with one set of intrinsics,
 * the update of mcx_integrate_tsdf.  Each call is one launch sequence over
 * all volumes and only launches kernels (capturable in a hipGraph).
 * mcx_integrate_tsdf_batched: from_init != 0 starts every voxel from
 *   (init_value, 0) without reading the arrays (else from their contents).
 *   K <= 512.  Workspace: mcx_integrate_workspace_size(W, H).
 * mcx_extract_diff_batched: mcx_extract_diff of every volume (float3
 *   positions, vertex ranges in d_vbase as mcx_extract_batched), keeping in
 *   the workspace what the backward needs.
 * mcx_extract_backward_batched: dL/dtsdf (d_grad_vol, the concatenated
 *   layout) from dL/dpos, with the workspace of the preceding
 *   mcx_extract_diff_batched of this plan and d_scratch of
 *   mcx_batch_backward_scratch_size bytes; d_maxabs (optional, device
 *   uint32) receives max |dL/dtsdf| as float bits.  Fixed-order sums.
 * mcx_integrate_tsdf_backward_batched: dL/d(depth value) [nvol][K][H][W]
 *   from dL/dtsdf, as mcx_integrate_tsdf_backward per volume; d_weight0: the
 *   weights before the fused frames (NULL: all 0); d_maxabs: max |dL/dtsdf|
 *   bits from the previous call, or NULL to compute it.  Workspace:
 *   mcx_integrate_tsdf_backward_batched_workspace_size. */
int mcx_integrate_tsdf_batched(const mcx_batch* plan, float* d_tsdf, int32_t* d_weight,
                               int from_init, float init_value, float trunc, int max_w,
                               const float* intr, const float* d_poses, const float* d_depth,
                               float depth_scale, int W, int H, int K, void* d_ws,
                               size_t ws_bytes, void* stream);
int mcx_extract_diff_batched(const mcx_batch* plan, float* d_pos, int64_t cap, int64_t* d_vbase,
                             int64_t* d_count, int64_t* h_count, void* d_ws, size_t ws_bytes,
                             void* stream);
size_t mcx_batch_backward_scratch_size(const mcx_batch* plan, size_t ws_bytes);
int mcx_extract_backward_batched(const mcx_batch* plan, const float* d_grad_pos,
                                 float* d_grad_vol, uint32_t* d_maxabs, void* d_ws,
                                 size_t ws_bytes, void* d_scratch, size_t scratch_bytes,
                                 void* stream);
size_t mcx_integrate_tsdf_backward_batched_workspace_size(const mcx_batch* plan, int W, int H,
                                                          int K);
int mcx_integrate_tsdf_backward_batched(const mcx_batch* plan, const float* d_grad_tsdf,
                                        const int32_t* d_weight0, const uint32_t* d_maxabs,
                                        float trunc, int max_w, const float* intr,
                                        const float* d_poses, const float* d_depth,
                                        float depth_scale, int W, int H, int K,
                                        float* d_grad_depth, void* d_ws, size_t ws_bytes,
                                        void* stream);
/* Chamfer-style loss of a batch of meshes: vertex v of volume b (rows
 * [d_vbase[b], d_vbase[b+1]) of d_pos, float3) to the nearest of volume b's
 * M targets (d_targets [nvol][M][3]): d_loss[b] = sum of squared distances
 * (fixed summation order), d_grad_pos = 2 (p - nearest) per row (0 for rows
 * past the total, up to cap rows).  M <= 1024.  Workspace:
 * mcx_chamfer_workspace_size(cap). */
size_t mcx_chamfer_workspace_size(int64_t cap);
int mcx_chamfer_batched(const float* d_pos, const int64_t* d_vbase, int nvol,
                        const float* d_targets, int M, int64_t cap, float* d_loss,
                        float* d_grad_pos, void* d_ws, size_t ws_bytes, void* stream);

/* ---- indexed mesh (replaces the OBJ writer's de-duplication by printed
 *      coordinates, cu_kinect_fusion_v1.cu:3668-3719, with exact sharing by
 *      grid edge).  Input: the V vertices of an extraction (positions with a
 *      stride of stride_floats floats: 7 for vertex_voxel, 3 for float3) and
 *      their edge keys (mcx_extract_ex).  Output: the unique vertices in order
 *      of first occurrence (d_upos, float3; each takes the position of its
 *      first occurrence), d_indices (int32 per input vertex: its unique
 *      vertex; consecutive triples are the triangles) and the unique count
 *      (d_nunique device int64 and/or h_nunique host, which synchronises).
 *      Deterministic.  V < 2^31.  Keys may be any int64 (equal keys weld;
 *      extraction keys are >= 0).  Workspace: mcx_weld_workspace_size(V). */
size_t mcx_weld_workspace_size(int64_t V);
int mcx_weld(const int64_t* d_edge_key, const float* d_pos, int64_t stride_floats, int64_t V,
             float* d_upos, int32_t* d_indices, int64_t* d_nunique, int64_t* h_nunique,
             void* d_ws, size_t ws_bytes, void* stream);

/* ---- per-vertex normals (replaces d_voxel_extract_surface_normals,
 *      cu_kinect_fusion_v1.cu:3146-3238).  d_pos: n positions with a stride
 *      of stride_floats floats (7 for vertex_voxel records, 3 for float3);
 *      d_normals: n float3.  SHORT2 volumes. */
int mcx_normals(const void* d_vol, int X, int Y, int Z, float h, const float* d_pos,
                int64_t stride_floats, int64_t n, float* d_normals, void* stream);

/* ---- differentiable extraction (no counterpart in the reference) -------
 * mcx_extract_diff: extraction of an fp32 field (d_pos: float3 positions, same
 * order and values as mcx_extract with d_pos) that also keeps in the
 * workspace what the backward needs (active-cube bits and records).  With
 * h_count it returns MCX_ERR_WORKSPACE when the record scratch was too small
 * for some column (grow the workspace and call again).
 * mcx_extract_backward: d_grad_vol[i] = sum over the vertices on edges at
 * voxel i of dL/dp . dp/df_i, from d_grad_pos (float3 per vertex), using the
 * workspace of the preceding mcx_extract_diff call on the same volume
 * (unchanged) and d_scratch of mcx_backward_scratch_size() bytes.
 * Deterministic: fixed-order sums, no atomics. */
int mcx_extract_diff(const float* d_vol, int X, int Y, int Z, float h, float* d_pos,
                     int64_t cap, int64_t* d_count, int64_t* h_count, void* d_ws,
                     size_t ws_bytes, void* stream);
size_t mcx_backward_scratch_size(int X, int Y, int Z, size_t ws_bytes);
int mcx_extract_backward(const float* d_vol, int X, int Y, int Z, float h,
                         const float* d_grad_pos, float* d_grad_vol, void* d_ws,
                         size_t ws_bytes, void* d_scratch, size_t scratch_bytes,
                         void* stream);
/* Same result without relying on the workspace contents: the per-cube
 * records of the forward are recomputed from d_vol (the count and scan
 * kernels again, no emission) before the backward kernels run.  The
 * workspace can then serve other extractions between the forward and the
 * backward (records are not kept alive per forward).  d_ws must have at least
 * the size that the forward's mcx_extract_diff call checked (so no record
 * scratch overflows); its contents are overwritten. */
int mcx_extract_backward_recompute(const float* d_vol, int X, int Y, int Z, float h,
                                   const float* d_grad_pos, float* d_grad_vol, void* d_ws,
                                   size_t ws_bytes, void* d_scratch, size_t scratch_bytes,
                                   void* stream);

/* ---- TSDF volume (replaces cu_voxel_new / d_voxel_initialize,
 *      cu_kinect_fusion_v1.cu:364-398): every voxel = (tsdf -32766, w 0). */
int mcx_volume_init(void* d_vol, int X, int Y, int Z, void* stream);

/* ---- projective integration (replaces _cu_voxel_volumetric_integration,
 *      cu_kinect_fusion_v1.cu:534-549, and its kernels :416-532).
 * d_vol: packed short2 volume, voxel i at ((i+0.5)h) per axis.
 * intr (host): {fx, fy, cx, cy}.  d_poses: K x 12 floats, camera-to-world
 * {R row-major, t}.  d_depth: K x H x W uint16, metres = value * depth_scale,
 * 0 = no measurement.  Frames are fused in order, all K in one sweep of the
 * volume.  Workspace: mcx_integrate_workspace_size(W, H) bytes (per-pixel ray
 * factors, recomputed by every call).  Frames of W * H * 4 >= 2^31 bytes:
 * MCX_ERR_UNSUPPORTED. */
size_t mcx_integrate_workspace_size(int W, int H);
int mcx_integrate(void* d_vol, int X, int Y, int Z, float h, float trunc, int max_w,
                  const float* intr, const float* d_poses, const uint16_t* d_depth,
                  float depth_scale, int W, int H, int K, void* d_ws, size_t ws_bytes,
                  void* stream);
/* Same with float depth frames (metres = value * depth_scale), the form the
 * reference's integration takes (float *d_depth_ptr, :534). */
int mcx_integrate_f32(void* d_vol, int X, int Y, int Z, float h, float trunc, int max_w,
                      const float* intr, const float* d_poses, const float* d_depth,
                      float depth_scale, int W, int H, int K, void* d_ws, size_t ws_bytes,
                      void* stream);
/* Both with the reference's two truncation distances (cu_kinect_fusion_v1.cu:
 * 455-463, the volume's min_t and max_t): a voxel is skipped where sdf <
 * -trunc_neg, and t = sdf > 0 ? min(sdf / trunc_pos, 1) : sdf / trunc_neg
 * (equal distances give the update above, bit for bit).  depth_kind:
 * MCX_DEPTH_U16 (uint16 frames) or MCX_DEPTH_F32 (float frames). */
#define MCX_DEPTH_U16 0
#define MCX_DEPTH_F32 1
int mcx_integrate_ex(void* d_vol, int X, int Y, int Z, float h, float trunc_neg, float trunc_pos,
                     int max_w, const float* intr, const float* d_poses, const void* d_depth,
                     int depth_kind, float depth_scale, int W, int H, int K, void* d_ws,
                     size_t ws_bytes, void* stream);

/* ---- fp32 TSDF volumes, differentiable integration (no counterpart in the
 *      reference).  tsdf (float) and weight (int32) in separate [z][y][x]
 *      arrays; the observation of mcx_integrate (same projection, ray-distance
 *      sdf, skips and weight cap), stored unpacked:
 *        f <- fma(f, w, t) / (w + 1),  w <- min(w + 1, max_w),
 *        t = clamp(sdf / trunc, -1, 1),
 *      with float depth frames (metres = value * depth_scale).
 * mcx_tsdf_init: every voxel = (tsdf_value, 0).
 * mcx_integrate_tsdf: fuse K frames (workspace: mcx_integrate_workspace_size).
 * mcx_integrate_tsdf_backward: d_grad_depth (K x H x W) = dL/d(depth value)
 *   from d_grad_tsdf = dL/df after the call that fused these K frames into a
 *   volume whose weights before that call were d_weight0 (NULL: all 0).  Per
 *   voxel the observations are recomputed; each one's df/dt is the closed form
 *   of the capped running average; dt/dd is 0 where t is clamped.  Sums over
 *   voxels use int64 fixed-point atomics (exact, independent of order), so the
 *   result is deterministic.  K <= 512.  Workspace:
 *   mcx_integrate_tsdf_backward_workspace_size(W, H, K) bytes. */
int mcx_tsdf_init(float* d_tsdf, int32_t* d_weight, int64_t n, float tsdf_value, void* stream);
int mcx_integrate_tsdf(float* d_tsdf, int32_t* d_weight, int X, int Y, int Z, float h,
                       float trunc, int max_w, const float* intr, const float* d_poses,
                       const float* d_depth, float depth_scale, int W, int H, int K,
                       void* d_ws, size_t ws_bytes, void* stream);
size_t mcx_integrate_tsdf_backward_workspace_size(int W, int H, int K);
int mcx_integrate_tsdf_backward(const float* d_grad_tsdf, const int32_t* d_weight0, int X,
                                int Y, int Z, float h, float trunc, int max_w,
                                const float* intr, const float* d_poses, const float* d_depth,
                                float depth_scale, int W, int H, int K, float* d_grad_depth,
                                void* d_ws, size_t ws_bytes, void* stream);

/* ---- raycast: depth and normal maps from a packed short2 volume
 *      (replaces _cu_voxel_extract_depth_and_normal_map,
 *      cu_kinect_fusion_v1.cu:972-986, and its kernels :551-970).
 * Voxel i sits at world (i + origin) * h per axis: origin 0 is the
 * reference's convention, 0.5 that of mcx_integrate's voxel centres.
 * intr (host): {fx, fy, cx, cy}; pose (host): 12 floats, camera-to-world
 * {R row-major, t} in metres.  Ray march as the reference: step 1 voxel,
 * 0.25 inside the band |sdf| < 1, trilinear samples (+-1 at a corner holding
 * +-32766), linear zero-crossing refinement, back faces end a ray.
 * Outputs (device, row-major by pixel): d_depth[H*W] z-depth in metres,
 * 0 where no surface is hit; d_normals[H*W*3] (may be NULL) camera-frame
 * unit normals, 0 where no surface is hit or a corner is +-32766.  Colour is
 * not produced.  MCX_ERR_ARG when K R^T is singular (|det| < FLT_EPSILON;
 * the reference skips the launch).  Every ray ends: a march stops once it
 * has left the volume for good, and after 2^20 steps. */
int mcx_raycast(const void* d_vol, int X, int Y, int Z, float h, float origin,
                const float* intr, const float* pose, int W, int H, float* d_depth,
                float* d_normals, void* stream);

/* ---- `.vol` run-length codec (replaces the slice codec of
 *      cu_voxel_run_length_encode / _decode, cu_kinect_fusion_v1.cu:1362-1600,
 *      for the 4-byte voxel {int16 tsdf; uint16 w}).
 * A dump is, per z slice, a uint32 byte count then the slice's records:
 * u16 {n : 14 (low bits), v : 2 (high bits)}, v = 1 a run of (+32766, w 0),
 * v = 2 a run of (-32766, w 0), v = 3 a run of the 4-byte voxel that follows
 * (v = 0 decodes as v = 3).  Runs are maximal stretches of equal tsdf in
 * [y][x] order (weights not compared; a run keeps its first voxel), at most
 * 0x3fff voxels.  The file adds a 40-byte header in front: "VOLUME_TSDF_RGB\0",
 * int32 size[3], float min_t, max_t, grid_s (:1303-1310).
 * mcx_rle_encode: the dump of slices [z0, z0+nz) of d_vol (X, Y, Z) into
 *   d_dump (2-byte aligned, dump_cap bytes; mcx_rle_encode_bound(X, Y, nz)
 *   always suffices); the size goes to *d_dump_bytes and *h_dump_bytes (each
 *   may be NULL).  MCX_ERR_CAPACITY (nothing written) if it exceeds dump_cap.
 *   Workspace: mcx_rle_encode_workspace_size(X, Y, nz) bytes (16 per voxel:
 *   encode large volumes in slabs).  Synchronises the stream.
 * mcx_rle_decode: nz slices into d_vol[nz][Y][X]; slice j's uint32 byte
 *   count sits at byte h_slice_offsets[j] (host array) of d_dump (2-byte
 *   aligned, dump_bytes).  MCX_ERR_FORMAT for a malformed slice (offset
 *   outside the dump, odd or zero size, a run of length 0, runs past the
 *   slice, a truncated record, bytes left over); d_vol is then undefined.
 *   Workspace: mcx_rle_decode_workspace_size(dump_bytes, X, Y, nz).
 *   Synchronises the stream. */
size_t mcx_rle_encode_bound(int X, int Y, int nz);
size_t mcx_rle_encode_workspace_size(int X, int Y, int nz);
int mcx_rle_encode(const void* d_vol, int X, int Y, int Z, int z0, int nz, void* d_dump,
                   size_t dump_cap, int64_t* d_dump_bytes, int64_t* h_dump_bytes, void* d_ws,
                   size_t ws_bytes, void* stream);
size_t mcx_rle_decode_workspace_size(int64_t dump_bytes, int X, int Y, int nz);
int mcx_rle_decode(const void* d_dump, int64_t dump_bytes, const int64_t* h_slice_offsets, int nz,
                   int X, int Y, void* d_vol, void* d_ws, size_t ws_bytes, void* stream);

/* ---- synthetic depth frames (test/benchmark input; no reference
 *      counterpart: the reference reads OpenNI frames).  Ray-casts nobj
 *      objects (8 floats each: kind 0 sphere {c, r} / 1 box {c, half sizes})
 *      from each pose; z-depth plus Gaussian noise (sigma, metres), rounded
 *      to depth_unit_m; a pixel is dropped (0) with probability dropout. */
int mcx_synth_depth(uint16_t* d_depth, int W, int H, int K, const float* intr,
                    const float* d_poses, const float* d_objs, int nobj, float noise_sigma_m,
                    float dropout, uint32_t seed, float depth_unit_m, void* stream);

/* ---- launch-shape tunables (no counterpart in the reference) -----------
 * Process-wide settings read by every launch on the host (not synchronised:
 * set them before launching, from one thread).  Outputs do not depend on
 * them; only the decomposition of the work does.  marching_cube_amd.autotune
 * times candidates on the caller's volume and sets the fastest. */
#define MCX_TUNE_COLUMNS_PER_CU 0     /* count kernel columns (waves) per CU, single volume, at
                                         most (smaller volumes: about 8192 cubes a column, at
                                         least 4 per CU; a column's planes + 4 stay within 2^31
                                         bytes, so large planes take more, shorter ones); 56 */
#define MCX_TUNE_EMIT_GROUPS_PER_CU 1 /* emit kernel persistent workgroups per CU; 16 */
#define MCX_TUNE_BATCH_SLAB 2         /* planes per column of batched plans (at creation; fewer
                                         where the planes + 4 would pass 2^31 bytes); 64 */
int mcx_set_tuning(int key, int value);
int mcx_get_tuning(int key); /* -1 for an unknown key */

#ifdef __cplusplus
}
#endif
#endif /* MCX_H */
