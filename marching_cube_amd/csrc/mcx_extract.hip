// mcx_extract.hip -- marching-cubes extraction for gfx950 in three kernels.
//
// Replaces the reference's three synchronous stages in MarchingCubes::run
// (cu_marching_cubes_v1.cu:417-467):
//   getOccupiedVoxels              :97-206  classify + atomic compaction
//   computeOffsetsAndTotalVertexes :208-225  thrust exclusive scan + 2 D2H reads
//   generateTriangles              :228-396  per-active-cube triangle emission
//
// K1 count   Each wave owns a column: a range of kTile consecutive in-plane
//            cube ids and a slab of consecutive z planes.  It marches z with
//            the two current planes in registers and the next plane's loads in
//            flight, so the volume is read once.  Classification is bit
//            parallel: per 64 cubes, eight 64-bit ballot masks (corner signs)
//            and eight more (zero weights) give the active cubes with a few
//            scalar ops.  Active cubes get their case and a record (offset,
//            case, 8 corner words) in the column's scratch quota; every tile
//            (kTile cubes of one plane) gets its vertex count, and the tiles
//            with vertices are appended to a list (one atomic per column).
// K2 scan    Exclusive scan of the tile counts (single pass, look-back over
//            blocks of 4096 tiles) -> tile vertex offsets and the total.
// K3 emit    Wave-level emission: persistent waves take tiles from the list
//            (strided by wave); a tile's records are taken 64 at a time, one
//            per lane, their vertex counts scanned across the wave, and each
//            lane produces one output vertex, so neighbouring lanes write
//            neighbouring records.  A tile whose records did not fit in its
//            column's quota is re-read from the volume into the same form.
//
// The output is in canonical order: ascending cube linear id (z*X*Y + y*X + x),
// then the reference's triangle slot order with reversed winding (:319-332).
// Per-cube semantics follow the reference exactly:
//   cube index   computeCubeIndex :70-94 (any zero weight -> 0; bit k = f_k < 0)
//   numVerts     0 for cases 0 and 255, else numVertsTable (:127-131)
//   node coords  (i + 0.5f) * h (getNodeCoo :243-254)
//   interpolate  t = (0 - f0) / (f1 - f0 + 1e-15f); p = fma(t, p1 - p0, p0)
//                (vertex_interp :256-264; nvcc contracts the multiply-add)
//   record       vox = floor(p / h - 0.5f); tsdf/w of that voxel; int3 map;
//                ver_voxel scatter (store_point :336-367)
// A vox outside the volume (the reference reads out of bounds there) yields
// tsdf = w = 0 and no ver_voxel write.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "mc_tables.inc"
#include "mcx_internal.h"
#include "mcx_scan_part.h"

namespace mcx {

struct VolDesc {
  const uint32_t* vol;   // X*Y*Z raw voxels (short2 bits or fp32 bits), [z][y][x]
  int X, Y, Z;
  int R;                 // tiles per plane = ceil(X*Y / kTile)
  float h;
  int64_t tile_begin;    // first tile of this volume in the launch
  int16_t* ver_voxel;    // optional
  FastDiv divX, divR;
};

struct Column {          // one K1 wave: volume, in-plane tile range, planes [z0, z1)
  int vol, r, z0, z1;
};

namespace {

__constant__ unsigned long long dCase[256] = MCX_CASE_TABLE_INIT;
__constant__ unsigned char dNumVerts[256] = MCX_NUM_VERTS_INIT;

constexpr int kCPW = 8;                // 64-cube chunks per wave and plane
constexpr int kTile = 64 * kCPW;       // cubes per tile (one plane of a column)
constexpr int kRecWords = 9;           // record: (offset << 8 | case), 8 corner words
constexpr int kScanPer = 4;            // tiles per scan thread
constexpr int kScanItems = 256 * kScanPer;  // tiles per scan block
constexpr int kEmitThreads = 256;
// Launch-shape tunables (process-wide, set by mcx_set_tuning / the
// autotuner; read by every launch on the host): K1 columns (waves) per CU for
// single volumes, K3 persistent workgroups per CU, planes per column of
// batched plans (taken when a plan is created).
struct Tuning {
  int columns_per_cu = 56;
  int emit_groups_per_cu = 16;
  int batch_slab = 64;
};
Tuning g_tuning;
constexpr int kTilesPerContribGroup = 8;  // backward BW1: tiles per workgroup
// third word of a gfx9 raw buffer resource (stride 0, 32-bit data format)
constexpr int kBufferWord3 = 0x00020000;

// Edge e runs from corner kEa[e] to kEb[e] (:291-302), 3 bits per edge.
constexpr uint64_t kEdgeA = (0ull << 0) | (1ull << 3) | (2ull << 6) | (3ull << 9) | (4ull << 12) |
                            (5ull << 15) | (6ull << 18) | (7ull << 21) | (0ull << 24) |
                            (1ull << 27) | (2ull << 30) | (3ull << 33);
constexpr uint64_t kEdgeB = (1ull << 0) | (2ull << 3) | (3ull << 6) | (0ull << 9) | (5ull << 12) |
                            (6ull << 15) | (7ull << 18) | (4ull << 21) | (4ull << 24) |
                            (5ull << 27) | (6ull << 30) | (7ull << 33);

__device__ __forceinline__ int corner_dx(int c) { return ((c + 1) >> 1) & 1; }
__device__ __forceinline__ int corner_dy(int c) { return (c >> 1) & 1; }
__device__ __forceinline__ int corner_dz(int c) { return c >> 2; }
// corner id of offset (dx, dy, dz)
__device__ __forceinline__ int corner_of(int dx, int dy, int dz) {
  return 4 * dz + (dy ? (dx ? 2 : 3) : dx);
}

// Per case and output slot q: the edge's corners and axis,
// ca | cb << 3 | axis << 6.  Output slot q = 3*tri + r holds table corner
// 2 - r of triangle tri (:324-332); slots past the case's vertex count are 0.
struct SlotTable {
  uint8_t v[256 * 16];
};
constexpr SlotTable make_slot_table() {
  constexpr unsigned long long cs[256] = MCX_CASE_TABLE_INIT;
  SlotTable t{};
  for (int ci = 0; ci < 256; ++ci) {
    for (int q = 0; q < 15; ++q) {
      const int m = 3 * (q / 3) + 2 - (q % 3);
      const int e = (int)((cs[ci] >> (4 * m)) & 0xfu);
      if (e >= 12) continue;
      const int ca = (int)((kEdgeA >> (3 * e)) & 7u), cb = (int)((kEdgeB >> (3 * e)) & 7u);
      const int ax = (((ca + 1) >> 1) & 1) != (((cb + 1) >> 1) & 1) ? 0
                     : ((ca >> 1) & 1) != ((cb >> 1) & 1)          ? 1
                                                                    : 2;
      t.v[ci * 16 + q] = (uint8_t)(ca | (cb << 3) | (ax << 6));
    }
  }
  return t;
}
__constant__ SlotTable dSlot = make_slot_table();


struct ExtractParams {
  VolDesc vol0;          // single-volume launches
  const VolDesc* vols;   // batched launches: [nvol] in device memory
  const BatchVol* bvol;  // batched launches: [nvol] per-voxel layout (backward)
  const Column* cols;    // batched launches: [ncols]
  int nvol;
  int64_t ncols;
  int slab;              // single volume: planes per column
  int slab_max;          // most planes of any column
  bool wide;             // all volumes take the 128-bit count kernel
  int64_t n_tiles;
  int64_t quota;         // records per column
  int64_t ovf_begin;     // shared overflow pool of records (wide count kernel)
  int64_t ovf_cap;
  mcx_vertex_voxel* vv;
  float* pos;
  int32_t* v2vox;
  int64_t* ekey;         // optional: grid edge of each vertex (min-corner voxel * 3 + axis)
  int64_t cap;
  int64_t* d_count;
  int64_t* vbase;
  // workspace
  WsHeader* hdr;
  uint64_t* desc;
  int64_t cap_desc;
  int32_t* tileT;        // vertex count per tile
  int32_t* tile_nrec;    // records per tile, -1 = staged path
  int64_t* tile_rec;     // first record of the tile
  int64_t* voff;         // exclusive vertex offset per tile
  uint64_t* masks;       // active-cube bits, kCPW words per tile
  uint32_t* nz_list;     // ids of the tiles with vertices (in append order)
  // emission in units of 64 records (single volumes): the scan also scans
  // the records of the list's tiles in list order
  bool list_scan;
  uint32_t* clist;       // the list's tiles with records, in list order
  int64_t* lpre;         // [n + 1] first record of clist entry j (lpre[n]: all records)
  int64_t* lrs;          // tile_rec of clist entry j
  uint32_t* units;       // clist entry holding record 64 k, for every unit k
  uint32_t* dense_list;  // the list's tiles that took the staged path
  uint32_t* nz_vol;      // batched launches: volume of each list entry (nz_list order)
  uint32_t* clist_vol;   // batched launches: volume of each clist entry
  bool write_masks;
  int64_t first_block;   // gather: workgroup index of this launch's blockIdx.x 0
  uint32_t* rec;
  int64_t rec_cap;
};

// The fp32 value of a voxel: tsdf / 32766 (tsdf_unit, correctly rounded) or
// the fp32 bits.
template <int ELEM>
__device__ __forceinline__ float unpack_f(uint32_t v) {
  if constexpr (ELEM == MCX_ELEM_SHORT2) {
    return tsdf_unit((int16_t)(v & 0xffffu));
  } else {
    return __uint_as_float(v);
  }
}

template <int ELEM>
__device__ __forceinline__ bool is_neg(uint32_t v) {
  if constexpr (ELEM == MCX_ELEM_SHORT2) {
    return (v & 0x8000u) != 0;  // tsdf < 0  <=>  tsdf / 32766 < 0
  } else {
    return __uint_as_float(v) < 0.0f;
  }
}

template <int ELEM>
__device__ __forceinline__ bool is_w0(uint32_t v) {
  if constexpr (ELEM == MCX_ELEM_SHORT2) return v < 0x10000u;  // weight == 0
  return false;
}

// Cube index from the 8 corner words in the reference corner order.
template <int ELEM>
__device__ __forceinline__ uint32_t cube_index(const uint32_t c[8]) {
  uint32_t ci = 0;
  bool w0 = false;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    ci |= (is_neg<ELEM>(c[k]) ? 1u : 0u) << k;
    w0 = w0 || is_w0<ELEM>(c[k]);
  }
  return w0 ? 0u : ci;
}

template <bool BATCH>
__device__ __forceinline__ const VolDesc& vol_at(const ExtractParams& P, int vi) {
  if constexpr (BATCH) return P.vols[vi];
  return P.vol0;
}

// Volume of global tile t (batched: last volume with tile_begin <= t).
__device__ __forceinline__ int volume_of_tile(const ExtractParams& P, int64_t t) {
  int lo = 0, hi = P.nvol - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (P.vols[mid].tile_begin <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// The scan kernel's look-back descriptors of this call: their place in the
// workspace depends on the volume size, so a workspace reused with other
// sizes holds stale flags there.  Workgroup 0 of the count kernel (which
// completes before the scan starts) zeroes them.
// Zeroes both look-back descriptor sets (tiles, list) and the staged-tile
// list count; the previous call's emit kernel has finished (same stream).
__device__ __forceinline__ void clear_scan_descriptors(const ExtractParams& P) {
  if (blockIdx.x == 0) {
    for (int64_t i = threadIdx.x; i < 3 * P.cap_desc; i += blockDim.x) P.desc[i] = 0;
    if (threadIdx.x == 0) {
      P.hdr->dense_n = 0;
      P.hdr->scan_ticket = 0;
    }
  }
}

// ============================================================ K1: count
template <int ELEM, bool BATCH>
__global__ __launch_bounds__(256) void count_kernel(ExtractParams P) {
  __shared__ uint8_t s_nv[256];
  s_nv[threadIdx.x] = dNumVerts[threadIdx.x];
  clear_scan_descriptors(P);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  // wave index made wave-uniform so that everything derived from the column
  // (plane bases, loop bounds) lives in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t col = (int64_t)blockIdx.x * 4 + wave;
  if (col >= P.ncols) return;  // whole waves exit; no barrier follows

  int vi, r, z0, z1;
  if constexpr (BATCH) {
    const Column c = P.cols[col];
    vi = c.vol;
    r = c.r;
    z0 = c.z0;
    z1 = c.z1;
  } else {
    vi = 0;
    r = (int)(col % P.vol0.R);
    const int s = (int)(col / P.vol0.R);
    z0 = s * P.slab;
    z1 = min(z0 + P.slab, P.vol0.Z - 1);
  }
  const VolDesc& V = vol_at<BATCH>(P, vi);
  const int X = V.X;
  const int64_t XY = (int64_t)X * V.Y;
  const int64_t p0 = (int64_t)r * kTile;
  const int64_t rec_base = col * P.quota;

  // validity of the lane's cube in chunk j (identical for every plane): bit j
  uint32_t vbits = 0;
#pragma unroll
  for (int j = 0; j < kCPW; ++j) {
    const int64_t p = p0 + 64 * j + lane;
    if (p < XY) {
      const uint32_t y = V.divX.div((uint32_t)p);
      const uint32_t x = (uint32_t)p - y * (uint32_t)X;
      if (x + 1 < (uint32_t)X && y + 1 < (uint32_t)V.Y) vbits |= 1u << j;
    }
  }

  // plane q values at in-plane p (U) and p + X (W); chunk kCPW only on lane 0.
  // Raw buffer loads: the resource spans exactly plane q (X*Y words), so a
  // load past the plane returns 0, and every load shares one VGPR offset.
  uint32_t A[kCPW + 1], B[kCPW + 1], C[kCPW + 1], D[kCPW + 1];
  uint32_t E[kCPW + 1] = {}, F[kCPW + 1] = {};
  // Every offset is in the VGPR offset (plus a small immediate): the
  // hardware range-checks those against the resource size, not the SGPR
  // offset.  The host keeps X*Y + X + 1024 words below 2^30, so no offset
  // wraps.
  const uint32_t plane_bytes = (uint32_t)(XY * 4);
  const uint32_t voff = (uint32_t)(p0 + lane) * 4u;
  const uint32_t voffW = voff + 4u * (uint32_t)X;
  // chunk kCPW is only needed on lane 0; other lanes use an offset at the end
  // of the plane, which returns 0 without touching memory
  const uint32_t voff_x = lane == 0 ? voff + 256u * kCPW : plane_bytes;
  const uint32_t voffW_x = lane == 0 ? voffW + 256u * kCPW : plane_bytes;
  // Every load is unconditional (no branch around it), so the compiler can
  // keep the next plane's loads in flight while the current planes are
  // classified.  A plane that is not needed gets an empty resource.
  auto load_plane = [&](int q, bool need, uint32_t* U, uint32_t* W) {
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint32_t*>(V.vol + (need ? (int64_t)q * XY : 0)), (short)0,
        need ? (int)plane_bytes : 0, kBufferWord3);
#pragma unroll
    for (int j = 0; j < kCPW; ++j) {
      U[j] = __builtin_amdgcn_raw_buffer_load_b32(rs, voff + 256u * j, 0, 0);
      W[j] = __builtin_amdgcn_raw_buffer_load_b32(rs, voffW + 256u * j, 0, 0);
    }
    U[kCPW] = __builtin_amdgcn_raw_buffer_load_b32(rs, voff_x, 0, 0);
    W[kCPW] = __builtin_amdgcn_raw_buffer_load_b32(rs, voffW_x, 0, 0);
  };
  load_plane(z0, true, A, B);
  load_plane(z0 + 1, true, C, D);

  // sign and zero-weight masks of one chunk (wave-uniform)
  struct Masks {
    uint64_t n[4], w[4];
  };
  auto masks_of = [&](int j) {
    Masks M;
    M.n[0] = __ballot(is_neg<ELEM>(A[j]));
    M.n[1] = __ballot(is_neg<ELEM>(B[j]));
    M.n[2] = __ballot(is_neg<ELEM>(C[j]));
    M.n[3] = __ballot(is_neg<ELEM>(D[j]));
    if constexpr (ELEM == MCX_ELEM_SHORT2) {
      M.w[0] = __ballot(is_w0<ELEM>(A[j]));
      M.w[1] = __ballot(is_w0<ELEM>(B[j]));
      M.w[2] = __ballot(is_w0<ELEM>(C[j]));
      M.w[3] = __ballot(is_w0<ELEM>(D[j]));
    } else {
      M.w[0] = M.w[1] = M.w[2] = M.w[3] = 0;
    }
    return M;
  };

  int64_t used = 0;
  uint64_t nzmask = 0;  // planes zb + i of this column whose tile has vertices
  int zb = z0;
  for (int z = z0; z < z1; ++z) {
    load_plane(z + 2, z + 1 < z1, E, F);
    const int64_t t = V.tile_begin + (int64_t)z * V.R + r;
    const int64_t room = P.quota - used;  // records this tile may still write

    uint32_t vsum = 0;
    int64_t nrec = 0;
    Masks M0 = masks_of(0);
#pragma unroll
    for (int j = 0; j < kCPW; ++j) {
      const Masks M1 = masks_of(j + 1);
      // x+1 neighbour masks: shift down one lane, bit 63 from the next chunk
      uint64_t all = ~0ull, any = 0, w = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint64_t n1 = (M0.n[q] >> 1) | ((M1.n[q] & 1ull) << 63);
        all &= M0.n[q] & n1;
        any |= M0.n[q] | n1;
        w |= M0.w[q] | (M0.w[q] >> 1) | ((M1.w[q] & 1ull) << 63);
      }
      const uint64_t m = any & ~all & ~w & __ballot((vbits >> j) & 1u);
      M0 = M1;
      if (P.write_masks && lane == 0) P.masks[t * kCPW + j] = m;
      if (m == 0) continue;
      uint32_t a1 = __shfl_down(A[j], 1), b1 = __shfl_down(B[j], 1);
      uint32_t c1 = __shfl_down(C[j], 1), d1 = __shfl_down(D[j], 1);
      const uint32_t na = __shfl(A[j + 1], 0), nb = __shfl(B[j + 1], 0);
      const uint32_t nc = __shfl(C[j + 1], 0), nd = __shfl(D[j + 1], 0);
      if (lane == 63) {
        a1 = na;
        b1 = nb;
        c1 = nc;
        d1 = nd;
      }
      if ((m >> lane) & 1ull) {
        const uint32_t cw[8] = {A[j], a1, b1, B[j], C[j], c1, d1, D[j]};
        const uint32_t ci = cube_index<ELEM>(cw);
        vsum += s_nv[ci];
        const int64_t k = nrec + __popcll(m & ((1ull << lane) - 1ull));
        if (k < room) {
          uint32_t* rp = P.rec + (rec_base + used + k) * kRecWords;
          rp[0] = ((uint32_t)(64 * j + lane) << 8) | ci;
#pragma unroll
          for (int q = 0; q < 8; ++q) rp[1 + q] = cw[q];
        }
      }
      nrec += __popcll(m);
    }
    const bool dense = nrec > room;
    vsum = wave_sum32(vsum);
    if (lane == 0) {
      P.tileT[t] = (int32_t)vsum;
      P.tile_nrec[t] = dense ? -1 : (int32_t)nrec;
      P.tile_rec[t] = rec_base + used;
      if (dense) atomicAdd(&P.hdr->dense_acc, 1u);
    }
    if (!dense) used += nrec;
    // append the column's non-empty tiles to the list, 64 planes at a time
    if (vsum > 0) nzmask |= 1ull << (z - zb);
    if (z - zb == 63 || z + 1 == z1) {
      if (nzmask) {
        const uint32_t n = (uint32_t)__popcll(nzmask);
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&P.hdr->nz_count, n);
        base = __shfl(base, 0);
        if ((nzmask >> lane) & 1ull) {
          const uint32_t i = (uint32_t)__popcll(nzmask & ((1ull << lane) - 1ull));
          P.nz_list[base + i] = (uint32_t)(V.tile_begin + (int64_t)(zb + lane) * V.R + r);
          if constexpr (BATCH) P.nz_vol[base + i] = (uint32_t)vi;
        }
      }
      nzmask = 0;
      zb = z + 1;
    }

#pragma unroll
    for (int j = 0; j <= kCPW; ++j) {
      A[j] = C[j];
      B[j] = D[j];
      C[j] = E[j];
      D[j] = F[j];
    }
  }
}

// ============================================================ K1, wide form
// SHORT2 volumes of any width.  128-bit loads: in chunk c (256 cubes) of a
// tile, lane L holds the voxels 4L..4L+3 of each of the four corner rows
// (rows y, y+1 of planes z, z+1); tiles start at multiples of 512 voxels,
// rows and planes of a width X % 4 != 0 at 4-byte alignment, which the
// vector memory path accepts.  The rows' tsdf signs become bytes (0xff:
// negative); OR / AND over the four rows and over positions k and k + 1 (the
// next lane's position 0 by one DPP lane shift) mark the cubes with both
// signs, four per lane, masked by the cubes that exist (a row may end inside
// a lane); only chunks with such cubes test the weights.  Active cube
// 4L + k of the chunk is in cube order, so a record's rank within the tile
// is a sum of masked popcounts.  Records go to the column's quota, or, when
// the tile does not fit there, to a shared overflow pool (one atomic per
// overflowing tile); only when that is exhausted does the tile take the
// staged path.  No active-cube masks are written (the backward runs on fp32
// volumes, which take the 32-bit kernel).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4a4 __attribute__((ext_vector_type(4), aligned(4)));  // 4-byte aligned

// Workgroup id remapped so that runs of kXcdGroup consecutive workgroups
// (consecutive columns) execute on one XCD (dispatch is round-robin over the
// 8 XCDs: ids b, b + 8, ... run on XCD b % 8): neighbouring columns share a
// row through that XCD's L2, and the runs rotate over the XCDs so that every
// XCD gets work from every slab.  Ids of an incomplete last round map to
// themselves.  Runs of 4 (16 columns) measured fastest at 512^3: 141 us for
// K1 against 142 (1, no remap), 141-142 (2), 144-146 (8) and 150 (16, 32).
#ifndef MCX_XCD_GROUP
#define MCX_XCD_GROUP 4
#endif
constexpr uint32_t kXcdGroup = MCX_XCD_GROUP;
__device__ __forceinline__ uint32_t xcd_block() {
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t x = b & 7u, i = b >> 3;
  const uint32_t g = i / kXcdGroup, j = i - g * kXcdGroup;
  if (g >= nb / (8u * kXcdGroup)) return b;
  return (g * 8u + x) * kXcdGroup + j;
}

// DPP control: wave_shl:1 (lane i reads lane i + 1; lane 63 keeps `old`)
constexpr int kDppWaveShl1 = 0x130;

struct Plane4 {
  u32x4 u[2], w[2];  // rows y (u) and y + 1 (w) of chunks 0 and 1
  uint32_t xu, xw;   // word 512 of the two rows (the same in all lanes; neighbour of chunk 1's lane 63)
  // per chunk, byte k = 0xff / 0x00: the tsdf at position k of row u / w
  // negative (set once per plane, used by two steps)
  uint32_t su[2], sw[2];
};

// v_perm_b32 selectors: byte 8 / 10 = the sign of bits 15 / 47 of {src0, src1}
// (src1 = the lower word) as 0xff or 0x00, byte 12 = 0x00.
constexpr uint32_t kSignSel01 = 0x0C0C0A08u;  // bytes 0, 1 <- signs of src1, src0
constexpr uint32_t kSignSel23 = 0x0A080C0Cu;  // bytes 2, 3 <- signs of src1, src0

// Bits of m from a, the others from b (v_bfi_b32).
__device__ __forceinline__ uint32_t bfi(uint32_t m, uint32_t a, uint32_t b) {
  return (m & a) | (~m & b);
}

// Byte k = 0xff when the tsdf of voxel position k is negative.
__device__ __forceinline__ uint32_t sign_bytes(const u32x4& v) {
  return __builtin_amdgcn_perm(v[1], v[0], kSignSel01) | __builtin_amdgcn_perm(v[3], v[2], kSignSel23);
}

__device__ __forceinline__ void plane_signs(Plane4& L) {
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    L.su[c] = sign_bytes(L.u[c]);
    L.sw[c] = sign_bytes(L.w[c]);
  }
}
typedef const __attribute__((address_space(4))) uint32_t const_u32;  // scalar loads

template <bool BATCH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
void count4_kernel(ExtractParams P) {
  __shared__ uint8_t s_nv[256];
  __shared__ __attribute__((aligned(16))) uint32_t s_rec[4][256 * kRecWords];  // per wave: one chunk's records
  __shared__ uint32_t s_wg_used;  // records taken from the workgroup's region
  s_nv[threadIdx.x] = dNumVerts[threadIdx.x];
  if (threadIdx.x == 0) s_wg_used = 0;
  clear_scan_descriptors(P);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  uint32_t* const my_rec = s_rec[wave];
  const int64_t col = (int64_t)xcd_block() * 4 + wave;
  if (col >= P.ncols) return;  // whole waves exit; no barrier follows

  int vi, r, z0, z1;
  if constexpr (BATCH) {
    const Column c = P.cols[col];
    vi = c.vol;
    r = c.r;
    z0 = c.z0;
    z1 = c.z1;
  } else {
    vi = 0;
    r = (int)(col % P.vol0.R);
    const int s = (int)(col / P.vol0.R);
    z0 = s * P.slab;
    z1 = min(z0 + P.slab, P.vol0.Z - 1);
  }
  const VolDesc& V = vol_at<BATCH>(P, vi);
  const int X = V.X;
  const int64_t XY = (int64_t)X * V.Y;
  const int64_t p0 = (int64_t)r * kTile;
  // The quotas of the workgroup's (up to) four consecutive columns form one
  // region shared by its waves: a tile takes its records with an LDS atomic
  // (no wait on the vector memory counter, so the prefetched planes stay in
  // flight), and only when the region is full from the shared overflow pool.
  const int64_t col0 = col - wave;
  const int64_t wg_base = col0 * P.quota;
  const uint32_t wg_size = (uint32_t)(min((int64_t)4, P.ncols - col0) * P.quota);

  // byte k of cube_mask[c] = 0xff when cube k of the lane exists: in-plane
  // index p + k inside the plane, not in the last row and not the last of
  // its row (its x + 1 neighbour, position k + 1, is then in the same row;
  // a row may end inside a lane when X % 4 != 0)
  uint32_t cube_mask[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int64_t p = p0 + 256 * c + 4 * lane;
    uint32_t mk = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (p + k < XY) {
        const uint32_t y = V.divX.div((uint32_t)(p + k));
        const uint32_t x = (uint32_t)(p + k) - y * (uint32_t)X;
        if (y + 1 < (uint32_t)V.Y && x + 1 < (uint32_t)X) mk |= 0xffu << (8 * k);
      }
    }
    cube_mask[c] = mk;
  }

  // One buffer resource over the planes the column reads (z0 .. z1); the
  // plane offset is part of the VGPR offset, which the hardware range-checks
  // (the SGPR offset is not checked): planes past z1 read as zero.  Voxels of
  // a tile past the end of its plane read the next plane's words, which only
  // lanes without cubes see (they are masked).  The host keeps
  // (z1 - z0 + 4) planes under 2^31 bytes.
  const uint32_t plane_bytes = (uint32_t)(XY * 4);
  const uint32_t voff = (uint32_t)(p0 + 4 * lane) * 4u;
  const uint32_t voffW = voff + 4u * (uint32_t)X;
  // The plane loads are issued as inline assembly so that the compiler's
  // wait insertion does not see them: after a step's variable number of
  // record stores it could only wait for every outstanding load (including
  // the prefetched plane).  Each step instead waits for all but the newest
  // six vector memory operations (vector memory operations of a wave retire
  // in issue order), tied to the registers it is about to read.
  u32x4 rsv;  // buffer resource: base, stride 0, size, format word
  {
    const uint64_t base = reinterpret_cast<uint64_t>(V.vol + (int64_t)z0 * XY);
    rsv[0] = (uint32_t)base;
    rsv[1] = (uint32_t)(base >> 32) & 0xffffu;
    rsv[2] = (uint32_t)(z1 - z0 + 1) * plane_bytes;
    rsv[3] = (uint32_t)kBufferWord3;
    rsv[0] = __builtin_amdgcn_readfirstlane(rsv[0]);
    rsv[1] = __builtin_amdgcn_readfirstlane(rsv[1]);
    rsv[2] = __builtin_amdgcn_readfirstlane(rsv[2]);
    rsv[3] = __builtin_amdgcn_readfirstlane(rsv[3]);
  }
  // the word after the tile in each row (the same for all lanes; an offset
  // past the resource, hence 0, when it lies outside the plane)
  const uint32_t xoffU = p0 + kTile < XY ? (uint32_t)(p0 + kTile) * 4u : 0x80000000u;
  const uint32_t xoffW = p0 + kTile + X < XY ? (uint32_t)(p0 + kTile + X) * 4u : 0x80000000u;
  auto load_plane = [&](int q, Plane4& L) {
    const uint32_t po = (uint32_t)(q - z0) * plane_bytes;
    const uint32_t a = voff + po, b = voffW + po;
    asm volatile("buffer_load_dwordx4 %0, %1, %2, 0 offen" : "=v"(L.u[0]) : "v"(a), "s"(rsv) : "memory");
    asm volatile("buffer_load_dwordx4 %0, %1, %2, 0 offen offset:1024" : "=v"(L.u[1]) : "v"(a), "s"(rsv) : "memory");
    asm volatile("buffer_load_dwordx4 %0, %1, %2, 0 offen" : "=v"(L.w[0]) : "v"(b), "s"(rsv) : "memory");
    asm volatile("buffer_load_dwordx4 %0, %1, %2, 0 offen offset:1024" : "=v"(L.w[1]) : "v"(b), "s"(rsv) : "memory");
    const uint32_t xa = xoffU + po, xb = xoffW + po;
    asm volatile("buffer_load_dword %0, %1, %2, 0 offen" : "=v"(L.xu) : "v"(xa), "s"(rsv) : "memory");
    asm volatile("buffer_load_dword %0, %1, %2, 0 offen" : "=v"(L.xw) : "v"(xb), "s"(rsv) : "memory");
  };
  // all but the newest six vector memory operations done (they retire in
  // issue order); ties L's registers
  auto wait_plane6 = [&](Plane4& L) {
    asm volatile("s_waitcnt vmcnt(6)" : "+v"(L.u[0]), "+v"(L.u[1]), "+v"(L.w[0]), "+v"(L.w[1]), "+v"(L.xu), "+v"(L.xw) :: "memory");
  };
  // Four plane buffers whose roles rotate over an unrolled loop of four steps
  // (no register copies, which would wait for in-flight loads): the step for
  // cubes between planes z and z + 1 loads plane z + 3.
  Plane4 A, B, C, D;
  load_plane(z0, A);
  load_plane(z0 + 1, B);
  load_plane(z0 + 2, C);
  asm volatile("s_waitcnt vmcnt(12)" : "+v"(A.u[0]), "+v"(A.u[1]), "+v"(A.w[0]), "+v"(A.w[1]), "+v"(A.xu), "+v"(A.xw) :: "memory");
  plane_signs(A);

  uint64_t nzmask = 0;  // planes zb + i of this column whose tile has vertices
  int zb = z0;
  uint32_t t = (uint32_t)(V.tile_begin + (int64_t)z0 * V.R + r);  // this step's tile
  const uint32_t tR = (uint32_t)V.R;
  // The plane two steps ahead is loaded at the end of a step, after the
  // step's record stores: the data a step waits for is then always older
  // than the newest loads (a fixed count), however many stores were issued.
  auto step = [&](Plane4& A, Plane4& B, Plane4& L, int z) __attribute__((always_inline)) {

    // ---- pass 1: active masks m[c][k] (bit L: cube 256c + 4L + k).  Signs
    // first, four cubes per lane at once: byte k of O / N = some / every tsdf
    // of the four corner rows at position k negative (the planes' row pairs
    // combined); the bytes of positions k + 1 come from shifting O / N down
    // one byte and taking the next lane's byte 0 (a DPP lane shift; lane 63
    // takes the next chunk's lane 0, or the words after the tile).  A cube
    // has both signs where OR & ~AND; lanes without cubes and position 4 at a
    // row end are masked.  Only chunks with such cubes test the weights:
    // the minimum word over the eight corners below 0x10000 means some
    // weight is zero.
    wait_plane6(B);
    plane_signs(B);
    uint64_t m[2][4];
    uint32_t nxo, nxn, nxm;  // position 0 of the next chunk's lane 0 (wave-uniform)
    {
      const uint32_t xo = A.xu | A.xw | B.xu | B.xw, xa = A.xu & A.xw & B.xu & B.xw;
      nxo = (xo & 0x8000u) ? 0xffu : 0u;
      nxn = (xa & 0x8000u) ? 0xffu : 0u;
      nxm = min(min(A.xu, A.xw), min(B.xu, B.xw));
    }
#pragma unroll
    for (int c = 1; c >= 0; --c) {
      const uint32_t O = A.su[c] | A.sw[c] | B.su[c] | B.sw[c];
      const uint32_t N = A.su[c] & A.sw[c] & B.su[c] & B.sw[c];
      const uint32_t O2 = __builtin_amdgcn_alignbyte(
          __builtin_amdgcn_update_dpp(nxo, O, kDppWaveShl1, 0xf, 0xf, false), O, 1);
      const uint32_t N2 = __builtin_amdgcn_alignbyte(
          __builtin_amdgcn_update_dpp(nxn, N, kDppWaveShl1, 0xf, 0xf, false), N, 1);
      const uint32_t M = (O | O2) & ~(N & N2) & cube_mask[c];
      uint64_t sg[4];
      sg[0] = __ballot((int8_t)M < 0);
      sg[1] = __ballot((int16_t)M < 0);
      sg[2] = __ballot((int8_t)(M >> 16) < 0);
      sg[3] = __ballot((int32_t)M < 0);
      uint32_t mn0 = 0;
      if (sg[0] | sg[1] | sg[2] | sg[3]) {
        uint32_t mn[5];
#pragma unroll
        for (int pos = 0; pos < 4; ++pos)
          mn[pos] = min(min(min(A.u[c][pos], A.w[c][pos]), B.u[c][pos]), B.w[c][pos]);
        mn[4] = __builtin_amdgcn_update_dpp(nxm, mn[0], kDppWaveShl1, 0xf, 0xf, false);
        mn0 = mn[0];
#pragma unroll
        for (int k = 0; k < 4; ++k) m[c][k] = sg[k] & __ballot(min(mn[k], mn[k + 1]) >= 0x10000u);
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) m[c][k] = 0;
        if (c == 1)
          mn0 = min(min(min(A.u[1][0], A.w[1][0]), B.u[1][0]), B.w[1][0]);
      }
      if (c == 1) {
        nxo = __builtin_amdgcn_readlane(O, 0) & 0xffu;
        nxn = __builtin_amdgcn_readlane(N, 0) & 0xffu;
        nxm = __builtin_amdgcn_readlane(mn0, 0);
      }
    }
    const uint64_t any_c0 = m[0][0] | m[0][1] | m[0][2] | m[0][3];
    const uint64_t any_c1 = m[1][0] | m[1][1] | m[1][2] | m[1][3];
    if ((any_c0 | any_c1) == 0) {
      // no active cube: only the tile's vertex count (0) is read later
      if (lane == 0) P.tileT[t] = 0;
      load_plane(z + 3, L);
      t += tR;
      if (z - zb == 63 || z + 1 == z1) {
        if (nzmask) {
          const uint32_t n = (uint32_t)__popcll(nzmask);
          uint32_t base = 0;
          if (lane == 0) base = atomicAdd(&P.hdr->nz_count, n);
          base = __shfl(base, 0);
          if ((nzmask >> lane) & 1ull) {
            const uint32_t i = (uint32_t)__popcll(nzmask & ((1ull << lane) - 1ull));
            P.nz_list[base + i] = (uint32_t)(V.tile_begin + (int64_t)(zb + lane) * V.R + r);
            if constexpr (BATCH) P.nz_vol[base + i] = (uint32_t)vi;
          }
        }
        nzmask = 0;
        zb = z + 1;
      }
      return;
    }
    int nrec32 = 0;  // at most 512
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int k = 0; k < 4; ++k) nrec32 += __popcll(m[c][k]);
    const int64_t nrec = nrec32;

    // ---- destination of the tile's records (a failed take leaves the
    // counter past the region: later tiles of the workgroup use the pool)
    uint32_t got = 0;
    if (lane == 0) got = atomicAdd(&s_wg_used, (uint32_t)nrec32);
    got = __builtin_amdgcn_readfirstlane(got);
    int64_t rbase = wg_base + got;
    bool dense = false;
    if (got + (uint32_t)nrec32 > wg_size) {
      unsigned long long ob = 0;
      if (lane == 0) ob = atomicAdd(reinterpret_cast<unsigned long long*>(&P.hdr->ovf_next),
                                    (unsigned long long)nrec);
      ob = __shfl(ob, 0);
      if ((int64_t)ob + nrec <= P.ovf_cap) rbase = P.ovf_begin + (int64_t)ob;
      else dense = true;
    }

    // ---- pass 2: the active cubes of a chunk are staged in the wave's LDS
    // area in cube order (offset | case, 8 corner words) and written to the
    // scratch with consecutive lanes on consecutive words.  The cases of a
    // lane's four cubes come from the rows' sign bytes at positions k and
    // k + 1 (active cubes have no zero weight, so the case is the corner
    // signs), one bit per corner in corner order.
    uint32_t vsum = 0;
    int64_t cbase = 0;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const uint64_t any = m[c][0] | m[c][1] | m[c][2] | m[c][3];
      if (any == 0) continue;
      uint32_t pre = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        pre += __builtin_amdgcn_mbcnt_hi((uint32_t)(m[c][k] >> 32),
                                         __builtin_amdgcn_mbcnt_lo((uint32_t)m[c][k], 0u));
      // position 4 of the lane: the next lane's position 0; lane 63 takes
      // lane 0 of chunk 1 (c = 0) or the words after the tile (c = 1)
      uint32_t nau = 0, naw = 0, nbu = 0, nbw = 0;
      if (m[c][3]) {
        const uint32_t lau = c == 0 ? __builtin_amdgcn_readlane(A.u[1][0], 0) : A.xu;
        const uint32_t law = c == 0 ? __builtin_amdgcn_readlane(A.w[1][0], 0) : A.xw;
        const uint32_t lbu = c == 0 ? __builtin_amdgcn_readlane(B.u[1][0], 0) : B.xu;
        const uint32_t lbw = c == 0 ? __builtin_amdgcn_readlane(B.w[1][0], 0) : B.xw;
        nau = __builtin_amdgcn_update_dpp(lau, A.u[c][0], kDppWaveShl1, 0xf, 0xf, false);
        naw = __builtin_amdgcn_update_dpp(law, A.w[c][0], kDppWaveShl1, 0xf, 0xf, false);
        nbu = __builtin_amdgcn_update_dpp(lbu, B.u[c][0], kDppWaveShl1, 0xf, 0xf, false);
        nbw = __builtin_amdgcn_update_dpp(lbw, B.w[c][0], kDppWaveShl1, 0xf, 0xf, false);
      }
      uint32_t cases;
      {
        // the rows' sign bytes at positions k + 1: down one byte, the next
        // lane's byte 0 on top (lane 63: the next chunk's lane 0, or the
        // words after the tile)
        uint32_t x0, x1, x2, x3;
        if (c == 0) {
          x0 = __builtin_amdgcn_readlane(A.su[1], 0);
          x1 = __builtin_amdgcn_readlane(A.sw[1], 0);
          x2 = __builtin_amdgcn_readlane(B.su[1], 0);
          x3 = __builtin_amdgcn_readlane(B.sw[1], 0);
        } else {
          x0 = (A.xu & 0x8000u) ? 0xffu : 0u;
          x1 = (A.xw & 0x8000u) ? 0xffu : 0u;
          x2 = (B.xu & 0x8000u) ? 0xffu : 0u;
          x3 = (B.xw & 0x8000u) ? 0xffu : 0u;
        }
        auto next = [&](uint32_t v, uint32_t x) {
          return __builtin_amdgcn_alignbyte(
              __builtin_amdgcn_update_dpp(x, v, kDppWaveShl1, 0xf, 0xf, false), v, 1);
        };
        const uint32_t au1 = next(A.su[c], x0), aw1 = next(A.sw[c], x1);
        const uint32_t bu1 = next(B.su[c], x2), bw1 = next(B.sw[c], x3);
        cases = bfi(0x01010101u, A.su[c],
                bfi(0x02020202u, au1,
                bfi(0x04040404u, aw1,
                bfi(0x08080808u, A.sw[c],
                bfi(0x10101010u, B.su[c],
                bfi(0x20202020u, bu1, bfi(0x40404040u, bw1, B.sw[c])))))));
      }
      uint32_t own = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if ((m[c][k] >> lane) & 1ull) {
          uint32_t* rp = my_rec + (pre + own) * kRecWords;
          const uint32_t ci = (cases >> (8 * k)) & 0xffu;
          rp[0] = ((uint32_t)(256 * c + 4 * lane + k) << 8) | ci;
          vsum += s_nv[ci];
          rp[1] = A.u[c][k];
          rp[2] = k < 3 ? A.u[c][(k + 1) & 3] : nau;
          rp[3] = k < 3 ? A.w[c][(k + 1) & 3] : naw;
          rp[4] = A.w[c][k];
          rp[5] = B.u[c][k];
          rp[6] = k < 3 ? B.u[c][(k + 1) & 3] : nbu;
          rp[7] = k < 3 ? B.w[c][(k + 1) & 3] : nbw;
          rp[8] = B.w[c][k];
          ++own;
        }
      }
      const int nc = __popcll(m[c][0]) + __popcll(m[c][1]) + __popcll(m[c][2]) + __popcll(m[c][3]);
      wave_sync();
      if (!dense) {
        // the chunk's records are consecutive in scratch: 16 bytes a lane
        // (the staging area is 16-byte aligned, the scratch 4-byte aligned),
        // then the last words
        uint32_t* dst = P.rec + (rbase + cbase) * kRecWords;
        const int nw = nc * kRecWords, n4 = nw >> 2;
        u32x4a4* dst4 = reinterpret_cast<u32x4a4*>(dst);
        const u32x4* src4 = reinterpret_cast<const u32x4*>(my_rec);
        for (int w = lane; w < n4; w += 64) dst4[w] = src4[w];
        if (lane < nw - 4 * n4) dst[4 * n4 + lane] = my_rec[4 * n4 + lane];
        wave_sync();
      }
      cbase += nc;
    }
    vsum = wave_sum32(vsum);
    if (lane == 0) {
      P.tileT[t] = (int32_t)vsum;
      P.tile_nrec[t] = dense ? -1 : (int32_t)nrec;
      P.tile_rec[t] = rbase;
      if (dense) atomicAdd(&P.hdr->dense_acc, 1u);
    }
    load_plane(z + 3, L);
    t += tR;
    // append the column's non-empty tiles to the list, 64 planes at a time
    if (vsum > 0) nzmask |= 1ull << (z - zb);
    if (z - zb == 63 || z + 1 == z1) {
      if (nzmask) {
        const uint32_t n = (uint32_t)__popcll(nzmask);
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&P.hdr->nz_count, n);
        base = __shfl(base, 0);
        if ((nzmask >> lane) & 1ull) {
          const uint32_t i = (uint32_t)__popcll(nzmask & ((1ull << lane) - 1ull));
          P.nz_list[base + i] = (uint32_t)(V.tile_begin + (int64_t)(zb + lane) * V.R + r);
          if constexpr (BATCH) P.nz_vol[base + i] = (uint32_t)vi;
        }
      }
      nzmask = 0;
      zb = z + 1;
    }
  };
  for (int z = z0; z < z1; z += 4) {
    step(A, B, D, z);
    if (z + 1 < z1) step(B, C, A, z + 1);
    if (z + 2 < z1) step(C, D, B, z + 2);
    if (z + 3 < z1) step(D, A, C, z + 3);
  }
}

// ============================================================ K2: scan
// Exclusive scan of the tile vertex counts (tile order) into voff, single
// pass with decoupled look-back over blocks of 1024 tiles; a block publishes
// its tile aggregate as soon as its counts are in.  With list_scan, the
// blocks then split the non-empty-tile list evenly (contiguous ranges, list
// order): tiles that took the staged path move to dense_list, the others
// are compacted into clist with their first record (lrs) and the exclusive
// scan of their record counts (lpre), and units[k] is set to the clist entry
// holding record 64 k.  The three scans have descriptor sets of their own,
// looked back by waves 0, 1 and 2.  A workgroup's block of tiles is the
// ticket it takes when it starts (not its blockIdx), so the blocks it waits
// for belong to workgroups that are already running.  The workgroup with the
// last ticket publishes the totals (vertex count, list entries, list
// records) from its inclusive prefixes, the only writer of each.
__global__ __launch_bounds__(256) void scan_kernel(ExtractParams P) {
  __shared__ int64_t s_wave[3][4];
  __shared__ int64_t s_excl[3];
  __shared__ uint32_t s_ticket, s_nz;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) s_ticket = atomicAdd(&P.hdr->scan_ticket, 1u);
  __syncthreads();
  const int64_t blk = s_ticket;
  // The list length.  The last ticket resets this counter for the next call
  // once every other block has published, so the read must be complete
  // before the block's first publication: one thread reads it beside its
  // count loads and writes it to LDS before the barrier that precedes the
  // publication (the write waits for the read; that barrier waits for the
  // count loads anyway), and the block uses the LDS copy.
  uint32_t nz_read = 0;
  if (tid == 0) nz_read = P.hdr->nz_count;
  const bool last_blk = blk == (int64_t)gridDim.x - 1;
  const int64_t i0 = blk * kScanItems + (int64_t)tid * kScanPer;

  // ---- tiles: one 128-bit load of the thread's 4 counts (i0 is a multiple
  // of 4, the array is 256-byte aligned), element loads at the end
  static_assert(kScanPer == 4, "one int4 per thread");
  int32_t v[kScanPer];
  if (i0 + kScanPer <= P.n_tiles) {
    const int4 w = *reinterpret_cast<const int4*>(P.tileT + i0);
    v[0] = w.x;
    v[1] = w.y;
    v[2] = w.z;
    v[3] = w.w;
  } else {
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) v[k] = i0 + k < P.n_tiles ? P.tileT[i0 + k] : 0;
  }
  int64_t st = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) st += v[k];
  const uint64_t incl0 = wave_incl_scan((uint64_t)st);
  if (lane == 63) s_wave[0][wave] = (int64_t)incl0;
  if (tid == 0) s_nz = nz_read;
  __syncthreads();
  const uint32_t nz = s_nz;
  int64_t texcl0 = (int64_t)incl0 - st, total0 = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    if (w < wave) texcl0 += s_wave[0][w];
    total0 += s_wave[0][w];
  }
  // the descriptors were zeroed by the count kernel of this call
  if (wave == 0 && lane == 0)
    st_agent(P.desc + blk, (blk == 0 ? kStatInc : kStatAgg) | (uint64_t)total0);

  // ---- list: this block's contiguous range, c entries per thread
  const bool lst = P.list_scan && nz > 0;
  const ListPart part = list_part(lst ? (int64_t)nz : 0, gridDim.x, blk);  // c <= kScanPer
  const int64_t j0 = part.lb + (int64_t)tid * part.c;
  uint32_t tl[kScanPer], vl[kScanPer];
  int32_t n[kScanPer];
  int64_t rsl[kScanPer];
  const bool batch = P.vols != nullptr;  // entries carry their volume
  uint32_t keep = 0;  // bit k: entry j0 + k has records
  int64_t sn = 0, se = 0;
  if (lst) {
    // every load issued before any is waited for: clamped indices, the
    // entries outside the range read a valid one and are dropped
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) tl[k] = P.nz_list[min(j0 + k, (int64_t)nz - 1)];
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) vl[k] = batch ? P.nz_vol[min(j0 + k, (int64_t)nz - 1)] : 0u;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
      n[k] = P.tile_nrec[tl[k]];
      rsl[k] = P.tile_rec[tl[k]];
    }
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
      if (list_slot(part, tid, k) >= 0) {
        if (n[k] < 0) {  // the staged path (rare): off the list
          P.dense_list[atomicAdd(&P.hdr->dense_n, 1u)] = tl[k];
          n[k] = 0;
        } else {
          keep |= 1u << k;
        }
      } else {
        n[k] = 0;
      }
      sn += n[k];
    }
    se = __popc(keep);
  }
  const uint64_t incl1 = wave_incl_scan((uint64_t)sn);
  const uint64_t incl2 = wave_incl_scan((uint64_t)se);
  if (lane == 63) {
    s_wave[1][wave] = (int64_t)incl1;
    s_wave[2][wave] = (int64_t)incl2;
  }
  __syncthreads();
  int64_t texcl1 = (int64_t)incl1 - sn, texcl2 = (int64_t)incl2 - se, total1 = 0, total2 = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    if (w < wave) {
      texcl1 += s_wave[1][w];
      texcl2 += s_wave[2][w];
    }
    total1 += s_wave[1][w];
    total2 += s_wave[2][w];
  }
  if (wave == 0) {
    const int64_t excl = blk == 0 ? 0 : lookback<1>(P.desc, blk, &P.hdr->error);
    if (lane == 0) {
      if (blk > 0) st_agent(P.desc + blk, kStatInc | (uint64_t)(excl + total0));
      s_excl[0] = excl;
    }
  } else if (lst && wave < 3) {
    uint64_t* desc = P.desc + wave * P.cap_desc;
    const int64_t tot = wave == 1 ? total1 : total2;
    int64_t excl = 0;
    if (blk == 0) {
      if (lane == 0) st_agent(desc, kStatInc | (uint64_t)tot);
    } else {
      if (lane == 0) st_agent(desc + blk, kStatAgg | (uint64_t)tot);
      excl = lookback<1>(desc, blk, &P.hdr->error);
      if (lane == 0) st_agent(desc + blk, kStatInc | (uint64_t)(excl + tot));
    }
    if (lane == 0) s_excl[wave] = excl;
  }
  __syncthreads();
  if (last_blk && tid == 0) {
    // every other workgroup has published (the look-backs ended), so every
    // one has read nz_count: the counters of the count kernel are handed over
    // and reset for the next call
    const bool bad = __hip_atomic_load(&P.hdr->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    const int64_t cnt = bad ? -1 : s_excl[0] + total0;
    P.hdr->count = cnt;
    if (P.d_count) *P.d_count = cnt;
    P.hdr->dense_last = P.hdr->dense_acc;  // K1 has finished: hand the count over
    P.hdr->dense_acc = 0;
    P.hdr->nz_total = nz;                  // the emit kernel's work list
    P.hdr->nz_count = 0;
    P.hdr->ovf_next = 0;
    if (lst) {
      const int64_t nl = s_excl[2] + total2;  // list entries with records
      P.hdr->list_n = (uint32_t)nl;
      P.lpre[nl] = s_excl[1] + total1;        // records of the whole list
    }
  }
  int64_t run = s_excl[0] + texcl0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    if (i0 + k < P.n_tiles) P.voff[i0 + k] = run;
    run += v[k];
  }
  if (lst) {
    int64_t r = s_excl[1] + texcl1;
    int64_t e = s_excl[2] + texcl2;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
      if ((keep >> k) & 1u) {
        P.clist[e] = tl[k];
        P.lpre[e] = r;
        P.lrs[e] = rsl[k];
        if (batch) P.clist_vol[e] = vl[k];
        for (int64_t u = (r + 63) >> 6; (u << 6) < r + n[k]; ++u) P.units[u] = (uint32_t)e;
        r += n[k];
        ++e;
      }
    }
  }
}

// per-volume first vertex (batched launches), vbase[nvol] = total
__global__ __launch_bounds__(256) void vbase_kernel(ExtractParams P) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > P.nvol) return;
  const int64_t total = P.hdr->count;
  if (total < 0 || P.hdr->error) {  // the scan reported an error: every base is -1
    P.vbase[b] = -1;
    return;
  }
  if (b == P.nvol) {
    P.vbase[b] = total;
    return;
  }
  const int64_t t = P.vols[b].tile_begin;
  P.vbase[b] = t < P.n_tiles ? P.voff[t] : total;
}

// ============================================================ K3: emit
// Exclusive scan of one int per thread over a 256-thread workgroup; returns
// the thread's exclusive prefix and the total through *tot (s_tmp: 4 ints).
__device__ __forceinline__ int block_excl_scan(int v, int* s_tmp, int* tot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t incl = wave_incl_scan32((uint32_t)v);
  if (lane == 63) s_tmp[wave] = (int)incl;
  __syncthreads();
  int pre = 0, all = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    if (w < wave) pre += s_tmp[w];
    all += s_tmp[w];
  }
  __syncthreads();
  *tot = all;
  return (int)incl - v + pre;
}

struct TileInfo {      // backward BW1 per-tile info in LDS (10 words)
  int64_t delta;       // output index of the tile's first record vertex minus
                       // the item-list vertex prefix at that record
  int64_t rstart;      // first record in scratch
  int64_t t;           // tile id
  int32_t nrec;        // records (-1: staged path)
  int32_t T;           // vertices
  int32_t lbeg;        // first position in the item's record list
  int32_t vi;          // volume
};
static_assert(sizeof(TileInfo) == 40, "tile info is 10 words");

// Per-wave LDS area of the emit kernel (32-bit words): 64 records (cube
// offset << 8 | case, 8 corners), the vertex -> record map (u8, at most 15
// vertices a record), the first vertex of each record, and per record its
// output index and cube (x, y, z) in 4 words (meta_hi).
constexpr int kWaveRec = 64 * kRecWords;
constexpr int kWaveMap = 64 * 15 / 4;
constexpr int kWaveEx = 64;
constexpr int kWaveMeta = 64 * 4;
constexpr int kPackedMap = kWaveRec, kPackedEx = kPackedMap + kWaveMap;
constexpr int kPackedMeta = kPackedEx + kWaveEx;
constexpr int kPackedWords = kPackedMeta + kWaveMeta;

// Word 1 of a record's metadata in the emit kernels: bits 32..39 of its
// output index (saturated: an index of 2^40 or more lies past any capacity)
// and the cube's z (< 2^24).  Words 0, 2, 3: the index's low bits, x, y.
__device__ __forceinline__ uint32_t meta_hi(uint64_t out, uint32_t z) {
  return (uint32_t)min(out >> 32, (uint64_t)0xff) | (z << 8);
}

// One output vertex, slot q of the record at rc (LDS: word 0 = cube offset
// << 8 | case, words 1..8 = the corner words in corner order) of cube
// (x, y, z), in the fp32 operations of cu_marching_cubes_v1.cu:243-264, :336-367 (
// node coordinate (i + 0.5) h, t = (0 - fa) / (fb - fa + 1e-15f), fma
// interpolation, voxel floor(p / h - 0.5)); returns the grid
// edge key (lower corner voxel * 3 + axis) when want_key: the edge's corners
// come from the slot table and their words are read directly; along the two
// axes the edge does not run on, the interpolated coordinate is exactly the
// corner's node coordinate (fma(t, +0, p) = p), so only the edge axis is
// interpolated.
// Grid of a record's volume: dimensions, spacing and (batched launches) its
// index in the volume table; in registers (batched launches fetch them from
// the record's lane).
struct GridDims {
  int X, Y, Z;
  float h;
  uint32_t vi;
};

// vol_of() gives the volume's voxels (only called where the record's voxel
// is not one of its cube's corners).
template <int ELEM, typename VolOf>
__device__ __forceinline__ int64_t record_vertex(const GridDims& V, VolOf vol_of,
                                                 int16_t* ver_voxel, const uint8_t* s_slot,
                                                 const uint32_t* rc, int x, int y, int z, int q,
                                                 uint32_t* o, bool want_key) {
  const uint32_t sl = s_slot[(rc[0] & 0xffu) * 16 + q];
  const int ca = (int)(sl & 7u), cb = (int)((sl >> 3) & 7u), ax = (int)(sl >> 6);
  const uint32_t wa = rc[1 + ca], wb = rc[1 + cb];
  const int ix = x + corner_dx(ca), iy = y + corner_dy(ca), iz = z + corner_dz(ca);
  const int ia = ax == 0 ? ix : (ax == 1 ? iy : iz);          // ca along the edge
  const int ib = ia + (ax == 0 ? corner_dx(cb) - corner_dx(ca)
                               : (ax == 1 ? corner_dy(cb) - corner_dy(ca)
                                          : corner_dz(cb) - corner_dz(ca)));
  int64_t key = 0;
  if (want_key) {
    // the grid edge: its lower corner and its axis
    const int lo = min(ia, ib);
    const int mx = ax == 0 ? lo : ix, my = ax == 1 ? lo : iy, mz = ax == 2 ? lo : iz;
    key = (((int64_t)mz * V.Y + my) * V.X + mx) * 3 + ax;
  }
  const float h = V.h;
  const float fa = unpack_f<ELEM>(wa);
  const float fb = unpack_f<ELEM>(wb);
  const float num = 0.0f - fa, den = fb - fa + 1e-15f;
  float tt;
  if constexpr (ELEM == MCX_ELEM_SHORT2) tt = div_weight_short2(num, den);
  else tt = num / den;
  const float pax = ((float)ix + 0.5f) * h;
  const float pay = ((float)iy + 0.5f) * h;
  const float paz = ((float)iz + 0.5f) * h;
  const float pae = ax == 0 ? pax : (ax == 1 ? pay : paz);
  const float pbe = ((float)ib + 0.5f) * h;
  const float pe = __fmaf_rn(tt, pbe - pae, pae);
  const float px = ax == 0 ? pe : pax;
  const float py = ax == 1 ? pe : pay;
  const float pz = ax == 2 ? pe : paz;
  // record fields (store_point :336-367)
  int vx, vy, vz;
  {
    float cx = px / h, cy = py / h, cz = pz / h;
    cx = cx - 0.5f;
    cy = cy - 0.5f;
    cz = cz - 0.5f;
    vx = (int)floorf(cx);
    vy = (int)floorf(cy);
    vz = (int)floorf(cz);
  }
  const bool inb = vx >= 0 && vy >= 0 && vz >= 0 && vx < V.X && vy < V.Y && vz < V.Z;
  uint32_t tw = 0;
  auto lin = [&]() { return ((int64_t)vz * V.Y + vy) * V.X + vx; };
  if (inb) {
    if constexpr (ELEM == MCX_ELEM_SHORT2) {
      // the voxel is almost always one of the cube's corners; otherwise it
      // is loaded and consumed in the branch (a VALU copy), so that no load
      // is outstanding where the paths join: a wait there would also wait
      // for every store issued before it (one counter for both)
      const unsigned dx = (unsigned)(vx - x), dy = (unsigned)(vy - y), dz = (unsigned)(vz - z);
      if ((dx | dy | dz) <= 1u) {
        tw = rc[1 + corner_of((int)dx, (int)dy, (int)dz)];
      } else {
        const uint32_t t = vol_of()[lin()];
        asm volatile("v_mov_b32 %0, %1" : "=v"(tw) : "v"(t));
      }
    }
  }
  o[0] = __float_as_uint(px);
  o[1] = __float_as_uint(py);
  o[2] = __float_as_uint(pz);
  o[3] = (uint32_t)vx;
  o[4] = (uint32_t)vy;
  o[5] = (uint32_t)vz;
  o[6] = tw;
  if (ver_voxel && inb) ver_voxel[lin()] = (int16_t)(tw & 0xffffu);
  return key;
}

// Emit the vertices of a packed block: nb records (up to 64, possibly of
// several tiles) in the wave's record area, record r with its output index
// (of its first vertex) and cube (x, y, z) in s_meta[4r..4r+3] (words: index
// bits 0..31, meta_hi, x, y).  The records' vertex counts are scanned across the wave,
// every record writes its index into the vertex -> record map, and each lane
// produces one vertex of a round of 64 and stores its record fields at its
// output index (the vertex's own slot; consecutive lanes mostly write
// consecutive records).
// Batched launches: lane r < nb holds the grid and index of record r's
// volume in `dl` (each vertex lane fetches its record's with lane shuffles,
// not from the volume table in memory).
template <int ELEM, bool BATCH>
__device__ __forceinline__ void emit_block(const ExtractParams& P, const VolDesc& V,
                                           const uint8_t* s_slot, const uint8_t* s_nv,
                                           uint32_t* s_w, int nb, const GridDims& dl) {
  const uint32_t* s_rec = s_w;
  uint8_t* s_map = reinterpret_cast<uint8_t*>(s_w + kPackedMap);
  uint32_t* s_ex = s_w + kPackedEx;
  const uint32_t* s_meta = s_w + kPackedMeta;
  const int lane = threadIdx.x & 63;
  wave_sync();
  const int nv = lane < nb ? s_nv[s_rec[lane * kRecWords] & 0xffu] : 0;
  const uint32_t incl = wave_incl_scan32((uint32_t)nv);
  const int ex = (int)incl - nv;
  const int tot = (int)__builtin_amdgcn_readlane(incl, 63);
  s_ex[lane] = (uint32_t)ex;
  for (int q = 0; q < nv; ++q) s_map[ex + q] = (uint8_t)lane;
  wave_sync();
  uint32_t* vv = reinterpret_cast<uint32_t*>(P.vv);
  for (int vb = 0; vb < tot; vb += 64) {
    const int j = vb + lane;
    const int r = j < tot ? (int)s_map[j] : 0;
    GridDims g;
    if constexpr (BATCH) {
      // every lane takes part in the shuffles (lanes past the round read
      // record 0's grid and do not use it)
      g.X = __shfl(dl.X, r);
      g.Y = __shfl(dl.Y, r);
      g.Z = __shfl(dl.Z, r);
      g.h = __shfl(dl.h, r);
      g.vi = (uint32_t)__shfl((int)dl.vi, r);
    } else {
      g = GridDims{V.X, V.Y, V.Z, V.h, 0u};
    }
    if (j < tot) {
      const uint32_t* rc = s_rec + r * kRecWords;
      const uint32_t* mt = s_meta + 4 * r;
      const int slot = j - (int)s_ex[r];
      const int64_t out = (int64_t)(((uint64_t)(mt[1] & 0xffu) << 32) | mt[0]) + slot;
      if (out < P.cap) {
        uint32_t o[7];
        auto vol_of = [&]() { return BATCH ? P.vols[g.vi].vol : V.vol; };
        const int64_t key = record_vertex<ELEM>(g, vol_of, BATCH ? nullptr : V.ver_voxel, s_slot,
                                                rc, (int)mt[2], (int)mt[3],
                                                (int)(mt[1] >> 8), slot, o, P.ekey != nullptr);
        if (vv) {
#pragma unroll
          for (int f = 0; f < 7; ++f) vv[7 * out + f] = o[f];
        }
        if (P.pos) {
#pragma unroll
          for (int f = 0; f < 3; ++f) P.pos[3 * out + f] = __uint_as_float(o[f]);
        }
        if (P.v2vox) {
#pragma unroll
          for (int f = 0; f < 3; ++f) P.v2vox[3 * out + f] = (int32_t)o[3 + f];
        }
        if (P.ekey) P.ekey[out] = key;
      }
    }
  }
  wave_sync();
}

// A tile whose records did not fit the scratch, re-read from the volume 64
// cubes at a time (cube order) into the record form and emitted from voff[t].
template <int ELEM, bool BATCH>
__device__ __forceinline__ void emit_staged_tile(const ExtractParams& P, const uint8_t* s_slot,
                                                 const uint8_t* s_nv, uint32_t* s_w,
                                                 uint32_t t) {
  const int lane = threadIdx.x & 63;
  const int vi = BATCH ? volume_of_tile(P, t) : 0;
  const VolDesc& V = vol_at<BATCH>(P, vi);
  uint32_t* const s_meta = s_w + kPackedMeta;
  const uint32_t local = (uint32_t)(t - V.tile_begin);
  const int z = (int)V.divR.div(local);
  const int64_t p0 = (int64_t)(local - (uint32_t)z * (uint32_t)V.R) * kTile;
  int64_t out = P.voff[t];
  const int X = V.X;
  const int64_t XY = (int64_t)X * V.Y;
  const uint32_t* vol = V.vol + (int64_t)z * XY;
  for (int cb = 0; cb < kTile; cb += 64) {
    const int o = cb + lane;
    const int64_t p = p0 + o;
    uint32_t cw[8] = {}, ci = 0;
    if (p < XY) {
      const uint32_t y = V.divX.div((uint32_t)p);
      const uint32_t x = (uint32_t)p - y * (uint32_t)X;
      if (x + 1 < (uint32_t)X && y + 1 < (uint32_t)V.Y) {
        const uint32_t* pa = vol + p;
        const uint32_t* pb = pa + XY;
        cw[0] = pa[0]; cw[1] = pa[1]; cw[2] = pa[X + 1]; cw[3] = pa[X];
        cw[4] = pb[0]; cw[5] = pb[1]; cw[6] = pb[X + 1]; cw[7] = pb[X];
        ci = cube_index<ELEM>(cw);
      }
    }
    const uint32_t nvc = s_nv[ci];
    const uint64_t act = __ballot(nvc != 0);
    const int n = __popcll(act);
    const uint32_t vincl = wave_incl_scan32(nvc);  // lane order = record order
    if ((act >> lane) & 1ull) {
      const int rk = __popcll(act & ((1ull << lane) - 1ull));
      uint32_t* rp = s_w + rk * kRecWords;
      rp[0] = ((uint32_t)o << 8) | ci;
#pragma unroll
      for (int q = 0; q < 8; ++q) rp[1 + q] = cw[q];
      const uint64_t ob = (uint64_t)(out + (int64_t)(vincl - nvc));
      const uint32_t y = V.divX.div((uint32_t)p);
      uint32_t* mt = s_meta + 4 * rk;
      mt[0] = (uint32_t)ob;
      mt[1] = meta_hi(ob, (uint32_t)z);
      mt[2] = (uint32_t)p - y * (uint32_t)X;
      mt[3] = y;
    }
    out += (int64_t)__builtin_amdgcn_readlane(vincl, 63);
    if (n)
      emit_block<ELEM, BATCH>(P, V, s_slot, s_nv, s_w, n,
                              GridDims{V.X, V.Y, V.Z, V.h, (uint32_t)vi});
  }
}

// Single volumes, balanced by records: the records of the list's tiles, in
// list order, are cut into units of 64 (the scan's clist / lpre / units
// tables) and wave w takes units w, w + nwaves, ..., then the staged-path
// tiles.  In a unit, lane l holds clist entry units[k] + l (every entry has
// records, so the at most 64 entries overlapping the unit are consecutive;
// the first may start before it); each entry marks its
// first record in the unit and a max-scan gives every record lane its entry.
// A record's output index is its tile's first vertex, plus the vertices of
// the tile's records before it: those in the unit come from a scan of the
// unit's vertex counts, those before the unit (first entry only) are summed
// from their case bytes.
// Seven workgroups per CU (22.8 KB of LDS each).  Batched launches: every
// clist entry carries its volume, and each record lane holds its volume's
// grid and index for the vertex lanes.
template <int ELEM, bool BATCH>
__global__ __launch_bounds__(kEmitThreads) __attribute__((amdgpu_waves_per_eu(6)))
void emit_units_kernel(ExtractParams P) {
  __shared__ __attribute__((aligned(16))) uint8_t s_slot[256 * 16];
  __shared__ uint8_t s_nv[256];
  __shared__ uint32_t s_wave[kEmitThreads / 64][kPackedWords];
  __shared__ uint8_t s_mark[kEmitThreads / 64][64];  // lane + 1 of the entry starting here
  static_assert(kEmitThreads * 16 == sizeof(SlotTable), "16 table bytes per thread");
  reinterpret_cast<u32x4*>(s_slot)[threadIdx.x] =
      reinterpret_cast<const u32x4*>(dSlot.v)[threadIdx.x];
  s_nv[threadIdx.x] = dNumVerts[threadIdx.x];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  s_mark[wv][lane] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0 && P.hdr->error) {
    // a look-back of the scan (any workgroup) hit its wait limit: the
    // offsets are not to be trusted, and the count says so
    P.hdr->count = -1;
    if (P.d_count) *P.d_count = -1;
  }
  const VolDesc& V = P.vol0;  // single volumes
  __syncthreads();  // the last workgroup barrier
  uint32_t* const s_w = s_wave[wv];
  uint8_t* const mark = s_mark[wv];
  uint32_t* const s_meta = s_w + kPackedMeta;
  const uint32_t nl = P.hdr->nz_total ? P.hdr->list_n : 0u;  // clist entries
  const int64_t total = nl ? P.lpre[nl] : 0;  // records of the list
  const int64_t n_units = (total + 63) >> 6;
  const int64_t n_work = n_units + (int64_t)P.hdr->dense_n;
  const int64_t nwaves = (int64_t)gridDim.x * (kEmitThreads / 64);
  for (int64_t k = (int64_t)blockIdx.x * (kEmitThreads / 64) + wv; k < n_work; k += nwaves) {
    if (k >= n_units) {
      emit_staged_tile<ELEM, BATCH>(P, s_slot, s_nv, s_w, P.dense_list[k - n_units]);
      continue;
    }
    const int64_t g0 = k << 6;
    const int R = (int)min((int64_t)64, total - g0);
    // lane l: clist entry units[k] + l, its records [pre, pre + nrec)
    const uint32_t i = P.units[k] + (uint32_t)lane;
    const uint32_t ic = min(i, nl);
    const int64_t pre = P.lpre[ic];
    const int64_t nrec = i < nl ? P.lpre[ic + 1] - pre : 0;
    const uint32_t t = i < nl ? P.clist[i] : 0u;
    const int64_t trec = i < nl ? P.lrs[i] : 0;  // the tile's first record
    const uint32_t vi = BATCH && i < nl ? P.clist_vol[i] : 0u;
    const bool inu = nrec > 0 && pre < g0 + R;  // the entry has records in the unit
    const int first = (int)max(pre - g0, (int64_t)0);  // its first record in the unit
    int64_t rs = 0, off = 0;
    uint32_t p0 = 0, z = 0;
    if (inu) {
      rs = trec + (g0 + first - pre);  // the record of unit position `first`
      off = P.voff[t];
      const VolDesc& Vl = vol_at<BATCH>(P, (int)vi);
      const uint32_t local = (uint32_t)(t - Vl.tile_begin);
      z = Vl.divR.div(local);
      p0 = (local - z * (uint32_t)Vl.R) * (uint32_t)kTile;
      mark[first] = (uint8_t)(lane + 1);
    }
    // records of the first entry's tile before the unit: their vertices
    const int64_t head = g0 - __shfl(pre, 0);
    const int64_t rs0 = __shfl(rs, 0) - head;  // the tile's first record
    wave_sync();
    uint32_t mk = 0;
    if (lane < R) {
      mk = mark[lane];
      mark[lane] = 0;
    }
    const int kk = lane < R ? (int)wave_incl_max32(mk) - 1 : 0;
    const int fr = __shfl(first, kk);
    const int64_t rsk = __shfl(rs, kk);
    uint32_t rec[kRecWords];
    if (lane < R) {
      const uint32_t* src = P.rec + (rsk + (lane - fr)) * kRecWords;
#pragma unroll
      for (int q = 0; q < kRecWords; ++q) rec[q] = src[q];
    } else {
#pragma unroll
      for (int q = 0; q < kRecWords; ++q) rec[q] = 0;
    }
    // a tile has at most kTile records: head < 512, eight loads per lane
    uint32_t hv = 0;
    if (head > 0) {
      uint32_t hw[kTile / 64];
#pragma unroll
      for (int q = 0; q < kTile / 64; ++q) {
        const int64_t j = lane + 64 * q;
        hw[q] = j < head ? P.rec[(rs0 + j) * kRecWords] : 0u;
      }
#pragma unroll
      for (int q = 0; q < kTile / 64; ++q) hv += lane + 64 * q < head ? s_nv[hw[q] & 0xffu] : 0u;
      hv = wave_sum32(hv);
    }
    const uint32_t nv = lane < R ? s_nv[rec[0] & 0xffu] : 0u;
    const uint32_t vincl = wave_incl_scan32(nv);
    const uint32_t vex = vincl - nv;
    const uint32_t vfirst = (uint32_t)__shfl((int)vex, fr);
    const int64_t offk = __shfl(off, kk) + (kk == 0 ? (int64_t)hv : 0);
    const uint32_t p0k = (uint32_t)__shfl((int)p0, kk);
    const uint32_t zk = (uint32_t)__shfl((int)z, kk);
    const uint32_t vik = BATCH ? (uint32_t)__shfl((int)vi, kk) : 0u;
    GridDims dl{0, 0, 0, 0.0f, 0u};
    if (lane < R) {
#pragma unroll
      for (int q = 0; q < kRecWords; ++q) s_w[lane * kRecWords + q] = rec[q];
      const uint64_t o = (uint64_t)(offk + (int64_t)(vex - vfirst));
      const uint32_t p = p0k + (rec[0] >> 8);
      const VolDesc& Vk = vol_at<BATCH>(P, (int)vik);
      const uint32_t y = Vk.divX.div(p);
      uint32_t* mt = s_meta + 4 * lane;
      mt[0] = (uint32_t)o;
      mt[1] = meta_hi(o, zk);
      mt[2] = p - y * (uint32_t)Vk.X;
      mt[3] = y;
      if constexpr (BATCH) dl = GridDims{Vk.X, Vk.Y, Vk.Z, Vk.h, vik};
    }
    emit_block<ELEM, BATCH>(P, V, s_slot, s_nv, s_w, R, dl);
  }
}


// ============================================================ backward
// Differentiable form (no counterpart in the reference).  For a vertex on
// edge (a -> b), p = p_a + t (p_b - p_a) with t = -f_a / d, d = f_b - f_a + eps,
// so dp/df_a = (p_b - p_a) * (-(f_b + eps) / d^2) and dp/df_b = (p_b - p_a) *
// (f_a / d^2).  BW1: one thread per active cube reduces g . dp/df over the
// cube's vertices, in slot order, into 8 per-corner sums.  BW2: every voxel
// sums the corner sums of the (at most 8) active cubes it is a corner of, in
// corner order.  No atomics, so the result is bitwise reproducible.
template <bool BATCH>
__global__ __launch_bounds__(kEmitThreads) void contrib_kernel(ExtractParams P, const float* gpos,
                                                               float* contrib) {
  __shared__ uint64_t s_case[256];
  __shared__ uint8_t s_nv[256];
  __shared__ int s_tmp[8];
  __shared__ TileInfo s_tile[kTilesPerContribGroup];
  const int tid = threadIdx.x;
  const int64_t t0 = (P.first_block + blockIdx.x) * kTilesPerContribGroup;
  s_case[tid] = dCase[tid];
  s_nv[tid] = dNumVerts[tid];
  if (tid < kTilesPerContribGroup) {
    const int64_t t = t0 + tid;
    TileInfo ti{};
    if (t < P.n_tiles) {
      ti.T = P.tileT[t];
      ti.nrec = P.tile_nrec[t];
      ti.rstart = P.tile_rec[t];
      ti.delta = P.voff[t];
      ti.vi = BATCH ? volume_of_tile(P, t) : 0;
    }
    s_tile[tid] = ti;
  }
  __syncthreads();
  if (tid == 0) {
    int lpos = 0;
    int64_t vpre = 0;
    for (int u = 0; u < kTilesPerContribGroup; ++u) {
      TileInfo& ti = s_tile[u];
      ti.lbeg = lpos;
      ti.delta -= vpre;
      if (ti.T > 0 && ti.nrec >= 0) {
        lpos += ti.nrec;
        vpre += ti.T;
      }
    }
    s_tmp[7] = lpos;
  }
  __syncthreads();
  const int nlist = s_tmp[7];
  // records were dropped (staged path) by the forward: no exact gradient;
  // make it loud (NaN) rather than silently wrong
  const float poison = P.hdr->dense_last ? __uint_as_float(0x7fc00000u) : 0.0f;
  int64_t carry = 0;
  for (int rb = 0; rb < nlist; rb += kEmitThreads) {
    const int k = rb + tid;
    int nv = 0, u = 0;
    uint32_t rc[kRecWords] = {};
    int64_t ri = 0;
    if (k < nlist) {
      while (u + 1 < kTilesPerContribGroup &&
             !(s_tile[u].T > 0 && s_tile[u].nrec >= 0 && k < s_tile[u].lbeg + s_tile[u].nrec))
        ++u;
      ri = s_tile[u].rstart + (k - s_tile[u].lbeg);
      const uint32_t* src = P.rec + ri * kRecWords;
#pragma unroll
      for (int q = 0; q < kRecWords; ++q) rc[q] = src[q];
      nv = s_nv[rc[0] & 0xffu];
    }
    int tot;
    const int ex = block_excl_scan(nv, s_tmp, &tot);
    if (k < nlist) {
      const TileInfo& ti = s_tile[u];
      const VolDesc& V = vol_at<BATCH>(P, ti.vi);
      const float h = V.h;
      const uint32_t ci = rc[0] & 0xffu;
      const int64_t first = ti.delta + carry + ex;
      float c[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) c[q] = 0.0f;
      for (int q = 0; q < nv; ++q) {
        const int m = 3 * (q / 3) + 2 - (q % 3);
        const int e = (int)((s_case[ci] >> (4 * m)) & 0xfu);
        const int ca = (int)((kEdgeA >> (3 * e)) & 7u);
        const int cb = (int)((kEdgeB >> (3 * e)) & 7u);
        float fa = 0.f, fb = 0.f;
#pragma unroll
        for (int w = 0; w < 8; ++w) {
          fa = ca == w ? __uint_as_float(rc[1 + w]) : fa;
          fb = cb == w ? __uint_as_float(rc[1 + w]) : fb;
        }
        // p_b - p_a is one signed grid step along one axis (same fp32 values
        // as the forward's node coordinates)
        const float dpx = (float)(corner_dx(cb) - corner_dx(ca)) * h;
        const float dpy = (float)(corner_dy(cb) - corner_dy(ca)) * h;
        const float dpz = (float)(corner_dz(cb) - corner_dz(ca)) * h;
        const float* g = gpos + 3 * (first + q);
        const float sdot = g[0] * dpx + g[1] * dpy + g[2] * dpz;
        const float d = fb - fa + 1e-15f;
        const float inv = 1.0f / (d * d);
        const float ga = sdot * (-(fb + 1e-15f) * inv);
        const float gb = sdot * (fa * inv);
#pragma unroll
        for (int w = 0; w < 8; ++w) c[w] += (ca == w ? ga : 0.0f) + (cb == w ? gb : 0.0f);
      }
      float4* dst = reinterpret_cast<float4*>(contrib + ri * 8);
      dst[0] = make_float4(c[0] + poison, c[1] + poison, c[2] + poison, c[3] + poison);
      dst[1] = make_float4(c[4] + poison, c[5] + poison, c[6] + poison, c[7] + poison);
    }
    carry += tot;
  }
}

// Backward BW2: dL/dv of every voxel = the sum, over the up to 8 cubes it
// is corner k of (k = 0..7 in corner order), of the cube's contribution k.
// A workgroup takes kGatherSpan consecutive voxels of one plane z.  The
// cubes they are corners of lie in planes z - 1 and z and, per plane, in two
// in-plane windows (rows y and y - 1: cube ids [q0 - 1, q0 + span) and the
// same minus X), each within 4 consecutive tiles: the workgroup stages those
// tiles' active-cube mask words, the popcount prefix of each word and the
// tiles' first records in LDS, so a voxel's lookups read LDS and only active
// cubes read their contribution (record index = tile's first record + rank
// of the cube's bit).  A wave takes 64 consecutive voxels at a time: each
// lane finds which of its (at most 8) cubes are active (the cube (x - dx,
// y - dy) has in-plane index q - dx - dy X), the wave packs the active
// (voxel, corner) pairs onto consecutive lanes (about 1.6 % of the cubes are
// active, so one round of 64 lanes usually takes them all) to rank them and
// read their contributions, and each voxel lane then adds its values in
// corner order, the order of the per-corner sum, so the bits are the same.
constexpr int kGatherSpan = 1024;  // voxels per workgroup (4 per thread)
// Batched launches: workgroup blk belongs to the volume whose gather range
// [gblk_begin, next) holds it, and writes that volume's part of the
// concatenated gradient (from vox_begin).  With maxg, the largest |dL/dv| of
// the launch goes there as float bits (atomicMax; order-independent).
template <bool BATCH>
__global__ __launch_bounds__(256) void gather_kernel(ExtractParams P, const float* contrib,
                                                     float* gvol, uint32_t* maxg) {
  __shared__ uint64_t s_mask[2][2][4][kCPW];  // [dz][dy][tile slot][word]
  __shared__ uint32_t s_pre[2][2][4][kCPW];   // set bits of the tile's earlier words
  __shared__ int64_t s_rec[2][2][4];          // first record of the tile
  __shared__ uint32_t s_max[4];
  __shared__ uint16_t s_pair[4][64 * 8];  // per wave: lane << 3 | corner of each active pair
  __shared__ float s_val[4][64 * 8];      // the pair's contribution
  int vi = 0;
  const int64_t gblk = P.first_block + blockIdx.x;
  uint32_t blk = (uint32_t)gblk;  // single volumes: below 2^31
  if constexpr (BATCH) {
    vi = batch_volume_of<&BatchVol::gblk_begin>(P.bvol, P.nvol, gblk);
    blk = (uint32_t)(gblk - P.bvol[vi].gblk_begin);
    gvol += P.bvol[vi].vox_begin;
  }
  const VolDesc& V = vol_at<BATCH>(P, vi);
  const int X = V.X, Y = V.Y, Z = V.Z;
  const uint32_t XY = (uint32_t)X * (uint32_t)Y;  // < 2^30 (plane_ok)
  const uint32_t per_plane = (XY + kGatherSpan - 1) / kGatherSpan;
  const int z = (int)(blk / per_plane);
  const uint32_t q0 = (blk - (uint32_t)z * per_plane) * kGatherSpan;
  const int tid = threadIdx.x;
  // first tile of each window (cube ids below 0 are never looked up)
  int base[2];
#pragma unroll
  for (int dy = 0; dy < 2; ++dy) {
    const int64_t lo = (int64_t)q0 - 1 - (dy ? X : 0);
    base[dy] = (int)(max(lo, (int64_t)0) >> 9);
  }
  uint64_t m = 0;  // the staged mask word of threads 0..127
  if (tid < 128) {
    const int dz = tid >> 6, dy = (tid >> 5) & 1, slot = (tid >> 3) & 3, word = tid & 7;
    const int cz = z - dz, r = base[dy] + slot;
    if (cz >= 0 && cz <= Z - 2 && r < V.R)
      m = P.masks[(V.tile_begin + (int64_t)cz * V.R + r) * kCPW + word];
    s_mask[dz][dy][slot][word] = m;
  } else if (tid < 128 + 16) {
    const int e = tid - 128, dz = e >> 3, dy = (e >> 2) & 1, slot = e & 3;
    const int cz = z - dz, r = base[dy] + slot;
    int64_t rs = 0;
    if (cz >= 0 && cz <= Z - 2 && r < V.R) rs = P.tile_rec[V.tile_begin + (int64_t)cz * V.R + r];
    s_rec[dz][dy][slot] = rs;
  }
  if (!__syncthreads_or(m != 0)) {
    // no active cube has a corner in the span: its gradient is zero (the
    // sums below would all be empty), and it raises no maximum
    float* const gz0 = gvol + (int64_t)z * XY;
#pragma unroll
    for (int j = 0; j < kGatherSpan / 256; ++j) {
      const uint32_t q = q0 + (uint32_t)(j * 256 + tid);
      if (q < XY) gz0[q] = 0.0f;
    }
    return;
  }
  if (tid < 16) {
    const int dz = tid >> 3, dy = (tid >> 2) & 1, slot = tid & 3;
    uint32_t pre = 0;
#pragma unroll
    for (int w = 0; w < kCPW; ++w) {
      s_pre[dz][dy][slot][w] = pre;
      pre += (uint32_t)__popcll(s_mask[dz][dy][slot][w]);
    }
  }
  __syncthreads();
  float* const gz = gvol + (int64_t)z * XY;
  uint32_t gmax = 0;  // largest |g| bits of this thread
  const int lane = tid & 63, wv = tid >> 6;
  uint16_t* const pairs = s_pair[wv];
  float* const vals = s_val[wv];
#pragma unroll 1
  for (int j = 0; j < kGatherSpan / 256; ++j) {
    // which of the (at most 8) cubes having this voxel as a corner are active
    const uint32_t qb = q0 + (uint32_t)(j * 256 + wv * 64);
    const uint32_t q = qb + (uint32_t)lane;
    const bool inq = q < XY;
    uint32_t act = 0;
    if (inq) {
      const uint32_t y = V.divX.div(q);
      const int x = (int)(q - y * (uint32_t)X);
      // cube (x - dx, y - dy) has in-plane index q - dx - dy X
      const bool okx[2] = {x <= X - 2, x >= 1}, oky[2] = {(int)y <= Y - 2, y >= 1};
      const bool okz[2] = {z <= Z - 2, z >= 1};
#pragma unroll
      for (int dy = 0; dy < 2; ++dy) {
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
          if (!(okx[dx] && oky[dy])) continue;
          const uint32_t p = q - (uint32_t)dx - (dy ? (uint32_t)X : 0u);
          const int slot = (int)(p >> 9) - base[dy];
          const uint32_t word = (p >> 6) & 7u, bit = p & 63u;
          const int c = dy ? (dx ? 2 : 3) : dx;  // corner index in the plane
#pragma unroll
          for (int dz = 0; dz < 2; ++dz)
            if (okz[dz]) act |= (uint32_t)((s_mask[dz][dy][slot][word] >> bit) & 1ull) << (4 * dz + c);
        }
      }
    }
    // the wave's active (voxel, corner) pairs, packed in corner order per voxel
    const uint32_t cnt = (uint32_t)__popc(act);
    const uint32_t incl = wave_incl_scan32(cnt);
    const uint32_t excl = incl - cnt;
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane(incl, 63);
    {
      uint32_t a = act, e = excl;
      while (a) {
        const int k = __builtin_ctz(a);
        a &= a - 1u;
        pairs[e++] = (uint16_t)((lane << 3) | k);
      }
    }
    wave_sync();
    for (uint32_t pr = (uint32_t)lane; pr < total; pr += 64) {
      const uint32_t d = pairs[pr];
      const int k = (int)(d & 7u);
      const uint32_t ql = qb + (d >> 3);
      const int dx = corner_dx(k), dy = corner_dy(k), dz = corner_dz(k);
      const uint32_t p = ql - (uint32_t)dx - (dy ? (uint32_t)X : 0u);
      const int slot = (int)(p >> 9) - base[dy];
      const int word = (int)((p >> 6) & 7u), bit = (int)(p & 63u);
      const uint64_t m = s_mask[dz][dy][slot][word];
      const uint32_t rank = s_pre[dz][dy][slot][word] + (uint32_t)__popcll(m & ((1ull << bit) - 1ull));
      vals[pr] = contrib[(s_rec[dz][dy][slot] + rank) * 8 + k];
    }
    wave_sync();
    if (inq) {
      // the voxel's sum in corner order, as the per-corner loop adds it
      float g = 0.0f;
      uint32_t a = act, e = excl;
      while (a) {
        a &= a - 1u;
        g += vals[e++];
      }
      gz[q] = g;
      gmax = max(gmax, __float_as_uint(g) & 0x7fffffffu);
    }
    wave_sync();
  }
  if (maxg) {
    // one atomic per workgroup, only where it can raise the maximum (the
    // plain read is a filter: a stale value only lets an atomic through)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) gmax = max(gmax, (uint32_t)__shfl_xor((int)gmax, o));
    if ((tid & 63) == 0) s_max[tid >> 6] = gmax;
    __syncthreads();
    if (tid == 0) {
      const uint32_t m = max(max(s_max[0], s_max[1]), max(s_max[2], s_max[3]));
      if (m > __hip_atomic_load(maxg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(maxg, m);
    }
  }
}

// ---- per-cube classification (independent of the pipeline) ---------------
template <int ELEM>
__global__ __launch_bounds__(256) void classify_kernel(const uint32_t* __restrict__ vol, int X,
                                                       int Y, int Z, uint8_t* ci_out,
                                                       uint8_t* nv_out) {
  const int64_t N = (int64_t)X * Y * Z;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t rr = i / X;
    const int x = (int)(i - rr * X);
    const int z = (int)(rr / Y);
    const int y = (int)(rr - (int64_t)z * Y);
    uint32_t ci = 0, nv = 0;
    if (x < X - 1 && y < Y - 1 && z < Z - 1) {
      const int64_t XY = (int64_t)X * Y;
      const uint32_t c[8] = {vol[i], vol[i + 1], vol[i + X + 1], vol[i + X],
                             vol[i + XY], vol[i + XY + 1], vol[i + XY + X + 1], vol[i + XY + X]};
      ci = cube_index<ELEM>(c);
      nv = dNumVerts[ci];
    }
    if (ci_out) ci_out[i] = (uint8_t)ci;
    if (nv_out) nv_out[i] = (uint8_t)nv;
  }
}

// ================================================================ host side
size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// The backward's two launches over nb workgroups, in launches of at most
// kMaxLaunchBlocks workgroups (each given its first workgroup).
template <bool BATCH>
int launch_contrib(ExtractParams P, int64_t nb, const float* gpos, float* contrib,
                   hipStream_t s) {
  for (int64_t b0 = 0; b0 < nb; b0 += kMaxLaunchBlocks) {
    P.first_block = b0;
    hipLaunchKernelGGL(contrib_kernel<BATCH>, dim3((unsigned)std::min(kMaxLaunchBlocks, nb - b0)),
                       dim3(kEmitThreads), 0, s, P, gpos, contrib);
    MCX_HIP(hipGetLastError());
  }
  return MCX_OK;
}

template <bool BATCH>
int launch_gather(ExtractParams P, int64_t nb, const float* contrib, float* gvol,
                  uint32_t* maxg, hipStream_t s) {
  for (int64_t b0 = 0; b0 < nb; b0 += kMaxLaunchBlocks) {
    P.first_block = b0;
    hipLaunchKernelGGL(gather_kernel<BATCH>, dim3((unsigned)std::min(kMaxLaunchBlocks, nb - b0)),
                       dim3(256), 0, s, P, contrib, gvol, maxg);
    MCX_HIP(hipGetLastError());
  }
  return MCX_OK;
}

// Byte offsets of the count kernel's plane loads (in-plane word + X + 1024)
// stay below 2^32.
bool plane_ok(int X, int Y) { return (int64_t)X * Y + X + 1024 < ((int64_t)1 << 30); }
// ... and the emit kernels pack a cube's z in 24 bits.
bool size_ok(int X, int Y, int Z) { return plane_ok(X, Y) && Z <= (1 << 24); }

// The 128-bit count kernel addresses a column's planes (slab + 1, plus the
// prefetched ones past it) through one buffer resource with 32-bit offsets.
bool span_ok(int64_t XY, int slab) { return ((int64_t)slab + 4) * XY * 4 <= ((int64_t)1 << 31); }
// The most planes a column of that kernel may take (< 1: planes too large).
int64_t span_slab(int64_t XY) { return ((int64_t)1 << 31) / (XY * 4) - 4; }

int64_t tiles_per_plane(int X, int Y) { return ((int64_t)X * Y + kTile - 1) / kTile; }

int64_t tiles_of(int X, int Y, int Z) {
  if (X < 2 || Y < 2 || Z < 2) return 0;
  return tiles_per_plane(X, Y) * (Z - 1);
}

int64_t cubes_of(int X, int Y, int Z) {
  return (X > 1 && Y > 1 && Z > 1) ? (int64_t)(X - 1) * (Y - 1) * (Z - 1) : 0;
}

size_t fixed_ws_bytes(int64_t n_tiles) {
  const int64_t nscan = n_tiles > 0 ? (n_tiles + kScanItems - 1) / kScanItems : 1;
  return kWsHeader + al256(24 * (size_t)nscan) + 7 * al256(4 * (size_t)n_tiles) +
         3 * al256(8 * (size_t)n_tiles) + al256(8 * (size_t)kCPW * (size_t)n_tiles) +
         al256(8 * (size_t)(n_tiles + 1));
}

// The record area: records of 9 words and the unit table (one word per 64
// records, plus two).  rec_area_bytes(r) always holds r records.
size_t rec_area_bytes(int64_t records) {
  return (size_t)records * kRecWords * 4 + (size_t)(records + 15) / 16 + 512;
}
int64_t rec_cap_of(size_t rest) {
  return rest > 512 ? (int64_t)((rest - 512) * 16 / (kRecWords * 4 * 16 + 1)) : 0;
}

int layout_ws(void* d_ws, size_t ws_bytes, int64_t n_tiles, ExtractParams& P) {
  if (n_tiles >= ((int64_t)1 << 32)) return MCX_ERR_UNSUPPORTED;  // 32-bit tile ids
  const size_t fixed = fixed_ws_bytes(n_tiles);
  if (!d_ws || ws_bytes < fixed) return MCX_ERR_WORKSPACE;
  char* b = static_cast<char*>(d_ws);
  P.hdr = reinterpret_cast<WsHeader*>(b);
  size_t o = kWsHeader;
  P.cap_desc = n_tiles > 0 ? (n_tiles + kScanItems - 1) / kScanItems : 1;
  P.desc = reinterpret_cast<uint64_t*>(b + o);
  o += al256(24 * (size_t)P.cap_desc);  // three descriptor sets: tiles, list records, list entries
  P.tileT = reinterpret_cast<int32_t*>(b + o);
  o += al256(4 * (size_t)n_tiles);
  P.tile_nrec = reinterpret_cast<int32_t*>(b + o);
  o += al256(4 * (size_t)n_tiles);
  P.tile_rec = reinterpret_cast<int64_t*>(b + o);
  o += al256(8 * (size_t)n_tiles);
  P.voff = reinterpret_cast<int64_t*>(b + o);
  o += al256(8 * (size_t)n_tiles);
  P.masks = reinterpret_cast<uint64_t*>(b + o);
  o += al256(8 * (size_t)kCPW * (size_t)n_tiles);
  P.nz_list = reinterpret_cast<uint32_t*>(b + o);
  o += al256(4 * (size_t)n_tiles);
  P.dense_list = reinterpret_cast<uint32_t*>(b + o);
  o += al256(4 * (size_t)n_tiles);
  P.clist = reinterpret_cast<uint32_t*>(b + o);
  o += al256(4 * (size_t)n_tiles);
  P.nz_vol = reinterpret_cast<uint32_t*>(b + o);
  o += al256(4 * (size_t)n_tiles);
  P.clist_vol = reinterpret_cast<uint32_t*>(b + o);
  o += al256(4 * (size_t)n_tiles);
  P.lpre = reinterpret_cast<int64_t*>(b + o);
  o += al256(8 * (size_t)(n_tiles + 1));
  P.lrs = reinterpret_cast<int64_t*>(b + o);
  o += al256(8 * (size_t)n_tiles);
  P.list_scan = false;  // set by run_extract for single-volume emission
  const int64_t rc = rec_cap_of(ws_bytes - o);
  P.units = reinterpret_cast<uint32_t*>(b + o);
  o += al256(4 * (size_t)(rc / 64 + 2));
  P.rec = reinterpret_cast<uint32_t*>(b + o);
  P.rec_cap = rc;
  return MCX_OK;
}

// Compute units of the current device (cached per device).
int cu_count() {
  static int cache[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  if (dev >= 0 && dev < 64 && cache[dev] > 0) return cache[dev];
  int n = 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
    n = 256;
  if (dev >= 0 && dev < 64) cache[dev] = n;
  return n;
}

// Launch K1, K2, (vbase), K3 for prepared parameters; emit = false stops
// after the scan (records, masks and tile offsets only: the backward's
// recompute of the forward's per-cube records).
int run_extract(ExtractParams& P, int elem, bool batch, int maxX, hipStream_t s,
                bool emit = true) {
  (void)maxX;
  // record scratch: per-column quotas; the wide count kernel keeps a quarter
  // of the scratch as a shared overflow pool
  const bool wide = P.wide && elem == MCX_ELEM_SHORT2 && !P.write_masks;
  P.list_scan = emit;
  const int64_t pool = wide ? P.rec_cap / 4 : 0;
  P.quota = P.ncols > 0 ? (P.rec_cap - pool) / P.ncols : 0;
  if (P.quota > (int64_t)kTile * P.slab_max) P.quota = (int64_t)kTile * P.slab_max;
  P.ovf_begin = P.quota * P.ncols;
  P.ovf_cap = wide ? P.rec_cap - P.ovf_begin : 0;
  const unsigned k1_blocks = (unsigned)((P.ncols + 3) / 4);
  const unsigned k2_blocks = (unsigned)((P.n_tiles + kScanItems - 1) / kScanItems);
  const int64_t items = (P.n_tiles + 3) / 4;  // a workgroup is 4 waves, a wave takes tiles
  const int64_t k3_max = (int64_t)cu_count() * g_tuning.emit_groups_per_cu;
  const unsigned k3_blocks = (unsigned)(items < k3_max ? items : k3_max);
#define MCX_LAUNCH_SET(E, B)                                                                     \
  do {                                                                                           \
    if (wide) hipLaunchKernelGGL(count4_kernel<B>, dim3(k1_blocks), dim3(256), 0, s, P);        \
    else hipLaunchKernelGGL((count_kernel<E, B>), dim3(k1_blocks), dim3(256), 0, s, P);         \
    MCX_HIP(hipGetLastError());                                                                  \
    hipLaunchKernelGGL(scan_kernel, dim3(k2_blocks), dim3(256), 0, s, P);                        \
    MCX_HIP(hipGetLastError());                                                                  \
    if (B) {                                                                                     \
      hipLaunchKernelGGL(vbase_kernel, dim3((unsigned)((P.nvol + 256) / 256)), dim3(256), 0, s, \
                         P);                                                                     \
      MCX_HIP(hipGetLastError());                                                                \
    }                                                                                            \
    if (!emit) break;                                                                            \
    hipLaunchKernelGGL((emit_units_kernel<E, B>), dim3(k3_blocks), dim3(kEmitThreads), 0, s, P); \
    MCX_HIP(hipGetLastError());                                                                  \
  } while (0)
  if (elem == MCX_ELEM_SHORT2) {
    if (batch) MCX_LAUNCH_SET(MCX_ELEM_SHORT2, true);
    else MCX_LAUNCH_SET(MCX_ELEM_SHORT2, false);
  } else {
    if (batch) MCX_LAUNCH_SET(MCX_ELEM_F32, true);
    else MCX_LAUNCH_SET(MCX_ELEM_F32, false);
  }
#undef MCX_LAUNCH_SET
  return MCX_OK;
}

VolDesc make_vol(const void* d_vol, int X, int Y, int Z, float h, int16_t* ver_voxel,
                 int64_t tile_begin) {
  VolDesc v{};
  v.vol = static_cast<const uint32_t*>(d_vol);
  v.X = X;
  v.Y = Y;
  v.Z = Z;
  v.h = h;
  v.R = (int)tiles_per_plane(X, Y);
  v.tile_begin = tile_begin;
  v.ver_voxel = ver_voxel;
  v.divX = FastDiv((uint32_t)X);
  v.divR = FastDiv((uint32_t)v.R);
  return v;
}

// Single-volume launch parameters: workspace layout, the volume, and the
// column decomposition (R in-plane ranges x S slabs, about columns_per_cu
// waves per CU: several rounds of waves per SIMD, so that columns with much
// surface do not leave SIMDs idle at the end).
// Depends only on (X, Y, Z, ws_bytes), so a backward call sees the forward's
// record slots.
int prepare_single(ExtractParams& P, const void* d_vol, int X, int Y, int Z, float h,
                   int16_t* ver_voxel, void* d_ws, size_t ws_bytes) {
  if (!size_ok(X, Y, Z)) return MCX_ERR_UNSUPPORTED;
  const int64_t n_tiles = tiles_of(X, Y, Z);
  const int rc = layout_ws(d_ws, ws_bytes, n_tiles, P);
  if (rc) return rc;
  P.vol0 = make_vol(d_vol, X, Y, Z, h, ver_voxel, 0);
  P.vols = nullptr;
  P.cols = nullptr;
  P.nvol = 1;
  P.n_tiles = n_tiles;
  const int64_t R = P.vol0.R;
  int64_t target = (int64_t)cu_count() * g_tuning.columns_per_cu;
  // Smaller volumes take fewer, longer columns: about 8192 cubes a column,
  // at least 4 columns per CU.  A column reads one plane more than its slab
  // and fills its plane pipeline before its first step, costs that slabs of
  // 1-3 planes (what the per-CU target alone gives at 128^3 and 256^3) pay
  // every few planes: at 256^3 the fp32 forward took 0.115 ms with 3-plane
  // slabs and 0.079 ms with 16 (later 0.084 with 16 against 0.087 with 8
  // and 0.104 with 4, in one process); 128^3 0.039 ms with 1 and 0.032 with 4.
  const int64_t by_size =
      std::max((int64_t)cu_count() * 4, cubes_of(X, Y, Z) / 8192);
  if (by_size < target) target = by_size;
  int64_t S = (target + R - 1) / R;
  // large planes take more, shorter slabs, so that the 128-bit count kernel
  // still reaches a column's planes through one buffer resource
  const int64_t smax = span_slab((int64_t)X * Y);
  if (smax >= 1) S = std::max(S, (Z - 1 + smax - 1) / smax);
  if (S > Z - 1) S = Z - 1;
  P.slab = (int)((Z - 1 + S - 1) / S);
  S = (Z - 1 + P.slab - 1) / P.slab;
  P.ncols = R * S;
  P.slab_max = P.slab;
  P.wide = (reinterpret_cast<uintptr_t>(d_vol) & 3u) == 0 &&
           span_ok((int64_t)X * Y, P.slab);
  return MCX_OK;
}

int read_header(void* d_ws, WsHeader* hh, hipStream_t s) {
  MCX_HIP(hipMemcpyAsync(hh, d_ws, sizeof(*hh), hipMemcpyDeviceToHost, s));
  MCX_HIP(hipStreamSynchronize(s));
  return MCX_OK;
}

}  // namespace
}  // namespace mcx

using namespace mcx;

extern "C" size_t mcx_extract_workspace_size_ex(int X, int Y, int Z, int64_t records) {
  if (records < 0) records = 0;
  return fixed_ws_bytes(tiles_of(X, Y, Z)) + rec_area_bytes(records);
}

// Default record scratch: one record per 16 cubes (6.25 % of the cubes of a
// column may be active before its tiles fall back to the staged path).
extern "C" size_t mcx_extract_workspace_size(int X, int Y, int Z) {
  int64_t rec = cubes_of(X, Y, Z) / 16;
  if (rec < 65536) rec = 65536;
  return mcx_extract_workspace_size_ex(X, Y, Z, rec);
}

extern "C" int mcx_extract(const void* d_vol, int elem, int X, int Y, int Z, float h,
                           mcx_vertex_voxel* d_vv, float* d_pos, int32_t* d_v2vox,
                           int16_t* d_ver_voxel, int64_t cap, int64_t* d_count,
                           int64_t* h_count, void* d_ws, size_t ws_bytes, void* stream) {
  return mcx_extract_ex(d_vol, elem, X, Y, Z, h, d_vv, d_pos, d_v2vox, d_ver_voxel, nullptr, cap,
                        d_count, h_count, d_ws, ws_bytes, stream);
}

extern "C" int mcx_extract_ex(const void* d_vol, int elem, int X, int Y, int Z, float h,
                              mcx_vertex_voxel* d_vv, float* d_pos, int32_t* d_v2vox,
                              int16_t* d_ver_voxel, int64_t* d_edge_key, int64_t cap,
                              int64_t* d_count, int64_t* h_count, void* d_ws, size_t ws_bytes,
                              void* stream) {
  if (!d_vol || X < 1 || Y < 1 || Z < 1 || cap < 0) return MCX_ERR_ARG;
  if (elem != MCX_ELEM_SHORT2 && elem != MCX_ELEM_F32) return MCX_ERR_ARG;
  if (elem == MCX_ELEM_F32 && (d_vv || d_ver_voxel)) return MCX_ERR_ARG;
  if (!size_ok(X, Y, Z)) return MCX_ERR_UNSUPPORTED;
  hipStream_t s = as_stream(stream);
  int rc = MCX_OK;
  const int64_t N = (int64_t)X * Y * Z;
  if (d_ver_voxel) {
    rc = fill_bytes_zero(d_ver_voxel, (size_t)N * sizeof(int16_t), s);
    if (rc) return rc;
  }
  const int64_t n_tiles = tiles_of(X, Y, Z);
  if (n_tiles == 0) {
    if (d_count) {
      rc = fill_bytes_zero(d_count, sizeof(int64_t), s);
      if (rc) return rc;
    }
    if (h_count) {
      MCX_HIP(hipStreamSynchronize(s));
      *h_count = 0;
    }
    return MCX_OK;
  }
  ExtractParams P{};
  rc = prepare_single(P, d_vol, X, Y, Z, h, d_ver_voxel, d_ws, ws_bytes);
  if (rc) return rc;
  P.vv = d_vv;
  P.pos = d_pos;
  P.v2vox = d_v2vox;
  P.ekey = d_edge_key;
  P.cap = cap;
  P.d_count = d_count;
  P.vbase = nullptr;
  rc = run_extract(P, elem, false, X, s);
  if (rc) return rc;
  if (h_count) {
    WsHeader hh;
    MCX_HIP(hipMemcpyAsync(&hh, d_ws, sizeof(hh), hipMemcpyDeviceToHost, s));
    MCX_HIP(hipStreamSynchronize(s));
    *h_count = hh.count;
    if (hh.error) return take_error(d_ws, s);
    if (hh.count > cap) return MCX_ERR_CAPACITY;
  }
  return MCX_OK;
}

// ---- extraction without allocation (no counterpart in the reference) ----
// For callers that cannot allocate device memory (code under graph capture,
// tight serving loops): the workspace is this file-scope device array, one
// copy per GPU, held until the process ends.  Host code never passes the
// array's name to a kernel (on the host it only stands for the symbol): its
// device address is looked up once per device with hipGetSymbolAddress and
// kept below.  Every call on a device uses the same array, so calls that
// overlap on one GPU (different streams) race; calls ordered on one stream
// do not.
namespace {
constexpr size_t kStaticScratchBytes = MCX_STATIC_SCRATCH_BYTES;
__device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch array (looked up on first use).
int static_scratch(void** out) {
  int dev = 0;
  MCX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return MCX_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    MCX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return MCX_ERR_HIP;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return MCX_OK;
}
}  // namespace

extern "C" size_t mcx_extract_static_limit(int X, int Y, int Z) {
  if (X < 1 || Y < 1 || Z < 1 || !size_ok(X, Y, Z)) return 0;
  const size_t need = fixed_ws_bytes(tiles_of(X, Y, Z)) + rec_area_bytes(kStaticMinRecords);
  return need <= kStaticScratchBytes ? kStaticScratchBytes : 0;
}

extern "C" int mcx_extract_static(const void* d_vol, int elem, int X, int Y, int Z, float h,
                                  mcx_vertex_voxel* d_vv, float* d_pos, int32_t* d_v2vox,
                                  int16_t* d_ver_voxel, int64_t cap, int64_t* d_count,
                                  int64_t* h_count, void* stream) {
  if (!d_vol || X < 1 || Y < 1 || Z < 1 || cap < 0) return MCX_ERR_ARG;
  if (!mcx_extract_static_limit(X, Y, Z)) return MCX_ERR_UNSUPPORTED;
  void* ws = nullptr;
  int rc = static_scratch(&ws);
  if (rc) return rc;
  // the header's counters must start at zero: cleared by a fill kernel
  // (no memset node, no allocation); the rest is written before it is read
  rc = fill_bytes_zero(ws, kWsHeader, as_stream(stream));
  if (rc) return rc;
  return mcx_extract_ex(d_vol, elem, X, Y, Z, h, d_vv, d_pos, d_v2vox, d_ver_voxel, nullptr, cap,
                        d_count, h_count, ws, kStaticScratchBytes, stream);
}

extern "C" int mcx_classify(const void* d_vol, int elem, int X, int Y, int Z, uint8_t* d_ci,
                            uint8_t* d_nv, void* stream) {
  if (!d_vol || X < 1 || Y < 1 || Z < 1) return MCX_ERR_ARG;
  if (elem != MCX_ELEM_SHORT2 && elem != MCX_ELEM_F32) return MCX_ERR_ARG;
  hipStream_t s = as_stream(stream);
  const int64_t N = (int64_t)X * Y * Z;
  int64_t blocks = (N + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (elem == MCX_ELEM_SHORT2)
    hipLaunchKernelGGL(classify_kernel<MCX_ELEM_SHORT2>, dim3((unsigned)blocks), dim3(256), 0, s,
                       static_cast<const uint32_t*>(d_vol), X, Y, Z, d_ci, d_nv);
  else
    hipLaunchKernelGGL(classify_kernel<MCX_ELEM_F32>, dim3((unsigned)blocks), dim3(256), 0, s,
                       static_cast<const uint32_t*>(d_vol), X, Y, Z, d_ci, d_nv);
  MCX_HIP(hipGetLastError());
  return MCX_OK;
}

// ---- differentiable form (no counterpart in the reference) ---------------
extern "C" size_t mcx_backward_scratch_size(int X, int Y, int Z, size_t ws_bytes) {
  ExtractParams P{};
  const int64_t n_tiles = tiles_of(X, Y, Z);
  if (ws_bytes < fixed_ws_bytes(n_tiles)) return 0;
  char dummy;
  if (layout_ws(&dummy, ws_bytes, n_tiles, P)) return 0;
  return (size_t)P.rec_cap * 8 * sizeof(float);
}

extern "C" int mcx_extract_diff(const float* d_vol, int X, int Y, int Z, float h, float* d_pos,
                                int64_t cap, int64_t* d_count, int64_t* h_count, void* d_ws,
                                size_t ws_bytes, void* stream) {
  if (!d_vol || X < 1 || Y < 1 || Z < 1 || cap < 0) return MCX_ERR_ARG;
  if (!size_ok(X, Y, Z)) return MCX_ERR_UNSUPPORTED;
  hipStream_t s = as_stream(stream);
  const int64_t n_tiles = tiles_of(X, Y, Z);
  if (n_tiles == 0) {
    if (d_count && fill_bytes_zero(d_count, sizeof(int64_t), s)) return MCX_ERR_HIP;
    if (h_count) {
      MCX_HIP(hipStreamSynchronize(s));
      *h_count = 0;
    }
    return MCX_OK;
  }
  ExtractParams P{};
  int rc = prepare_single(P, d_vol, X, Y, Z, h, nullptr, d_ws, ws_bytes);
  if (rc) return rc;
  P.pos = d_pos;
  P.cap = cap;
  P.d_count = d_count;
  P.write_masks = true;
  rc = run_extract(P, MCX_ELEM_F32, false, X, s);
  if (rc) return rc;
  if (h_count) {
    WsHeader hh;
    rc = read_header(d_ws, &hh, s);
    if (rc) return rc;
    *h_count = hh.count;
    if (hh.error) return take_error(d_ws, s);
    if (hh.dense_last) return MCX_ERR_WORKSPACE;  // backward needs every record
    if (hh.count > cap) return MCX_ERR_CAPACITY;
  }
  return MCX_OK;
}

namespace {
int backward_impl(const float* d_vol, int X, int Y, int Z, float h, const float* d_grad_pos,
                  float* d_grad_vol, void* d_ws, size_t ws_bytes, void* d_scratch,
                  size_t scratch_bytes, void* stream, bool recompute) {
  if (!d_vol || !d_grad_vol || X < 1 || Y < 1 || Z < 1) return MCX_ERR_ARG;
  if (!size_ok(X, Y, Z)) return MCX_ERR_UNSUPPORTED;
  hipStream_t s = as_stream(stream);
  const int64_t N = (int64_t)X * Y * Z;
  const int64_t n_tiles = tiles_of(X, Y, Z);
  if (n_tiles == 0) return fill_bytes_zero(d_grad_vol, (size_t)N * sizeof(float), s);
  if (!d_grad_pos) return MCX_ERR_ARG;
  ExtractParams P{};
  int rc = prepare_single(P, d_vol, X, Y, Z, h, nullptr, d_ws, ws_bytes);
  if (rc) return rc;
  if (recompute) {
    // rebuild the per-cube records, active-cube masks and tile offsets of
    // the forward (count + scan, no emission); same layout for the same
    // (X, Y, Z, ws_bytes), so the same records as a checked forward
    P.write_masks = true;
    P.cap = 0;
    rc = run_extract(P, MCX_ELEM_F32, false, X, s, /*emit=*/false);
    if (rc) return rc;
  }
  if (!d_scratch || scratch_bytes < (size_t)P.rec_cap * 8 * sizeof(float)) return MCX_ERR_WORKSPACE;
  float* contrib = static_cast<float*>(d_scratch);
  const int64_t g1 = (n_tiles + kTilesPerContribGroup - 1) / kTilesPerContribGroup;
  rc = launch_contrib<false>(P, g1, d_grad_pos, contrib, s);
  if (rc) return rc;
  const int64_t g2 = ((int64_t)X * Y + kGatherSpan - 1) / kGatherSpan * Z;
  if (g2 >= ((int64_t)1 << 31)) return MCX_ERR_UNSUPPORTED;
  return launch_gather<false>(P, g2, contrib, d_grad_vol, nullptr, s);
}
}  // namespace

extern "C" int mcx_extract_backward(const float* d_vol, int X, int Y, int Z, float h,
                                    const float* d_grad_pos, float* d_grad_vol, void* d_ws,
                                    size_t ws_bytes, void* d_scratch, size_t scratch_bytes,
                                    void* stream) {
  return backward_impl(d_vol, X, Y, Z, h, d_grad_pos, d_grad_vol, d_ws, ws_bytes, d_scratch,
                       scratch_bytes, stream, false);
}

extern "C" int mcx_extract_backward_recompute(const float* d_vol, int X, int Y, int Z, float h,
                                              const float* d_grad_pos, float* d_grad_vol,
                                              void* d_ws, size_t ws_bytes, void* d_scratch,
                                              size_t scratch_bytes, void* stream) {
  return backward_impl(d_vol, X, Y, Z, h, d_grad_pos, d_grad_vol, d_ws, ws_bytes, d_scratch,
                       scratch_bytes, stream, true);
}

// ---- batched extraction (no counterpart in the reference) -----------------
// A plan holds the per-volume descriptors and the K1 column list in device
// memory (uploaded once); mcx_extract_batched then only launches kernels, so
// it can be captured in a hipGraph and replayed.
extern "C" int mcx_batch_create(mcx_batch** out, int nvol, const int* dims_xyz, const float* h,
                                const void* const* d_vols, int elem) {
  if (!out || nvol < 1 || !dims_xyz || !h || !d_vols) return MCX_ERR_ARG;
  if (elem != MCX_ELEM_SHORT2 && elem != MCX_ELEM_F32) return MCX_ERR_ARG;
  std::vector<VolDesc> vols(nvol);
  std::vector<BatchVol> bvol(nvol);
  std::vector<Column> cols;
  int64_t tb = 0, cubes = 0, voxels = 0, blocks = 0, gblocks = 0;
  int maxX = 1;
  bool wide = true;
  for (int b = 0; b < nvol; ++b) {
    const int X = dims_xyz[3 * b], Y = dims_xyz[3 * b + 1], Z = dims_xyz[3 * b + 2];
    if (X < 1 || Y < 1 || Z < 1 || !d_vols[b]) return MCX_ERR_ARG;
    if (!size_ok(X, Y, Z)) return MCX_ERR_UNSUPPORTED;
    vols[b] = make_vol(d_vols[b], X, Y, Z, h[b], nullptr, tb);
    const int64_t nv = (int64_t)X * Y * Z;
    bvol[b] = BatchVol{X, Y, Z, h[b], voxels, blocks, gblocks, FastDiv((uint32_t)X),
                       FastDiv((uint32_t)Y)};
    voxels += nv;
    blocks += (nv + 255) / 256;
    gblocks += ((int64_t)X * Y + kGatherSpan - 1) / kGatherSpan * Z;
    const int64_t nt = tiles_of(X, Y, Z);
    if (nt > 0) {
      // one column per in-plane range, slabs of the plan's planes per column
      // (fewer where the planes are large)
      const int R = vols[b].R;
      const int64_t smax = span_slab((int64_t)X * Y);
      const int slab = (int)(smax >= 1 ? std::min((int64_t)g_tuning.batch_slab, smax)
                                       : (int64_t)g_tuning.batch_slab);
      for (int z0 = 0; z0 < Z - 1; z0 += slab)
        for (int r = 0; r < R; ++r) cols.push_back(Column{b, r, z0, std::min(z0 + slab, Z - 1)});
    }
    tb += nt;
    cubes += cubes_of(X, Y, Z);
    maxX = std::max(maxX, X);
    wide = wide && (reinterpret_cast<uintptr_t>(d_vols[b]) & 3u) == 0 &&
           span_slab((int64_t)X * Y) >= 1;
  }
  mcx_batch* p = new mcx_batch();
  p->wide = wide;
  p->slab = g_tuning.batch_slab;
  p->nvol = nvol;
  p->elem = elem;
  p->maxX = maxX;
  p->n_tiles = tb;
  p->ncols = (int64_t)cols.size();
  p->cubes = cubes;
  p->voxels = voxels;
  p->blocks = blocks;
  p->gather_blocks = gblocks;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&p->d_vols), sizeof(VolDesc) * nvol);
  if (e == hipSuccess)
    e = hipMalloc(reinterpret_cast<void**>(&p->d_bvol), sizeof(BatchVol) * nvol);
  if (e == hipSuccess && !cols.empty())
    e = hipMalloc(reinterpret_cast<void**>(&p->d_cols), sizeof(Column) * cols.size());
  if (e == hipSuccess)
    e = hipMemcpy(p->d_vols, vols.data(), sizeof(VolDesc) * nvol, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(p->d_bvol, bvol.data(), sizeof(BatchVol) * nvol, hipMemcpyHostToDevice);
  if (e == hipSuccess && !cols.empty())
    e = hipMemcpy(p->d_cols, cols.data(), sizeof(Column) * cols.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(p->d_vols);
    (void)hipFree(p->d_bvol);
    (void)hipFree(p->d_cols);
    delete p;
    return hip_status(e);
  }
  *out = p;
  return MCX_OK;
}

extern "C" void mcx_batch_destroy(mcx_batch* p) {
  if (!p) return;
  (void)hipFree(p->d_vols);
  (void)hipFree(p->d_bvol);
  (void)hipFree(p->d_cols);
  delete p;
}

extern "C" int64_t mcx_batch_voxels(const mcx_batch* p) { return p ? p->voxels : 0; }

extern "C" int64_t mcx_batch_tiles(const mcx_batch* p) { return p ? p->n_tiles : 0; }

extern "C" size_t mcx_batch_workspace_size(const mcx_batch* p, int64_t records) {
  if (!p) return 0;
  if (records < 0) records = std::max<int64_t>(65536, p->cubes / 16);
  return fixed_ws_bytes(p->n_tiles) + rec_area_bytes(records);
}

namespace {
// Launch parameters of a batched call: workspace layout and the plan's tables.
int prepare_batch(ExtractParams& P, const mcx_batch* plan, void* d_ws, size_t ws_bytes) {
  int rc = layout_ws(d_ws, ws_bytes, plan->n_tiles, P);
  if (rc) return rc;
  P.vols = plan->d_vols;
  P.bvol = plan->d_bvol;
  P.cols = plan->d_cols;
  P.nvol = plan->nvol;
  P.n_tiles = plan->n_tiles;
  P.ncols = plan->ncols;
  P.slab_max = plan->slab;
  P.wide = plan->wide;
  return MCX_OK;
}
}  // namespace

extern "C" int mcx_extract_batched(const mcx_batch* plan, mcx_vertex_voxel* d_vv, float* d_pos,
                                   int32_t* d_v2vox, int64_t cap, int64_t* d_vbase,
                                   int64_t* d_count, int64_t* h_count, void* d_ws,
                                   size_t ws_bytes, void* stream) {
  if (!plan || !d_vbase || cap < 0) return MCX_ERR_ARG;
  if (plan->elem == MCX_ELEM_F32 && d_vv) return MCX_ERR_ARG;
  hipStream_t s = as_stream(stream);
  ExtractParams P{};
  int rc = prepare_batch(P, plan, d_ws, ws_bytes);
  if (rc) return rc;
  P.vv = d_vv;
  P.pos = d_pos;
  P.v2vox = d_v2vox;
  P.cap = cap;
  P.d_count = d_count;
  P.vbase = d_vbase;
  if (plan->n_tiles == 0) {
    // no cubes anywhere: every base (and the total) is 0
    rc = fill_bytes_zero(d_vbase, sizeof(int64_t) * (plan->nvol + 1), s);
    if (!rc && d_count) rc = fill_bytes_zero(d_count, sizeof(int64_t), s);
    if (rc) return rc;
    if (h_count) {
      MCX_HIP(hipStreamSynchronize(s));
      *h_count = 0;
    }
    return MCX_OK;
  }
  rc = run_extract(P, plan->elem, true, plan->maxX, s);
  if (rc) return rc;
  if (h_count) {
    WsHeader hh;
    rc = read_header(d_ws, &hh, s);
    if (rc) return rc;
    *h_count = hh.count;
    if (hh.error) return take_error(d_ws, s);
    if (hh.count > cap) return MCX_ERR_CAPACITY;
  }
  return MCX_OK;
}

// ---- batched differentiable extraction (no counterpart in the reference) -
extern "C" int mcx_extract_diff_batched(const mcx_batch* plan, float* d_pos, int64_t cap,
                                        int64_t* d_vbase, int64_t* d_count, int64_t* h_count,
                                        void* d_ws, size_t ws_bytes, void* stream) {
  if (!plan || !d_vbase || cap < 0) return MCX_ERR_ARG;
  if (plan->elem != MCX_ELEM_F32) return MCX_ERR_ARG;
  hipStream_t s = as_stream(stream);
  ExtractParams P{};
  int rc = prepare_batch(P, plan, d_ws, ws_bytes);
  if (rc) return rc;
  P.pos = d_pos;
  P.cap = cap;
  P.d_count = d_count;
  P.vbase = d_vbase;
  P.write_masks = true;
  if (plan->n_tiles == 0) {
    rc = fill_bytes_zero(d_vbase, sizeof(int64_t) * (plan->nvol + 1), s);
    if (!rc && d_count) rc = fill_bytes_zero(d_count, sizeof(int64_t), s);
    if (rc) return rc;
    if (h_count) {
      MCX_HIP(hipStreamSynchronize(s));
      *h_count = 0;
    }
    return MCX_OK;
  }
  rc = run_extract(P, MCX_ELEM_F32, true, plan->maxX, s);
  if (rc) return rc;
  if (h_count) {
    WsHeader hh;
    rc = read_header(d_ws, &hh, s);
    if (rc) return rc;
    *h_count = hh.count;
    if (hh.error) return take_error(d_ws, s);
    if (hh.dense_last) return MCX_ERR_WORKSPACE;  // the backward needs every record
    if (hh.count > cap) return MCX_ERR_CAPACITY;
  }
  return MCX_OK;
}

extern "C" size_t mcx_batch_backward_scratch_size(const mcx_batch* plan, size_t ws_bytes) {
  if (!plan || ws_bytes < fixed_ws_bytes(plan->n_tiles)) return 0;
  ExtractParams P{};
  char dummy;
  if (layout_ws(&dummy, ws_bytes, plan->n_tiles, P)) return 0;
  return (size_t)P.rec_cap * 8 * sizeof(float);
}

extern "C" int mcx_extract_backward_batched(const mcx_batch* plan, const float* d_grad_pos,
                                            float* d_grad_vol, uint32_t* d_maxabs, void* d_ws,
                                            size_t ws_bytes, void* d_scratch,
                                            size_t scratch_bytes, void* stream) {
  if (!plan || !d_grad_vol) return MCX_ERR_ARG;
  if (plan->elem != MCX_ELEM_F32) return MCX_ERR_ARG;
  hipStream_t s = as_stream(stream);
  int rc = MCX_OK;
  if (d_maxabs) {
    rc = fill_u32(d_maxabs, 0u, 1, s);
    if (rc) return rc;
  }
  if (plan->n_tiles == 0)
    return fill_bytes_zero(d_grad_vol, (size_t)plan->voxels * sizeof(float), s);
  if (!d_grad_pos) return MCX_ERR_ARG;
  ExtractParams P{};
  rc = prepare_batch(P, plan, d_ws, ws_bytes);
  if (rc) return rc;
  if (!d_scratch || scratch_bytes < (size_t)P.rec_cap * 8 * sizeof(float)) return MCX_ERR_WORKSPACE;
  if (plan->gather_blocks >= ((int64_t)1 << 31)) return MCX_ERR_UNSUPPORTED;
  float* contrib = static_cast<float*>(d_scratch);
  const int64_t g1 = (plan->n_tiles + kTilesPerContribGroup - 1) / kTilesPerContribGroup;
  rc = launch_contrib<true>(P, g1, d_grad_pos, contrib, s);
  if (rc) return rc;
  return launch_gather<true>(P, plan->gather_blocks, contrib, d_grad_vol, d_maxabs, s);
}

// ---- launch-shape tunables (no counterpart in the reference) ------------
extern "C" int mcx_set_tuning(int key, int value) {
  switch (key) {
    case MCX_TUNE_COLUMNS_PER_CU:
      if (value < 1 || value > 1024) return MCX_ERR_ARG;
      g_tuning.columns_per_cu = value;
      return MCX_OK;
    case MCX_TUNE_EMIT_GROUPS_PER_CU:
      if (value < 1 || value > 64) return MCX_ERR_ARG;
      g_tuning.emit_groups_per_cu = value;
      return MCX_OK;
    case MCX_TUNE_BATCH_SLAB:
      if (value < 1 || value > 4096) return MCX_ERR_ARG;
      g_tuning.batch_slab = value;
      return MCX_OK;
    default:
      return MCX_ERR_ARG;
  }
}

extern "C" int mcx_get_tuning(int key) {
  switch (key) {
    case MCX_TUNE_COLUMNS_PER_CU: return g_tuning.columns_per_cu;
    case MCX_TUNE_EMIT_GROUPS_PER_CU: return g_tuning.emit_groups_per_cu;
    case MCX_TUNE_BATCH_SLAB: return g_tuning.batch_slab;
    default: return -1;
  }
}
