// mcx_extract.hip -- fused single-pass marching-cubes extraction for gfx950.
//
// Replaces the reference's three synchronous stages in MarchingCubes::run
// (cu_marching_cubes_v1.cu:417-467):
//   getOccupiedVoxels              :97-206  classify + atomic compaction
//   computeOffsetsAndTotalVertexes :208-225  thrust exclusive scan + 2 D2H reads
//   generateTriangles              :228-396  per-active-cube triangle emission
// with one kernel that reads the volume once.  Each workgroup owns a tile of
// L = 256 * 4 * G consecutive cube linear ids (z*X*Y + y*X + x), stages the two
// voxel ranges it needs (plane z and plane z+1, each L + X + 1 voxels) in LDS,
// classifies and counts, obtains its output offset by a decoupled look-back
// over the tiles before it, and writes its vertices.  The tile order is the
// ticket order taken at workgroup start (forward progress), so the output is
// in canonical order (ascending cube id, then the reference's triangle slot
// order with reversed winding, :319-332) and is deterministic.
//
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

#include "mc_tables.inc"
#include "mcx_internal.h"

namespace mcx {
namespace {

__constant__ unsigned long long dCase[256] = MCX_CASE_TABLE_INIT;
__constant__ unsigned char dNumVerts[256] = MCX_NUM_VERTS_INIT;

// Edge e runs from corner kEa[e] to kEb[e] (:291-302), 3 bits per edge.
constexpr uint64_t kEdgeA = (0ull << 0) | (1ull << 3) | (2ull << 6) | (3ull << 9) | (4ull << 12) |
                            (5ull << 15) | (6ull << 18) | (7ull << 21) | (0ull << 24) |
                            (1ull << 27) | (2ull << 30) | (3ull << 33);
constexpr uint64_t kEdgeB = (1ull << 0) | (2ull << 3) | (3ull << 6) | (0ull << 9) | (5ull << 12) |
                            (6ull << 15) | (7ull << 18) | (4ull << 21) | (4ull << 24) |
                            (5ull << 27) | (6ull << 30) | (7ull << 33);

struct VolDesc {
  const uint32_t* vol;  // X*Y*Z raw voxels (short2 bits or fp32 bits)
  int X, Y, Z;
  float h;
  int64_t tile_begin;   // first tile of this volume in the launch
  int16_t* ver_voxel;   // optional, per volume
  FastDiv divX, divY;
};

struct ExtractParams {
  VolDesc single;
  const VolDesc* descs;  // batched launches: [nvol], sorted by tile_begin
  int nvol;
  int64_t n_tiles;
  int L;                 // cubes per tile
  int lds_plane;         // words per staged plane range
  mcx_vertex_voxel* vv;
  float* pos;
  int32_t* v2vox;
  int64_t cap;
  int64_t* d_count;
  int64_t* vbase;        // batched: per-volume first vertex, [nvol + 1]
  int64_t* tile_excl;    // optional: exclusive vertex offset of every tile
  WsHeader* hdr;
  uint64_t* desc;        // [2][cap_tiles]
  int64_t cap_tiles;
};

constexpr int kThreads = 256;
// LDS word offsets of the fixed part
constexpr int kLdsCase = 0;      // 256 x u64
constexpr int kLdsNv = 512;      // 256 x u8
constexpr int kLdsRed = 576;     // 8 x u64 wave partials
constexpr int kLdsBc = 592;      // 8 x i64 broadcast slots
constexpr int kLdsPlanes = 608;  // two staged voxel ranges

template <int ELEM>
__device__ __forceinline__ float unpack_f(uint32_t v) {
  if constexpr (ELEM == MCX_ELEM_SHORT2) {
    return (float)(int16_t)(v & 0xffffu) / 32766.0f;
  } else {
    return __uint_as_float(v);
  }
}

template <int ELEM>
__device__ __forceinline__ uint32_t neg_bit(uint32_t v) {
  if constexpr (ELEM == MCX_ELEM_SHORT2) {
    return (v >> 15) & 1u;  // tsdf < 0  <=>  tsdf / 32766 < 0
  } else {
    return __uint_as_float(v) < 0.0f ? 1u : 0u;
  }
}

// Cube index from the 8 corner words c[0..7] in the reference corner order.
template <int ELEM>
__device__ __forceinline__ uint32_t cube_index(const uint32_t c[8]) {
  uint32_t ci = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) ci |= neg_bit<ELEM>(c[k]) << k;
  if constexpr (ELEM == MCX_ELEM_SHORT2) {
    uint32_t mn = c[0];
#pragma unroll
    for (int k = 1; k < 8; ++k) mn = min(mn, c[k]);
    if (mn < 0x10000u) ci = 0;  // some weight == 0
  }
  return ci;
}

// Advance (x, y, z) by k cubes along the linear order (k < 2X).
__device__ __forceinline__ void step_coords(int X, int Y, int& x, int& y, int& z, int k) {
  x += k;
  while (x >= X) {
    x -= X;
    ++y;
  }
  while (y >= Y) {
    y -= Y;
    ++z;
  }
}

template <int ELEM>
__device__ __forceinline__ void store_point(const ExtractParams& P, const VolDesc& V,
                                            int64_t idx, float px, float py, float pz) {
  if (idx >= P.cap) return;
  const float h = V.h;
  float cx = px / h, cy = py / h, cz = pz / h;
  cx = cx - 0.5f;
  cy = cy - 0.5f;
  cz = cz - 0.5f;
  const int vx = (int)floorf(cx), vy = (int)floorf(cy), vz = (int)floorf(cz);
  const bool inb = vx >= 0 && vy >= 0 && vz >= 0 && vx < V.X && vy < V.Y && vz < V.Z;
  uint32_t tw = 0;
  int64_t li = 0;
  if (inb) {
    li = ((int64_t)vz * V.Y + vy) * V.X + vx;
    if constexpr (ELEM == MCX_ELEM_SHORT2) tw = V.vol[li];
  }
  if (P.vv) {
    uint32_t* r = reinterpret_cast<uint32_t*>(P.vv) + 7 * idx;
    r[0] = __float_as_uint(px);
    r[1] = __float_as_uint(py);
    r[2] = __float_as_uint(pz);
    r[3] = (uint32_t)vx;
    r[4] = (uint32_t)vy;
    r[5] = (uint32_t)vz;
    r[6] = tw;
  }
  if (P.pos) {
    float* q = P.pos + 3 * idx;
    q[0] = px;
    q[1] = py;
    q[2] = pz;
  }
  if (P.v2vox) {
    int32_t* q = P.v2vox + 3 * idx;
    q[0] = vx;
    q[1] = vy;
    q[2] = vz;
  }
  if (V.ver_voxel && inb) V.ver_voxel[li] = (int16_t)(tw & 0xffffu);
}

template <int G, bool BATCH, int ELEM>
__global__ __launch_bounds__(kThreads) void extract_kernel(ExtractParams P) {
  extern __shared__ uint32_t smem[];
  uint64_t* s_case = reinterpret_cast<uint64_t*>(smem + kLdsCase);
  uint8_t* s_nv = reinterpret_cast<uint8_t*>(smem + kLdsNv);
  uint64_t* s_red = reinterpret_cast<uint64_t*>(smem + kLdsRed);
  int64_t* s_bc = reinterpret_cast<int64_t*>(smem + kLdsBc);
  uint32_t* s_p0 = smem + kLdsPlanes;
  uint32_t* s_p1 = s_p0 + P.lds_plane;

  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  constexpr int L = kThreads * 4 * G;

  s_case[tid] = dCase[tid];
  s_nv[tid] = dNumVerts[tid];
  if (tid == 0) {
    const uint32_t tk =
        __hip_atomic_fetch_add(&P.hdr->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t par = __hip_atomic_load(&P.hdr->parity, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int vi = 0;
    if constexpr (BATCH) {
      int lo = 0, hi = P.nvol - 1;  // last volume with tile_begin <= tk
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (P.descs[mid].tile_begin <= (int64_t)tk) lo = mid; else hi = mid - 1;
      }
      vi = lo;
    }
    s_bc[0] = tk;
    s_bc[1] = par & 1u;
    s_bc[3] = vi;
  }
  __syncthreads();
  const int64_t t = s_bc[0];
  const uint32_t parity = (uint32_t)s_bc[1];
  const int vi = (int)s_bc[3];
  const VolDesc V = BATCH ? P.descs[vi] : P.single;
  uint64_t* desc = P.desc + (int64_t)parity * P.cap_tiles;
  uint64_t* desc_other = P.desc + (int64_t)(parity ^ 1u) * P.cap_tiles;

  const int X = V.X, Y = V.Y, Z = V.Z;
  const int64_t XY = (int64_t)X * Y;
  const int64_t N = XY * Z;
  const int64_t ncube_lin = XY * (Z - 1);
  const int64_t i0 = (t - V.tile_begin) * (int64_t)L;

  // tile start coordinates (x0, y0, z0)
  int x0, y0, z0;
  {
    const int64_t r = i0 / X;
    x0 = (int)(i0 - r * X);
    z0 = (int)(r / Y);
    y0 = (int)(r - (int64_t)z0 * Y);
  }

  // ---- stage voxel ranges [i0, i0+L+X+1) of plane z and z+1 in LDS
  int shift[2];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const int64_t start = i0 + p * XY;
    int64_t len64 = N - start;
    const int64_t want = (int64_t)L + X + 1;
    if (len64 > want) len64 = want;
    if (len64 < 0) len64 = 0;
    const int len = (int)len64;
    const uint32_t* g = V.vol + (start < N ? start : 0);
    const int sh = (int)((reinterpret_cast<uintptr_t>(g) >> 2) & 3u);
    shift[p] = sh;
    const uint4* ga = reinterpret_cast<const uint4*>(g - sh);
    uint32_t* dst = p == 0 ? s_p0 : s_p1;
    const int nchunks = (len + sh + 3) >> 2;
    for (int c = tid; c < nchunks; c += kThreads) {
      const int j0 = 4 * c - sh;
      if (j0 >= 0 && j0 + 4 <= len) {
        *reinterpret_cast<uint4*>(dst + 4 * c) = ga[c];
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int jj = j0 + q;
          dst[4 * c + q] = (jj >= 0 && jj < len) ? g[jj] : 0u;
        }
      }
    }
  }
  __syncthreads();

  // ---- classify + count: thread owns cubes o = g*1024 + 4*tid + k
  uint32_t cis[G];
  uint64_t packed = 0;  // 16-bit vertex count per group
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int o = g * (kThreads * 4) + 4 * tid;
    uint32_t ciw = 0, cnt = 0;
    if (i0 + o < ncube_lin) {
      int x = x0, y = y0, z = z0;
      {
        const uint32_t xs = (uint32_t)(x0 + o);
        const uint32_t q = V.divX.div(xs);
        x = (int)(xs - q * (uint32_t)X);
        const uint32_t ys = (uint32_t)y0 + q;
        const uint32_t q2 = V.divY.div(ys);
        y = (int)(ys - q2 * (uint32_t)Y);
        z = z0 + (int)q2;
      }
      const uint32_t* a = s_p0 + shift[0] + o;
      const uint32_t* b = s_p1 + shift[1] + o;
      uint32_t r0[5], r1[5], r2[5], r3[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        r0[k] = a[k];
        r1[k] = a[X + k];
        r2[k] = b[k];
        r3[k] = b[X + k];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        int xk = x, yk = y, zk = z;
        if (k) step_coords(X, Y, xk, yk, zk, k);
        const bool valid = xk < X - 1 && yk < Y - 1 && zk < Z - 1;
        const uint32_t c[8] = {r0[k], r0[k + 1], r1[k + 1], r1[k],
                               r2[k], r2[k + 1], r3[k + 1], r3[k]};
        uint32_t ci = valid ? cube_index<ELEM>(c) : 0u;
        const uint32_t nv = (ci - 1u) < 254u ? (uint32_t)s_nv[ci] : 0u;
        ci = nv ? ci : 0u;
        ciw |= ci << (8 * k);
        cnt += nv;
      }
    }
    cis[g] = ciw;
    packed |= (uint64_t)cnt << (16 * g);
  }

  // ---- workgroup exclusive scan of the packed group counts
  const uint64_t incl = wave_incl_scan(packed);
  if (lane == 63) s_red[wave] = incl;
  __syncthreads();
  uint64_t wprefix = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kThreads / kWave; ++w) {
    const uint64_t v = s_red[w];
    if (w < wave) wprefix += v;
    total += v;
  }
  const uint64_t texcl = incl - packed + wprefix;
  int64_t gbase[G];
  int64_t T = 0;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    gbase[g] = T;
    T += (int64_t)((total >> (16 * g)) & 0xffffu);
  }

  // ---- decoupled look-back (wave 0); waves 1..3 clear the other parity
  if (wave == 0) {
    int64_t excl = 0;
    if (t == 0) {
      if (lane == 0) st_agent(desc + t, kStatInc | (uint64_t)T);
    } else {
      if (lane == 0) st_agent(desc + t, kStatAgg | (uint64_t)T);
      excl = lookback(desc, t, &P.hdr->error);
      if (lane == 0) st_agent(desc + t, kStatInc | (uint64_t)(excl + T));
    }
    if (lane == 0) {
      s_bc[2] = excl;
      if (P.tile_excl) P.tile_excl[t] = excl;
      if constexpr (BATCH) {
        if (t == V.tile_begin) {
          P.vbase[vi] = excl;
          for (int b = vi - 1; b >= 0 && P.descs[b].tile_begin == V.tile_begin; --b) P.vbase[b] = excl;
        }
      }
      if (t == P.n_tiles - 1) {
        P.hdr->count = excl + T;
        if (P.d_count) *P.d_count = excl + T;
        if constexpr (BATCH) {
          for (int b = P.nvol; b > vi; --b) P.vbase[b] = excl + T;
        }
      }
    }
  } else {
    const int64_t k = (P.cap_tiles + P.n_tiles - 1) / P.n_tiles;
    const int64_t beg = t * k;
    int64_t end = beg + k;
    if (end > P.cap_tiles) end = P.cap_tiles;
    for (int64_t s = beg + (tid - kWave); s < end; s += kThreads - kWave) desc_other[s] = 0;
  }
  __syncthreads();
  const int64_t excl = s_bc[2];

  // ---- emit
  const float h = V.h;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const uint32_t ciw = cis[g];
    if (ciw == 0) continue;
    int64_t off = excl + gbase[g] + (int64_t)((texcl >> (16 * g)) & 0xffffu);
    const int o = g * (kThreads * 4) + 4 * tid;
    int x, y, z;
    {
      const uint32_t xs = (uint32_t)(x0 + o);
      const uint32_t q = V.divX.div(xs);
      x = (int)(xs - q * (uint32_t)X);
      const uint32_t ys = (uint32_t)y0 + q;
      const uint32_t q2 = V.divY.div(ys);
      y = (int)(ys - q2 * (uint32_t)Y);
      z = z0 + (int)q2;
    }
    for (int k = 0; k < 4; ++k) {
      const uint32_t ci = (ciw >> (8 * k)) & 0xffu;
      if (ci == 0) continue;
      int xk = x, yk = y, zk = z;
      if (k) step_coords(X, Y, xk, yk, zk, k);
      const uint64_t cs = s_case[ci];
      const int nverts = 3 * (int)(cs >> 60);
      const uint32_t* pl[2] = {s_p0 + shift[0] + o + k, s_p1 + shift[1] + o + k};
      for (int m = 0; m < nverts; ++m) {
        const int e = (int)((cs >> (4 * m)) & 0xfu);
        const int ca = (int)((kEdgeA >> (3 * e)) & 7u);
        const int cb = (int)((kEdgeB >> (3 * e)) & 7u);
        const int ax = ((ca + 1) >> 1) & 1, ay = (ca >> 1) & 1, az = ca >> 2;
        const int bx = ((cb + 1) >> 1) & 1, by = (cb >> 1) & 1, bz = cb >> 2;
        const float fa = unpack_f<ELEM>(pl[az][ax + ay * X]);
        const float fb = unpack_f<ELEM>(pl[bz][bx + by * X]);
        const float pax = ((float)(xk + ax) + 0.5f) * h;
        const float pay = ((float)(yk + ay) + 0.5f) * h;
        const float paz = ((float)(zk + az) + 0.5f) * h;
        const float pbx = ((float)(xk + bx) + 0.5f) * h;
        const float pby = ((float)(yk + by) + 0.5f) * h;
        const float pbz = ((float)(zk + bz) + 0.5f) * h;
        const float tt = (0.0f - fa) / (fb - fa + 1e-15f);
        const float px = __fmaf_rn(tt, pbx - pax, pax);
        const float py = __fmaf_rn(tt, pby - pay, pay);
        const float pz = __fmaf_rn(tt, pbz - paz, paz);
        // triangle corner m%3 of triangle m/3 goes to slot 2 - m%3 (:324-332)
        const int tri = m / 3, corner = m - 3 * tri;
        store_point<ELEM>(P, V, off + 3 * tri + (2 - corner), px, py, pz);
      }
      off += nverts;
    }
  }

  // ---- last workgroup resets the ticket/done counters and flips parity
  if (tid == 0) {
    const uint32_t prev =
        __hip_atomic_fetch_add(&P.hdr->done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((int64_t)prev == P.n_tiles - 1) {
      __hip_atomic_store(&P.hdr->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&P.hdr->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&P.hdr->parity, parity ^ 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// ---- per-cube classification (independent of the fused kernel) -----------
template <int ELEM>
__global__ __launch_bounds__(256) void classify_kernel(const uint32_t* __restrict__ vol, int X,
                                                       int Y, int Z, uint8_t* ci_out,
                                                       uint8_t* nv_out) {
  const int64_t N = (int64_t)X * Y * Z;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / X;
    const int x = (int)(i - r * X);
    const int z = (int)(r / Y);
    const int y = (int)(r - (int64_t)z * Y);
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

// ---- launch plumbing -------------------------------------------------------
template <int G, bool BATCH, int ELEM>
int launch_extract(const ExtractParams& P, int lds_bytes, hipStream_t s) {
  auto k = extract_kernel<G, BATCH, ELEM>;
  MCX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes));
  hipLaunchKernelGGL(k, dim3((unsigned)P.n_tiles), dim3(kThreads), lds_bytes, s, P);
  MCX_HIP(hipGetLastError());
  return MCX_OK;
}

constexpr int kMaxLds = 160 * 1024;

// Words per staged range: L + X + 1 voxels, plus up to 3 words of alignment
// shift in front, plus the rounding of the last 16-byte chunk (the staging
// loop writes indices up to len + shift + 2 <= L + X + 6); reads stay below
// L + X + 4.
int lds_bytes_for(int G, int maxX) {
  const int L = kThreads * 4 * G;
  const int need = L + maxX + 1 + 3 + 3;
  const int plane = ((need + 3) / 4) * 4;
  return (kLdsPlanes + 2 * plane) * 4;
}

int64_t tiles_for(int X, int Y, int Z, int L) {
  if (X < 2 || Y < 2 || Z < 2) return 0;
  const int64_t ncl = (int64_t)X * Y * (Z - 1);
  return (ncl + L - 1) / L;
}

}  // namespace

// Default tile: G = 4 (4096 cubes per workgroup), smaller when X is so wide
// that the staged ranges would not fit in LDS.
static int pick_groups(int maxX) {
  for (int G = 4; G >= 1; G >>= 1)
    if (lds_bytes_for(G, maxX) <= kMaxLds) return G;
  return 0;
}

}  // namespace mcx

using namespace mcx;

extern "C" size_t mcx_extract_workspace_size(int X, int Y, int Z) {
  // smallest tile (G = 1) gives the most tiles
  return ws_bytes_for_tiles(tiles_for(X, Y, Z, kThreads * 4));
}

extern "C" int mcx_extract(const void* d_vol, int elem, int X, int Y, int Z, float h,
                           mcx_vertex_voxel* d_vv, float* d_pos, int32_t* d_v2vox,
                           int16_t* d_ver_voxel, int64_t cap, int64_t* d_count,
                           int64_t* h_count, void* d_ws, size_t ws_bytes, void* stream) {
  if (!d_vol || X < 1 || Y < 1 || Z < 1 || cap < 0) return MCX_ERR_ARG;
  if (elem != MCX_ELEM_SHORT2 && elem != MCX_ELEM_F32) return MCX_ERR_ARG;
  if (elem == MCX_ELEM_F32 && (d_vv || d_ver_voxel)) return MCX_ERR_ARG;
  if ((int64_t)X * Y * Z > (int64_t)1 << 40) return MCX_ERR_UNSUPPORTED;
  hipStream_t s = as_stream(stream);
  int rc = MCX_OK;
  const int64_t N = (int64_t)X * Y * Z;
  if (d_ver_voxel) {
    rc = fill_bytes_zero(d_ver_voxel, (size_t)N * sizeof(int16_t), s);
    if (rc) return rc;
  }
  const int G = pick_groups(X);
  if (G == 0) return MCX_ERR_UNSUPPORTED;
  const int L = kThreads * 4 * G;
  const int64_t n_tiles = tiles_for(X, Y, Z, L);
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
  if (!d_ws) return MCX_ERR_WORKSPACE;
  const int64_t cap_tiles = ws_cap_tiles(ws_bytes);
  if (cap_tiles < n_tiles) return MCX_ERR_WORKSPACE;
  if (n_tiles > 0xffffffffll) return MCX_ERR_UNSUPPORTED;

  ExtractParams P{};
  P.single.vol = static_cast<const uint32_t*>(d_vol);
  P.single.X = X;
  P.single.Y = Y;
  P.single.Z = Z;
  P.single.h = h;
  P.single.tile_begin = 0;
  P.single.ver_voxel = d_ver_voxel;
  P.single.divX = FastDiv((uint32_t)X);
  P.single.divY = FastDiv((uint32_t)Y);
  P.descs = nullptr;
  P.nvol = 1;
  P.n_tiles = n_tiles;
  P.L = L;
  const int lds = lds_bytes_for(G, X);
  P.lds_plane = (lds / 4 - kLdsPlanes) / 2;
  P.vv = d_vv;
  P.pos = d_pos;
  P.v2vox = d_v2vox;
  P.cap = cap;
  P.d_count = d_count;
  P.vbase = nullptr;
  P.tile_excl = nullptr;
  P.hdr = static_cast<WsHeader*>(d_ws);
  P.desc = reinterpret_cast<uint64_t*>(static_cast<char*>(d_ws) + kWsHeader);
  P.cap_tiles = cap_tiles;

  if (elem == MCX_ELEM_SHORT2) {
    switch (G) {
      case 4: rc = launch_extract<4, false, MCX_ELEM_SHORT2>(P, lds, s); break;
      case 2: rc = launch_extract<2, false, MCX_ELEM_SHORT2>(P, lds, s); break;
      default: rc = launch_extract<1, false, MCX_ELEM_SHORT2>(P, lds, s); break;
    }
  } else {
    switch (G) {
      case 4: rc = launch_extract<4, false, MCX_ELEM_F32>(P, lds, s); break;
      case 2: rc = launch_extract<2, false, MCX_ELEM_F32>(P, lds, s); break;
      default: rc = launch_extract<1, false, MCX_ELEM_F32>(P, lds, s); break;
    }
  }
  if (rc) return rc;
  if (h_count) {
    WsHeader hh;
    MCX_HIP(hipMemcpyAsync(&hh, d_ws, sizeof(hh), hipMemcpyDeviceToHost, s));
    MCX_HIP(hipStreamSynchronize(s));
    *h_count = hh.count;
    if (hh.error) return MCX_ERR_INTERNAL;
    if (hh.count > cap) return MCX_ERR_CAPACITY;
  }
  return MCX_OK;
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
