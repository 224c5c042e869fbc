// mcx_weld.hip -- indexed mesh from the canonical vertex list (gfx950).
//
// Replaces the reference's OBJ vertex de-duplication by printed coordinates
// ("%.4f_%.4f_%.4f" keys in a std::map, cu_kinect_fusion_v1.cu:3668-3719) with
// exact sharing by grid edge: every vertex of the extraction lies on one grid
// edge (its key: lower corner voxel * 3 + axis, written by the extraction), the
// up to 4 cubes around an edge each emit a vertex for it, and all of them
// become one vertex.  Its position is that of the first of them in the
// canonical order (the cubes interpolate the edge in different directions, so
// their fp32 positions can differ in the last bit).
//
// W1 insert  hash table (open addressing, power of two >= 2V slots): the
//            first lane of a wave holding a key (the smallest index of the
//            key in the wave; vertices come in cube order, so a key often
//            repeats within 64 vertices) claims the key's slot (atomicCAS)
//            and records its index with atomicMin, so the slot ends with the
//            first occurrence whatever the order of the atomics.
// W2 flag    each vertex looks its key up: rep = first occurrence; unique
//            vertices (rep == v) get flag 1; per-block flag sums.
// W3 scan    exclusive scan of the block sums (one workgroup).
// W4 place   unique index of every unique vertex; its position is copied;
//            then every vertex's triangle index = unique index of its rep.
// Deterministic: the table layout may vary between runs, what is read from
// it (the minimum index per key) does not.
#include <hip/hip_runtime.h>

#include "mcx_internal.h"

namespace mcx {
namespace {

// An empty slot holds kEmptyKey.  A vertex whose key equals it (-1 as int64:
// never a grid-edge key, but the entry point accepts any int64) is kept in
// the extra value slot T past the table, so that every key works.
constexpr uint64_t kEmptyKey = ~0ull;
constexpr int kWeldBlock = 1024;  // vertices per flag block (256 threads x 4)

__device__ __forceinline__ uint64_t mix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

__device__ __forceinline__ uint64_t find_slot(const uint64_t* keys, uint64_t mask, uint64_t key) {
  if (key == kEmptyKey) return mask + 1;  // its own slot past the table
  uint64_t h = mix64(key) & mask;
  while (keys[h] != key) h = (h + 1) & mask;  // the key is present (inserted by W1)
  return h;
}

__global__ __launch_bounds__(256) void weld_insert_kernel(const int64_t* __restrict__ ekey,
                                                          int64_t V, uint64_t* keys,
                                                          uint32_t* vals, uint64_t mask) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t key = v < V ? (uint64_t)ekey[v] : kEmptyKey;
  // a lower lane of the wave with the same key inserts for this one (lanes
  // past V are the highest of the last wave, so no live lane compares with
  // them)
  const int lane = (int)(threadIdx.x & 63);
  const uint32_t klo = (uint32_t)key, khi = (uint32_t)(key >> 32);
  bool dup = false;
  for (int i = 0; i < 63; ++i) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)klo, i);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)khi, i);
    dup |= i < lane && lo == klo && hi == khi;
  }
  if (v >= V || dup) return;
  if (key == kEmptyKey) {
    atomicMin(vals + mask + 1, (uint32_t)v);
    return;
  }
  uint64_t h = mix64(key) & mask;
  // at most V distinct keys in >= 2V slots: the probe ends
  for (;;) {
    const unsigned long long old =
        atomicCAS(reinterpret_cast<unsigned long long*>(keys + h), kEmptyKey, key);
    if (old == kEmptyKey || old == key) {
      atomicMin(vals + h, (uint32_t)v);
      return;
    }
    h = (h + 1) & mask;
  }
}

__global__ __launch_bounds__(256) void weld_flag_kernel(const int64_t* __restrict__ ekey,
                                                        int64_t V, const uint64_t* keys,
                                                        const uint32_t* vals, uint64_t mask,
                                                        uint32_t* rep, uint32_t* bsum) {
  __shared__ uint32_t s_w[4];
  const int64_t base = (int64_t)blockIdx.x * kWeldBlock;
  uint32_t cnt = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t v = base + k * 256 + threadIdx.x;
    if (v < V) {
      const uint32_t r = vals[find_slot(keys, mask, (uint64_t)ekey[v])];
      rep[v] = r;
      cnt += r == (uint32_t)v;
    }
  }
  uint32_t t = cnt;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) t += (uint32_t)__shfl_xor((int)t, o);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) bsum[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// Exclusive scan of nb block sums in place (one workgroup of 256 threads);
// bsum[nb] = total.
__global__ __launch_bounds__(256) void weld_scan_kernel(uint32_t* bsum, int64_t nb,
                                                        int64_t* d_nunique) {
  __shared__ uint32_t s_w[4];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int64_t b0 = 0; b0 < nb; b0 += 256) {
    const int64_t i = b0 + threadIdx.x;
    const uint32_t x = i < nb ? bsum[i] : 0u;
    const uint32_t inc = wave_incl_scan32(x);
    if ((threadIdx.x & 63) == 63) s_w[threadIdx.x >> 6] = inc;
    __syncthreads();
    uint32_t pre = s_carry;
    const int w = threadIdx.x >> 6;
    for (int k = 0; k < w; ++k) pre += s_w[k];
    if (i < nb) bsum[i] = pre + inc - x;
    __syncthreads();
    if (threadIdx.x == 255) s_carry = pre + inc;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    bsum[nb] = s_carry;
    if (d_nunique) *d_nunique = s_carry;
  }
}

// Unique index of each unique vertex (block offset + rank inside the block),
// its position copied to upos.  uidx doubles as the per-vertex unique index.
__global__ __launch_bounds__(256) void weld_place_kernel(int64_t V, const uint32_t* rep,
                                                         const uint32_t* bsum, const float* pos,
                                                         int64_t stride, uint32_t* uidx,
                                                         float* upos) {
  __shared__ uint32_t s_w[4];
  const int64_t base = (int64_t)blockIdx.x * kWeldBlock;
  const int tid = threadIdx.x;
  // thread owns vertices base + 4 tid .. +3 (consecutive: ranks in order)
  uint32_t f[4];
  uint32_t cnt = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t v = base + 4 * tid + k;
    f[k] = v < V && rep[v] == (uint32_t)v;
    cnt += f[k];
  }
  const uint32_t inc = wave_incl_scan32(cnt);
  if ((tid & 63) == 63) s_w[tid >> 6] = inc;
  __syncthreads();
  uint32_t pre = bsum[blockIdx.x];
  for (int k = 0; k < (tid >> 6); ++k) pre += s_w[k];
  uint32_t r = pre + inc - cnt;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t v = base + 4 * tid + k;
    if (f[k]) {
      uidx[v] = r;
      upos[3 * (int64_t)r + 0] = pos[v * stride + 0];
      upos[3 * (int64_t)r + 1] = pos[v * stride + 1];
      upos[3 * (int64_t)r + 2] = pos[v * stride + 2];
      ++r;
    }
  }
}

__global__ __launch_bounds__(256) void weld_index_kernel(int64_t V, const uint32_t* rep,
                                                         const uint32_t* uidx, int32_t* idx) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < V) {
    const uint32_t r = rep[v];  // the key's first occurrence: r <= v
    idx[v] = r < (uint64_t)V ? (int32_t)uidx[r] : -1;
  }
}

__global__ __launch_bounds__(256) void fill_u64_kernel(uint64_t* p, int64_t n, uint64_t val) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = val;
}

uint64_t table_slots(int64_t V) {
  uint64_t t = 1024;
  while (t < 2 * (uint64_t)V) t <<= 1;
  return t;
}

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

unsigned grid_for(int64_t n, int64_t per) { return (unsigned)((n + per - 1) / per); }

}  // namespace
}  // namespace mcx

using namespace mcx;

extern "C" size_t mcx_weld_workspace_size(int64_t V) {
  if (V < 0) return 0;
  const uint64_t T = table_slots(V);
  const int64_t nb = (V + kWeldBlock - 1) / kWeldBlock;
  return al256(T * 8) + al256((T + 1) * 4) + 3 * al256((size_t)V * 4 + 4) +
         al256((size_t)(nb + 1) * 4);
}

extern "C" int mcx_weld(const int64_t* d_edge_key, const float* d_pos, int64_t stride_floats,
                        int64_t V, float* d_upos, int32_t* d_indices, int64_t* d_nunique,
                        int64_t* h_nunique, void* d_ws, size_t ws_bytes, void* stream) {
  if (V < 0 || V >= ((int64_t)1 << 31) || stride_floats < 3) return MCX_ERR_ARG;
  hipStream_t s = as_stream(stream);
  if (V == 0) {
    if (d_nunique && fill_bytes_zero(d_nunique, sizeof(int64_t), s)) return MCX_ERR_HIP;
    if (h_nunique) {
      MCX_HIP(hipStreamSynchronize(s));
      *h_nunique = 0;
    }
    return MCX_OK;
  }
  if (!d_edge_key || !d_pos || !d_upos || !d_indices) return MCX_ERR_ARG;
  if (!d_ws || ws_bytes < mcx_weld_workspace_size(V)) return MCX_ERR_WORKSPACE;
  const uint64_t T = table_slots(V);
  const int64_t nb = (V + kWeldBlock - 1) / kWeldBlock;
  char* b = static_cast<char*>(d_ws);
  uint64_t* keys = reinterpret_cast<uint64_t*>(b);
  b += al256(T * 8);
  uint32_t* vals = reinterpret_cast<uint32_t*>(b);
  b += al256((T + 1) * 4);
  uint32_t* rep = reinterpret_cast<uint32_t*>(b);
  b += al256((size_t)V * 4 + 4);
  uint32_t* uidx = reinterpret_cast<uint32_t*>(b);
  b += al256((size_t)V * 4 + 4);
  b += al256((size_t)V * 4 + 4);  // reserved
  uint32_t* bsum = reinterpret_cast<uint32_t*>(b);
  int64_t fb = (int64_t)(T + 255) / 256;
  if (fb > 65536) fb = 65536;
  hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)fb), dim3(256), 0, s, keys, (int64_t)T,
                     kEmptyKey);
  MCX_HIP(hipGetLastError());
  int rc = fill_u32(vals, 0xffffffffu, (size_t)T + 1, s);
  if (rc) return rc;
  const uint64_t mask = T - 1;
  hipLaunchKernelGGL(weld_insert_kernel, dim3(grid_for(V, 256)), dim3(256), 0, s, d_edge_key, V,
                     keys, vals, mask);
  MCX_HIP(hipGetLastError());
  hipLaunchKernelGGL(weld_flag_kernel, dim3((unsigned)nb), dim3(256), 0, s, d_edge_key, V, keys,
                     vals, mask, rep, bsum);
  MCX_HIP(hipGetLastError());
  hipLaunchKernelGGL(weld_scan_kernel, dim3(1), dim3(256), 0, s, bsum, nb, d_nunique);
  MCX_HIP(hipGetLastError());
  hipLaunchKernelGGL(weld_place_kernel, dim3((unsigned)nb), dim3(256), 0, s, V, rep, bsum, d_pos,
                     stride_floats, uidx, d_upos);
  MCX_HIP(hipGetLastError());
  hipLaunchKernelGGL(weld_index_kernel, dim3(grid_for(V, 256)), dim3(256), 0, s, V, rep, uidx,
                     d_indices);
  MCX_HIP(hipGetLastError());
  if (h_nunique) {
    uint32_t total = 0;
    MCX_HIP(hipMemcpyAsync(&total, bsum + nb, sizeof(total), hipMemcpyDeviceToHost, s));
    MCX_HIP(hipStreamSynchronize(s));
    *h_nunique = total;
  }
  return MCX_OK;
}
