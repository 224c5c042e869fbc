// mcx_rle.hip -- the `.vol` run-length codec on the GPU (gfx950).
//
// Format (cu_kinect_fusion_v1.cu:1303-1447, for the 4-byte voxel
// cu_voxel_t {short tsdf; ushort w;} of structure.h:14-18): per z slice a
// uint32 byte count, then records u16 {n : 14, v : 2} covering the slice's
// voxels in [y][x] order: v = 1 a run of (+32766, w 0), v = 2 a run of
// (-32766, w 0), v = 3 (and 0 when decoding) a run of the 4-byte voxel that
// follows.  A run is a maximal stretch of equal tsdf (weights are not
// compared; the record keeps the run's first voxel), cut every 0x3fff voxels.
//
// Encode (flat over the slab's voxels, no sequential pass):
//   E1 count   tsdf change points per 1024 voxels (slice starts always);
//   scan       exclusive prefix of the counts;
//   E2 compact change-point positions in order;
//   E3 size    bytes of each change point's records (+4 for a slice start);
//   scan       byte offsets;
//   E4 write   records (and slice byte counts) at their offsets.
// Decode (records are 1 or 3 words, so a record boundary cannot be told
// from a word alone):
//   D0 index   each slice's payload range from the caller's offsets;
//   D1 resolve one workgroup per slice; every thread parses a 64-word
//              chunk from each of the three possible entry offsets (the
//              first record at or after the chunk start begins 0, 1 or 2
//              words in), giving a map entry offset -> (exit offset, voxels,
//              records); a block scan composes the maps, so each chunk
//              learns its true entry offset, voxel and record positions; it
//              re-parses and writes (first voxel, value) per record and, for
//              every 4096-voxel tile, the record covering the tile's start;
//   D2 fill    one workgroup per tile stages the tile's records in LDS and
//              writes the voxels (coalesced, balanced whatever the run
//              lengths).
// Malformed input (a run of length 0, runs past the slice, a truncated
// record, bytes left over, offsets outside the dump) is detected on the
// device and reported as MCX_ERR_FORMAT; every access stays in bounds.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mcx_internal.h"

namespace mcx {
namespace {

constexpr uint32_t kRunMax = 0x3fff;
constexpr int kMaxT = 32766;
constexpr int kBlk = 1024;        // scan block / encode count block (256 threads x 4)
constexpr int kChunkWords = 64;   // decode: words per speculative-parse chunk
constexpr int kFillTile = 4096;   // decode: voxels per fill workgroup
constexpr int kFillStage = kFillTile + 256;
constexpr uint64_t kSliceStart = 1ull << 63;
constexpr uint32_t kNoRecord = 0xffffffffu;

__device__ __forceinline__ int tsdf_of(uint32_t v) { return (int)(int16_t)(v & 0xffffu); }
__device__ __forceinline__ bool saturated(int t) { return t == kMaxT || t == -kMaxT; }

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// ------------------------------------------------------------ int64 scan
// Exclusive scan in place of a[0..n), a[n] = total (a holds n + 1).
__global__ __launch_bounds__(256) void scan_block_kernel(int64_t* a, int64_t n, int64_t* bsum) {
  __shared__ int64_t s_w[4];
  const int tid = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * kBlk + 4 * tid;
  int64_t v[4];
  int64_t t = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = base + k < n ? a[base + k] : 0;
    t += v[k];
  }
  const int64_t inc = (int64_t)wave_incl_scan((uint64_t)t);
  if ((tid & 63) == 63) s_w[tid >> 6] = inc;
  __syncthreads();
  int64_t pre = 0;
  for (int w = 0; w < (tid >> 6); ++w) pre += s_w[w];
  int64_t run = pre + inc - t;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) a[base + k] = run;
    run += v[k];
  }
  if (tid == 255) {
    if (bsum)
      bsum[blockIdx.x] = run;
    else
      a[n] = run;  // single block: the total
  }
}

__global__ __launch_bounds__(256) void scan_add_kernel(int64_t* a, int64_t n, const int64_t* boff,
                                                       int64_t nb) {
  const int64_t base = (int64_t)blockIdx.x * kBlk + 4 * threadIdx.x;
  const int64_t add = boff[blockIdx.x];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + k < n) a[base + k] += add;
  if (blockIdx.x == 0 && threadIdx.x == 0) a[n] = boff[nb];
}

int64_t scan_tmp_elems(int64_t n) {
  if (n <= kBlk) return 0;
  const int64_t nb = (n + kBlk - 1) / kBlk;
  return nb + 1 + scan_tmp_elems(nb);
}

int scan_excl(int64_t* a, int64_t n, int64_t* tmp, hipStream_t s) {
  if (n <= 0) return fill_bytes_zero(a, sizeof(int64_t), s);
  const int64_t nb = (n + kBlk - 1) / kBlk;
  if (nb == 1) {
    hipLaunchKernelGGL(scan_block_kernel, dim3(1), dim3(256), 0, s, a, n, (int64_t*)nullptr);
    MCX_HIP(hipGetLastError());
    return MCX_OK;
  }
  hipLaunchKernelGGL(scan_block_kernel, dim3((unsigned)nb), dim3(256), 0, s, a, n, tmp);
  MCX_HIP(hipGetLastError());
  int rc = scan_excl(tmp, nb, tmp + nb + 1, s);
  if (rc) return rc;
  hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)nb), dim3(256), 0, s, a, n, tmp, nb);
  MCX_HIP(hipGetLastError());
  return MCX_OK;
}

// ---------------------------------------------------------------- encode
typedef uint32_t u32x4a4 __attribute__((ext_vector_type(4), aligned(4)));  // 4-byte aligned

// grid (blocks per slice, nz): change points among 1024 voxels of slice z.
// Bit k: voxel i0 + k starts a run (the first of the slice, or a tsdf that
// differs from the voxel before it).
__device__ __forceinline__ uint32_t change_flags(const uint32_t* __restrict__ sl, int64_t XY,
                                                 int64_t i0) {
  uint32_t f = 0;
  if (i0 + 4 <= XY) {
    // the four voxels in one 128-bit load (the slice is 4-byte aligned,
    // which the vector memory path accepts), the one before in another
    const u32x4a4 v = *reinterpret_cast<const u32x4a4*>(sl + i0);
    int prev = i0 > 0 ? tsdf_of(sl[i0 - 1]) : 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int t = tsdf_of(v[k]);
      if (i0 + k == 0 || t != prev) f |= 1u << k;
      prev = t;
    }
    return f;
  }
  int prev = i0 > 0 && i0 - 1 < XY ? tsdf_of(sl[i0 - 1]) : 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t i = i0 + k;
    if (i < XY) {
      const int t = tsdf_of(sl[i]);
      if (i == 0 || t != prev) f |= 1u << k;
      prev = t;
    }
  }
  return f;
}

__global__ __launch_bounds__(256) void enc_count_kernel(const uint32_t* __restrict__ vol,
                                                        int64_t XY, int64_t* cnt) {
  __shared__ uint32_t s_w[4];
  const int tid = threadIdx.x;
  const int64_t z = blockIdx.y;
  const uint32_t f = change_flags(vol + z * XY, XY, (int64_t)blockIdx.x * kBlk + 4 * tid);
  uint32_t c = __popc(f);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += (uint32_t)__shfl_xor((int)c, o);
  if ((tid & 63) == 0) s_w[tid >> 6] = c;
  __syncthreads();
  if (tid == 0) cnt[z * gridDim.x + blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

__global__ __launch_bounds__(256) void enc_compact_kernel(const uint32_t* __restrict__ vol,
                                                          int64_t XY, const int64_t* ofs,
                                                          uint64_t* cp, int64_t* slice_first) {
  __shared__ uint32_t s_w[4];
  const int tid = threadIdx.x;
  const int64_t z = blockIdx.y;
  const int64_t i0 = (int64_t)blockIdx.x * kBlk + 4 * tid;
  const uint32_t f = change_flags(vol + z * XY, XY, i0);
  const uint32_t c = __popc(f);
  const uint32_t inc = wave_incl_scan32(c);
  if ((tid & 63) == 63) s_w[tid >> 6] = inc;
  __syncthreads();
  uint32_t pre = 0;
  for (int w = 0; w < (tid >> 6); ++w) pre += s_w[w];
  int64_t k = ofs[z * gridDim.x + blockIdx.x] + pre + inc - c;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (f & (1u << e)) {
      const int64_t i = i0 + e;
      cp[k] = (uint64_t)(z * XY + i) | (i == 0 ? kSliceStart : 0);
      if (i == 0) slice_first[z] = k;
      ++k;
    }
  }
}

__global__ __launch_bounds__(256) void enc_size_kernel(const uint32_t* __restrict__ vol,
                                                       const uint64_t* cp, int64_t M, int64_t N,
                                                       int64_t* bytes) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  const uint64_t c = cp[k];
  const int64_t s = (int64_t)(c & ~kSliceStart);
  const int64_t e = k + 1 < M ? (int64_t)(cp[k + 1] & ~kSliceStart) : N;
  const int64_t nrec = (e - s + kRunMax - 1) / kRunMax;
  const int64_t rs = saturated(tsdf_of(vol[s])) ? 2 : 6;
  bytes[k] = nrec * rs + ((c & kSliceStart) ? 4 : 0);
}

__device__ __forceinline__ void put16(uint16_t* out16, int64_t byte_pos, uint32_t v) {
  out16[byte_pos >> 1] = (uint16_t)v;
}

__global__ __launch_bounds__(256) void enc_write_kernel(const uint32_t* __restrict__ vol,
                                                        int64_t XY, const uint64_t* cp, int64_t M,
                                                        int64_t N, const int64_t* ofs,
                                                        const int64_t* slice_first, int nz,
                                                        uint16_t* out16, int64_t cap,
                                                        int64_t* d_total) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = ofs[M];
  if (k == 0 && d_total) *d_total = total;
  if (k >= M || total > cap) return;
  const uint64_t c = cp[k];
  const int64_t s = (int64_t)(c & ~kSliceStart);
  const int64_t e = k + 1 < M ? (int64_t)(cp[k + 1] & ~kSliceStart) : N;
  int64_t pos = ofs[k];
  if (c & kSliceStart) {
    const int64_t z = s / XY;
    const int64_t next = z + 1 < nz ? ofs[slice_first[z + 1]] : total;
    const uint32_t sz = (uint32_t)(next - pos - 4);
    put16(out16, pos, sz & 0xffffu);
    put16(out16, pos + 2, sz >> 16);
    pos += 4;
  }
  const int t = tsdf_of(vol[s]);
  const uint32_t v = t == kMaxT ? 1u : t == -kMaxT ? 2u : 3u;
  for (int64_t r = s; r < e; r += kRunMax) {
    const uint32_t n = (uint32_t)min((int64_t)kRunMax, e - r);
    put16(out16, pos, n | (v << 14));
    pos += 2;
    if (v == 3u) {
      const uint32_t val = vol[r];
      put16(out16, pos, val & 0xffffu);
      put16(out16, pos + 2, val >> 16);
      pos += 4;
    }
  }
}

// ---------------------------------------------------------------- decode
struct SliceInfo {
  int64_t word_base;  // first payload word within the dump
  uint32_t words;     // payload words
  uint32_t nrec;      // records (set by the resolve kernel)
};

// Map entry offset p (0..2) -> (exit offset, voxels, records) of a chunk.
struct Fn {
  uint32_t ex;  // 2 bits per entry offset
  uint32_t v[3];
  uint32_t r[3];
};

__device__ __forceinline__ uint32_t ex_at(uint32_t ex, uint32_t p) { return (ex >> (2 * p)) & 3u; }
__device__ __forceinline__ uint32_t sel3(const uint32_t* a, uint32_t p) {
  return p == 0 ? a[0] : p == 1 ? a[1] : a[2];
}
__device__ __forceinline__ uint32_t sat_add(uint32_t a, uint32_t b) {
  const uint32_t s = a + b;
  return s < a ? 0xffffffffu : s;
}

__device__ __forceinline__ Fn fn_identity() {
  Fn f;
  f.ex = 0u | (1u << 2) | (2u << 4);
#pragma unroll
  for (int p = 0; p < 3; ++p) f.v[p] = f.r[p] = 0;
  return f;
}

// a, then b
__device__ __forceinline__ Fn compose(const Fn& a, const Fn& b) {
  Fn o;
  o.ex = 0;
#pragma unroll
  for (uint32_t p = 0; p < 3; ++p) {
    const uint32_t e = ex_at(a.ex, p);
    o.ex |= ex_at(b.ex, e) << (2 * p);
    o.v[p] = sat_add(a.v[p], sel3(b.v, e));
    o.r[p] = a.r[p] + sel3(b.r, e);
  }
  return o;
}

__device__ __forceinline__ Fn shfl_up_fn(const Fn& f, int d) {
  Fn g;
  g.ex = __shfl_up(f.ex, d, kWave);
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    g.v[p] = __shfl_up(f.v[p], d, kWave);
    g.r[p] = __shfl_up(f.r[p], d, kWave);
  }
  return g;
}

__device__ __forceinline__ uint32_t record_words(uint32_t hdr) {
  const uint32_t v = hdr >> 14;
  return v == 1u || v == 2u ? 1u : 3u;
}

__global__ __launch_bounds__(256) void dec_index_kernel(const uint16_t* __restrict__ d16,
                                                        int64_t dump_bytes, const int64_t* offs,
                                                        int nz, SliceInfo* info, uint32_t* err) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nz) return;
  const int64_t off = offs[j];
  SliceInfo si{0, 0, 0};
  if (off >= 0 && (off & 1) == 0 && off + 4 <= dump_bytes) {
    const uint32_t size = (uint32_t)d16[off >> 1] | ((uint32_t)d16[(off >> 1) + 1] << 16);
    if ((size & 1u) == 0 && size >= 2 && (int64_t)size <= dump_bytes - off - 4) {
      si.word_base = (off + 4) >> 1;
      si.words = size >> 1;
    }
  }
  if (si.words == 0) atomicOr(err, 1u);
  info[j] = si;
}

__global__ __launch_bounds__(256) void dec_resolve_kernel(const uint16_t* __restrict__ d16,
                                                          int64_t XY, int tps, SliceInfo* info,
                                                          uint32_t* rec_start, uint32_t* rec_val,
                                                          uint32_t* tile_first, uint32_t* err) {
  __shared__ Fn s_fn[4];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int j = blockIdx.x;
  const SliceInfo si = info[j];
  if (si.words == 0) return;  // flagged by the index kernel
  const uint16_t* __restrict__ w = d16 + si.word_base;
  const int64_t L = si.words;
  const int64_t nch = (L + kChunkWords - 1) / kChunkWords;
  uint32_t* rs = rec_start + si.word_base;
  uint32_t* rv = rec_val + si.word_base;
  uint32_t* tf = tile_first + (int64_t)j * tps;
  uint32_t cp = 0, cv = 0, cr = 0;  // carry: entry offset, voxels, records
  bool bad = false;
  for (int64_t c0 = 0; c0 < nch; c0 += 256) {
    const int64_t c = c0 + tid;
    const int64_t beg = c * kChunkWords;
    const int64_t end = min(beg + kChunkWords, L);
    Fn f = fn_identity();
    if (c < nch) {
      f.ex = 0;
#pragma unroll 1
      for (uint32_t p = 0; p < 3; ++p) {
        int64_t pos = beg + p;
        uint32_t vox = 0, rec = 0;
        while (pos < end) {
          const uint32_t h = w[pos];
          vox = sat_add(vox, h & kRunMax);
          ++rec;
          pos += record_words(h);
        }
        f.ex |= (uint32_t)(pos - end) << (2 * p);
        f.v[p] = vox;
        f.r[p] = rec;
      }
    }
    // inclusive scan of the maps over the block
    Fn incl = f;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const Fn g = shfl_up_fn(incl, d);
      if (lane >= d) incl = compose(g, incl);
    }
    if (lane == 63) s_fn[wv] = incl;
    __syncthreads();
    Fn pre = fn_identity();
    for (int k = 0; k < wv; ++k) pre = compose(pre, s_fn[k]);
    const Fn up = shfl_up_fn(incl, 1);
    const Fn excl = lane == 0 ? pre : compose(pre, up);
    Fn total = s_fn[0];
    for (int k = 1; k < 4; ++k) total = compose(total, s_fn[k]);
    __syncthreads();  // s_fn is rewritten by the next batch
    if (c < nch) {
      // true entry state of this chunk, then its records
      int64_t pos = beg + ex_at(excl.ex, cp);
      uint64_t v0 = (uint64_t)sat_add(cv, sel3(excl.v, cp));
      uint32_t r = cr + sel3(excl.r, cp);
      while (pos < end) {
        const uint32_t h = w[pos];
        const uint32_t n = h & kRunMax, v = h >> 14;
        const uint32_t len = record_words(h);
        if (n == 0 || pos + len > L || v0 + n > (uint64_t)XY) {
          bad = true;
          break;
        }
        uint32_t val;
        if (v == 1u)
          val = (uint32_t)(uint16_t)kMaxT;
        else if (v == 2u)
          val = (uint32_t)(uint16_t)(int16_t)-kMaxT;
        else
          val = (uint32_t)w[pos + 1] | ((uint32_t)w[pos + 2] << 16);
        rs[r] = (uint32_t)v0;
        rv[r] = val;
        for (uint64_t t = (v0 + kFillTile - 1) / kFillTile; t * kFillTile < v0 + n; ++t)
          tf[t] = r;
        v0 += n;
        ++r;
        pos += len;
      }
    }
    const uint32_t np = ex_at(total.ex, cp);
    cv = sat_add(cv, sel3(total.v, cp));
    cr += sel3(total.r, cp);
    cp = np;
  }
  if (bad) atomicOr(err, 2u);
  if (tid == 0) {
    if (cp != 0 || (int64_t)cv != XY) atomicOr(err, 4u);
    info[j].nrec = cr;
  }
}

__global__ __launch_bounds__(256) void dec_fill_kernel(const SliceInfo* info, int64_t XY, int tps,
                                                       const uint32_t* rec_start,
                                                       const uint32_t* rec_val,
                                                       const uint32_t* tile_first,
                                                       const uint32_t* err, uint32_t* out,
                                                       int vec) {
  // staged runs: start - vb clamped to [0, kFillTile] (16 bits, so six
  // workgroups fit a CU's LDS; for every voxel q >= vb of the tile,
  // start <= q exactly when the clamped value <= q - vb), and voxel value
  __shared__ uint16_t s_start[kFillStage];
  __shared__ uint32_t s_val[kFillStage];
  static_assert(kFillTile <= 0xffff, "tile offsets in 16 bits");
  if (*err) return;  // malformed input: the call reports it, outputs are undefined
  const int tid = threadIdx.x;
  const int j = blockIdx.y, t = blockIdx.x;
  const SliceInfo si = info[j];
  const uint32_t r0 = tile_first[(int64_t)j * tps + t];
  if (r0 == kNoRecord || r0 >= si.nrec) return;
  const uint32_t* rs = rec_start + si.word_base;
  const uint32_t* rv = rec_val + si.word_base;
  const int64_t vb = (int64_t)t * kFillTile;
  const int64_t ve = min(vb + kFillTile, XY);
  int m = 0;
  for (int round = 0; round * 256 < kFillTile + 1; ++round) {
    const int64_t ri = (int64_t)r0 + round * 256 + tid;
    bool in = false;
    if (ri < si.nrec) {
      const uint32_t s = rs[ri];
      in = ri == r0 || (int64_t)s < ve;
      if (in) {
        s_start[round * 256 + tid] = (uint16_t)min(max((int64_t)s - vb, (int64_t)0), (int64_t)kFillTile);
        s_val[round * 256 + tid] = rv[ri];
      }
    }
    const int got = __syncthreads_count(in);
    m += got;
    if (got < 256) break;
  }
  __syncthreads();
  uint32_t* o = out + (int64_t)j * XY;
  for (int64_t q = vb + 4 * tid; q < ve; q += 4 * 256) {
    const int qr = (int)(q - vb);
    int lo = 0, hi = m - 1;  // last staged record with start <= q
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if ((int)s_start[mid] <= qr)
        lo = mid;
      else
        hi = mid - 1;
    }
    uint32_t v4[4];
    int idx = lo;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      while (idx + 1 < m && (int)s_start[idx + 1] <= qr + e) ++idx;
      v4[e] = s_val[idx];
    }
    if (vec && q + 3 < ve) {
      *reinterpret_cast<uint4*>(o + q) = make_uint4(v4[0], v4[1], v4[2], v4[3]);
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (q + e < ve) o[q + e] = v4[e];
    }
  }
}

unsigned grid_for(int64_t n, int64_t per) { return (unsigned)((n + per - 1) / per); }

}  // namespace
}  // namespace mcx

using namespace mcx;

// ------------------------------------------------------------------ C ABI
extern "C" size_t mcx_rle_encode_workspace_size(int X, int Y, int nz) {
  if (X < 1 || Y < 1 || nz < 1) return 0;
  const int64_t XY = (int64_t)X * Y, N = XY * nz;
  const int64_t nblk = ((XY + kBlk - 1) / kBlk) * nz;
  const int64_t tmp = std::max(scan_tmp_elems(nblk), scan_tmp_elems(N));
  return al256((size_t)(nblk + 1) * 8) + al256((size_t)N * 8) + al256((size_t)(N + 1) * 8) +
         al256((size_t)nz * 8) + al256((size_t)(tmp + 1) * 8) + 256;
}

extern "C" size_t mcx_rle_encode_bound(int X, int Y, int nz) {
  if (X < 1 || Y < 1 || nz < 1) return 0;
  return (size_t)nz * (4 + 6 * (size_t)X * Y);
}

extern "C" int mcx_rle_encode(const void* d_vol, int X, int Y, int Z, int z0, int nz, void* d_dump,
                              size_t dump_cap, int64_t* d_dump_bytes, int64_t* h_dump_bytes,
                              void* d_ws, size_t ws_bytes, void* stream) {
  if (!d_vol || X < 1 || Y < 1 || Z < 1 || z0 < 0 || nz < 1 || (int64_t)z0 + nz > Z)
    return MCX_ERR_ARG;
  if ((int64_t)X * Y >= ((int64_t)1 << 31)) return MCX_ERR_ARG;
  if (dump_cap > 0 && (!d_dump || (reinterpret_cast<uintptr_t>(d_dump) & 1))) return MCX_ERR_ARG;
  if (!d_ws || ws_bytes < mcx_rle_encode_workspace_size(X, Y, nz)) return MCX_ERR_WORKSPACE;
  hipStream_t s = as_stream(stream);
  const int64_t XY = (int64_t)X * Y, N = XY * nz;
  const int64_t bps = (XY + kBlk - 1) / kBlk, nblk = bps * nz;
  if (bps > 0x7fffffff || nz > 65535) return MCX_ERR_ARG;
  const uint32_t* vol = static_cast<const uint32_t*>(d_vol) + (int64_t)z0 * XY;
  char* b = static_cast<char*>(d_ws);
  int64_t* cnt = reinterpret_cast<int64_t*>(b);
  b += al256((size_t)(nblk + 1) * 8);
  uint64_t* cp = reinterpret_cast<uint64_t*>(b);
  b += al256((size_t)N * 8);
  int64_t* bytes = reinterpret_cast<int64_t*>(b);
  b += al256((size_t)(N + 1) * 8);
  int64_t* slice_first = reinterpret_cast<int64_t*>(b);
  b += al256((size_t)nz * 8);
  int64_t* tmp = reinterpret_cast<int64_t*>(b);
  const dim3 g2((unsigned)bps, (unsigned)nz);
  hipLaunchKernelGGL(enc_count_kernel, g2, dim3(256), 0, s, vol, XY, cnt);
  MCX_HIP(hipGetLastError());
  int rc = scan_excl(cnt, nblk, tmp, s);
  if (rc) return rc;
  int64_t M = 0;
  MCX_HIP(hipMemcpyAsync(&M, cnt + nblk, sizeof(M), hipMemcpyDeviceToHost, s));
  MCX_HIP(hipStreamSynchronize(s));
  if (M < nz || M > N) return MCX_ERR_INTERNAL;
  // one thread per run below: the grid (in work-items) must fit 32 bits
  if (M >= ((int64_t)1 << 31)) return MCX_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(enc_compact_kernel, g2, dim3(256), 0, s, vol, XY, cnt, cp, slice_first);
  MCX_HIP(hipGetLastError());
  hipLaunchKernelGGL(enc_size_kernel, dim3(grid_for(M, 256)), dim3(256), 0, s, vol, cp, M, N,
                     bytes);
  MCX_HIP(hipGetLastError());
  rc = scan_excl(bytes, M, tmp, s);
  if (rc) return rc;
  hipLaunchKernelGGL(enc_write_kernel, dim3(grid_for(M, 256)), dim3(256), 0, s, vol, XY, cp, M, N,
                     bytes, slice_first, nz, static_cast<uint16_t*>(d_dump), (int64_t)dump_cap,
                     d_dump_bytes);
  MCX_HIP(hipGetLastError());
  int64_t total = 0;
  MCX_HIP(hipMemcpyAsync(&total, bytes + M, sizeof(total), hipMemcpyDeviceToHost, s));
  MCX_HIP(hipStreamSynchronize(s));
  if (h_dump_bytes) *h_dump_bytes = total;
  return total > (int64_t)dump_cap ? MCX_ERR_CAPACITY : MCX_OK;
}

extern "C" size_t mcx_rle_decode_workspace_size(int64_t dump_bytes, int X, int Y, int nz) {
  if (dump_bytes < 0 || X < 1 || Y < 1 || nz < 1) return 0;
  const int64_t XY = (int64_t)X * Y;
  const int64_t tps = (XY + kFillTile - 1) / kFillTile;
  const size_t words = (size_t)(dump_bytes / 2) + 1;
  return 256 + al256((size_t)nz * sizeof(SliceInfo)) + al256((size_t)nz * 8) +
         al256((size_t)nz * tps * 4) + 2 * al256(words * 4);
}

extern "C" int mcx_rle_decode(const void* d_dump, int64_t dump_bytes,
                              const int64_t* h_slice_offsets, int nz, int X, int Y, void* d_vol,
                              void* d_ws, size_t ws_bytes, void* stream) {
  if (!d_dump || dump_bytes < 6 || !h_slice_offsets || nz < 1 || X < 1 || Y < 1 || !d_vol)
    return MCX_ERR_ARG;
  if (reinterpret_cast<uintptr_t>(d_dump) & 1) return MCX_ERR_ARG;
  const int64_t XY = (int64_t)X * Y;
  if (XY >= ((int64_t)1 << 31) || nz > 65535) return MCX_ERR_ARG;
  if (!d_ws || ws_bytes < mcx_rle_decode_workspace_size(dump_bytes, X, Y, nz))
    return MCX_ERR_WORKSPACE;
  hipStream_t s = as_stream(stream);
  const int64_t tps = (XY + kFillTile - 1) / kFillTile;
  const size_t words = (size_t)(dump_bytes / 2) + 1;
  char* b = static_cast<char*>(d_ws);
  uint32_t* err = reinterpret_cast<uint32_t*>(b);
  b += 256;
  SliceInfo* info = reinterpret_cast<SliceInfo*>(b);
  b += al256((size_t)nz * sizeof(SliceInfo));
  int64_t* offs = reinterpret_cast<int64_t*>(b);
  b += al256((size_t)nz * 8);
  uint32_t* tile_first = reinterpret_cast<uint32_t*>(b);
  b += al256((size_t)nz * tps * 4);
  uint32_t* rec_start = reinterpret_cast<uint32_t*>(b);
  b += al256(words * 4);
  uint32_t* rec_val = reinterpret_cast<uint32_t*>(b);
  MCX_HIP(hipMemcpyAsync(offs, h_slice_offsets, (size_t)nz * 8, hipMemcpyHostToDevice, s));
  int rc = fill_u32(err, 0u, 1, s);
  if (!rc) rc = fill_u32(tile_first, kNoRecord, (size_t)nz * tps, s);
  if (rc) return rc;
  const uint16_t* d16 = static_cast<const uint16_t*>(d_dump);
  hipLaunchKernelGGL(dec_index_kernel, dim3(grid_for(nz, 256)), dim3(256), 0, s, d16, dump_bytes,
                     offs, nz, info, err);
  MCX_HIP(hipGetLastError());
  hipLaunchKernelGGL(dec_resolve_kernel, dim3((unsigned)nz), dim3(256), 0, s, d16, XY, (int)tps,
                     info, rec_start, rec_val, tile_first, err);
  MCX_HIP(hipGetLastError());
  const int vec = (XY % 4 == 0) && ((reinterpret_cast<uintptr_t>(d_vol) & 15) == 0);
  hipLaunchKernelGGL(dec_fill_kernel, dim3((unsigned)tps, (unsigned)nz), dim3(256), 0, s, info, XY,
                     (int)tps, rec_start, rec_val, tile_first, err, static_cast<uint32_t*>(d_vol),
                     vec);
  MCX_HIP(hipGetLastError());
  uint32_t herr = 0;
  MCX_HIP(hipMemcpyAsync(&herr, err, sizeof(herr), hipMemcpyDeviceToHost, s));
  MCX_HIP(hipStreamSynchronize(s));
  return herr ? MCX_ERR_FORMAT : MCX_OK;
}
