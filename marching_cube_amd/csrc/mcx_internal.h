// mcx_internal.h -- shared device helpers and host plumbing for libmcx.
//
// gfx950 only (wave64).  Nothing here keeps device globals that hold state:
// look-back descriptors, tickets and counters live in caller workspace memory.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mcx.h"

namespace mcx {

constexpr int kWave = 64;

// ------------------------------------------------------- exact fp32 forms
// s / 32766 correctly rounded for every 16-bit s: a multiply by the rounded
// reciprocal and one fused correction step (equal to IEEE division on all
// 65536 values; tests/cpp/exact_div_check.hip checks it on the device).
__device__ __forceinline__ float tsdf_unit(int s) {
  constexpr float r = 1.0f / 32766.0f;
  const float x = (float)s;
  const float q = x * r;
  return __fmaf_rn(__fmaf_rn(-q, 32766.0f, x), r, q);
}

// s / 32767 correctly rounded for every 16-bit s (the stored tsdf of the
// packed integration as a fraction), in the same form as tsdf_unit (equal to
// IEEE division on all 65536 values; checked by the same device test).
__device__ __forceinline__ float tsdf_fraction(int s) {
  constexpr float r = 1.0f / 32767.0f;
  const float x = (float)s;
  const float q = x * r;
  return __fmaf_rn(__fmaf_rn(-q, 32767.0f, x), r, q);
}

// num / den correctly rounded, for the marching-cubes interpolation weight
// of two short2 corners of opposite classification (num = 0 - fa, den = fb -
// fa + 1e-15f, fa = tsdf_unit(ka), fb = tsdf_unit(kb), ka < 0 <= kb or
// kb < 0 <= ka): a reciprocal refined once and one quotient correction.
// Equal to IEEE division on all 2^31 such pairs (tests/cpp/exact_div_check.hip
// checks every pair on the device); fp32 volumes keep IEEE division.
__device__ __forceinline__ float div_weight_short2(float num, float den) {
  float r = __builtin_amdgcn_rcpf(den);
  r = __fmaf_rn(__fmaf_rn(-den, r, 1.0f), r, r);
  const float q = num * r;
  return __fmaf_rn(__fmaf_rn(-den, q, num), r, q);
}

// ---------------------------------------------------------------- host side
// Map a HIP error to a status code; keep the first HIP error for strerror.
int hip_status(hipError_t e);
void set_last_hip_error(hipError_t e);

#define MCX_HIP(expr)                               \
  do {                                              \
    hipError_t _e = (expr);                         \
    if (_e != hipSuccess) return ::mcx::hip_status(_e); \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Stream-ordered byte fill (a kernel, so it is safe inside captured graphs).
int fill_u32(void* d_ptr, uint32_t value, size_t n_words, hipStream_t s);
int fill_bytes_zero(void* d_ptr, size_t n_bytes, hipStream_t s);

// A workspace's error word was found set: clear it (stream-ordered fill,
// then a synchronisation) and return MCX_ERR_INTERNAL, so that the error is
// reported once.
int take_error(void* d_ws, hipStream_t s);

// ------------------------------------------------------ workspace layout
// [0, 256)   header: u32 dense_acc, u32 dense_last, u32 scan_ticket, u32 error, i64 count, ...
// [256, ...) u64 scan look-back descriptors (zeroed by each call's count kernel)
constexpr size_t kWsHeader = 256;
// The device array behind mcx_extract_static (bytes per GPU) and the fewest
// records it must leave room for beside the tile tables.
#define MCX_STATIC_SCRATCH_BYTES (64ull << 20)
constexpr int64_t kStaticMinRecords = 65536;
struct WsHeader {
  uint32_t dense_acc;   // tiles of the running call that took the staged path
  uint32_t dense_last;  // the same count for the last completed call
  uint32_t scan_ticket; // scan workgroups started by the running call (their order)
  uint32_t error;       // a look-back wait that hit its limit since the last report
  int64_t count;        // vertices of the last call; -1 when its scan reported an error
  uint32_t nz_count;    // non-empty tiles appended by the running count kernel
  uint32_t nz_total;    // the same, frozen by the scan for the emit kernel
  uint64_t ovf_next;    // records taken from the shared overflow pool
  uint32_t dense_n;     // staged-path tiles listed by the scan for the emit kernel
  uint32_t list_n;      // the list's tiles with records, compacted by the scan
};
static_assert(sizeof(WsHeader) <= 256, "header fits its slot");

// Look-back descriptor: status in the top two bits, value below.
constexpr uint64_t kStatAgg = 1ull << 62;
constexpr uint64_t kStatInc = 2ull << 62;
constexpr uint64_t kValMask = (1ull << 62) - 1;
constexpr int kSpinLimit = 1 << 20;

// ------------------------------------------------- division by invariants
// n / d for 32-bit unsigned n via one mulhi (Granlund-Montgomery).
struct FastDiv {
  uint32_t d, m;
  int s;  // shift
  FastDiv() = default;
  __host__ __device__ explicit FastDiv(uint32_t div) {
    d = div;
    if (div <= 1) {
      m = 0;
      s = 0;
      return;
    }
    int l = 0;
    while ((1ull << l) < div) ++l;
    m = (uint32_t)(((1ull << 32) * ((1ull << l) - div)) / div + 1);
    s = l;
  }
  __host__ __device__ __forceinline__ uint32_t div(uint32_t n) const {
    if (d <= 1) return n;
    uint32_t t = (uint32_t)(((uint64_t)m * n) >> 32);
    return (t + ((n - t) >> 1)) >> (s - 1);
  }
};

// ----------------------------------------------------------- wave helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t shfl_up_u64(uint64_t v, int d) {
  uint32_t lo = __shfl_up((uint32_t)v, d, kWave);
  uint32_t hi = __shfl_up((uint32_t)(v >> 32), d, kWave);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
  uint32_t lo = __shfl_xor((uint32_t)v, m, kWave);
  uint32_t hi = __shfl_xor((uint32_t)(v >> 32), m, kWave);
  return ((uint64_t)hi << 32) | lo;
}


// DPP controls (gfx9 wave64): row_shr:n shifts within rows of 16 lanes,
// row_bcast:15 / :31 pass lane 15 of each row / lane 31 on to the next
// row(s); a lane without a source takes `old` (0 below).
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114,
              kDppRowShr8 = 0x118, kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;

// Inclusive prefix sum over the 64 lanes of a wave in registers: four row
// shifts give each row of 16 its own scan, then lane 15's total goes to
// rows 1 and 3 and lane 31's to rows 2 and 3.
__device__ __forceinline__ uint32_t wave_incl_scan32(uint32_t v) {
  v += __builtin_amdgcn_update_dpp(0u, v, kDppRowShr1, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, kDppRowShr2, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, kDppRowShr4, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, kDppRowShr8, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, kDppRowBcast15, 0xa, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, kDppRowBcast31, 0xc, 0xf, false);
  return v;
}

// 64-bit inclusive prefix sum over the lanes of a wave (the same DPP
// pattern on both halves, a 64-bit add per step).
template <int CTRL, int ROWS>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_update_dpp(0u, (uint32_t)v, CTRL, ROWS, 0xf, false);
  const uint32_t hi = __builtin_amdgcn_update_dpp(0u, (uint32_t)(v >> 32), CTRL, ROWS, 0xf, false);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v) {
  v += dpp_u64<kDppRowShr1, 0xf>(v);
  v += dpp_u64<kDppRowShr2, 0xf>(v);
  v += dpp_u64<kDppRowShr4, 0xf>(v);
  v += dpp_u64<kDppRowShr8, 0xf>(v);
  v += dpp_u64<kDppRowBcast15, 0xa>(v);
  v += dpp_u64<kDppRowBcast31, 0xc>(v);
  return v;
}

// Inclusive maximum over the lanes of a wave (same pattern).
__device__ __forceinline__ uint32_t wave_incl_max32(uint32_t v) {
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0u, v, kDppRowShr1, 0xf, 0xf, false));
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0u, v, kDppRowShr2, 0xf, 0xf, false));
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0u, v, kDppRowShr4, 0xf, 0xf, false));
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0u, v, kDppRowShr8, 0xf, 0xf, false));
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0u, v, kDppRowBcast15, 0xa, 0xf, false));
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0u, v, kDppRowBcast31, 0xc, 0xf, false));
  return v;
}

// Sum over the 64 lanes, the same in every lane.
__device__ __forceinline__ uint32_t wave_sum32(uint32_t v) {
  return __builtin_amdgcn_readlane(wave_incl_scan32(v), 63);
}
__device__ __forceinline__ uint64_t wave_sum(uint64_t v) {
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) v += shfl_xor_u64(v, m);
  return v;
}

// Orders a wave's own LDS writes before its later LDS reads of other lanes'
// data (wave-private staging areas; no workgroup barrier needed).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t ld_agent(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Decoupled look-back executed by one full wave for tile t (t > 0): returns
// the sum of all tiles before t.  Lane l inspects descriptors
// j - l*W - w (w < W), so one round trip to memory covers 64*W tiles and the
// chain of inclusive prefixes advances faster than tiles are issued.
// Tiles are numbered by a ticket each workgroup takes when it starts (an
// atomic counter), so every predecessor of a waiting tile is already running,
// and every tile publishes its aggregate before it looks back itself: every
// wait ends.  A wait beyond kSpinLimit (a broken invariant, never expected)
// sets *err and counts the tile as inclusive 0, so the wave still exits; the
// caller then publishes -1 as its vertex count and the error stays set in the
// workspace until reported (mcx_workspace_status or a synchronous call).
template <int W>
__device__ __forceinline__ int64_t lookback(const uint64_t* desc, int64_t t, uint32_t* err) {
  const int lane = lane_id();
  int64_t acc = 0;
  int64_t j = t - 1;
  while (true) {
    uint64_t d[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const int64_t idx = j - (int64_t)lane * W - w;
      d[w] = idx >= 0 ? ld_agent(desc + idx) : kStatInc;  // virtual inclusive 0 before tile 0
    }
#pragma unroll
    for (int w = 0; w < W; ++w) {
      int spins = 0;
      while ((d[w] >> 62) == 0) {
        if (++spins >= kSpinLimit) {
          atomicOr(err, 1u);
          d[w] = kStatInc;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        d[w] = ld_agent(desc + (j - (int64_t)lane * W - w));
      }
    }
    int first = W;  // nearest inclusive descriptor of this lane
#pragma unroll
    for (int w = W - 1; w >= 0; --w)
      if ((d[w] >> 62) == 2) first = w;
    const uint64_t incmask = __ballot(first < W);
    const int fl = incmask ? __ffsll((unsigned long long)incmask) - 1 : kWave;
    uint64_t v = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const bool take = lane < fl || (lane == fl && w <= first);
      v += take ? (d[w] & kValMask) : 0;
    }
    acc += (int64_t)wave_sum(v);
    if (incmask) break;
    j -= (int64_t)kWave * W;
  }
  return acc;
}

// ------------------------------------------------------------ batch plans
// Per-volume table of a batch plan (device memory), shared by the batched
// extraction, MC backward and integration kernels.  Per-voxel arrays of a
// batch (fp32 TSDF, weights, gradients) are concatenated in volume order:
// volume b's voxels start at vox_begin.  Workgroups of the per-voxel kernels
// take 256 consecutive voxels of one volume (blk_begin: the volume's first
// such workgroup); the batched gather takes 1024 voxels of one plane
// (gblk_begin).
struct BatchVol {
  int X, Y, Z;
  float h;
  int64_t vox_begin;
  int64_t blk_begin;
  int64_t gblk_begin;
  FastDiv divX, divY;
};

struct VolDesc;  // extraction descriptors (mcx_extract.hip)
struct Column;

// Volume of workgroup `blk` (the last b with begin[b] <= blk), by binary
// A dispatch carries its grid size in work-items as a 32-bit number, so a
// grid of one workgroup per 256 or more items (volumes of 2^32 voxels and
// more) is issued as several launches of at most this many workgroups, each
// told the index of its first workgroup.
constexpr int64_t kMaxLaunchBlocks = (int64_t)1 << 23;

// search over a BatchVol table; the member selects which begin.
template <int64_t BatchVol::*BEGIN>
__device__ __forceinline__ int batch_volume_of(const BatchVol* bv, int nvol, int64_t blk) {
  int lo = 0, hi = nvol - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bv[mid].*BEGIN <= blk) lo = mid; else hi = mid - 1;
  }
  return lo;
}

}  // namespace mcx

// A batch of volumes (opaque in the C ABI): built once by mcx_batch_create.
struct mcx_batch {
  int nvol = 0;
  int elem = MCX_ELEM_SHORT2;
  int maxX = 0;
  bool wide = true;  // every volume suits the 128-bit count kernel
  int slab = 64;     // planes per column
  int64_t n_tiles = 0;
  int64_t ncols = 0;
  int64_t cubes = 0;
  int64_t voxels = 0;       // all volumes' voxels
  int64_t blocks = 0;       // 256-voxel workgroups of the per-voxel kernels
  int64_t gather_blocks = 0;
  mcx::VolDesc* d_vols = nullptr;
  mcx::Column* d_cols = nullptr;
  mcx::BatchVol* d_bvol = nullptr;
};
