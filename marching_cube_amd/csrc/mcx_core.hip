// mcx_core.hip -- library-level C ABI: version, status strings, tables,
// workspace initialisation and the stream-ordered fill kernels.
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <string.h>

#include "mc_tables.inc"
#include "mcx_internal.h"

namespace mcx {

namespace {
thread_local hipError_t t_last_hip = hipSuccess;

__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t* p, uint32_t v, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ __launch_bounds__(256) void fill_u128_kernel(uint4* p, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = make_uint4(0u, 0u, 0u, 0u);
}

__global__ void fill_bytes_kernel(uint8_t* p, size_t n) {
  for (size_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

unsigned grid_for(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b == 0) b = 1;
  return (unsigned)b;
}
}  // namespace

int hip_status(hipError_t e) {
  t_last_hip = e;
  return MCX_ERR_HIP;
}
void set_last_hip_error(hipError_t e) { t_last_hip = e; }

int fill_u32(void* d_ptr, uint32_t value, size_t n_words, hipStream_t s) {
  if (n_words == 0) return MCX_OK;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(n_words)), dim3(256), 0, s,
                     static_cast<uint32_t*>(d_ptr), value, n_words);
  MCX_HIP(hipGetLastError());
  return MCX_OK;
}

// Zero n_bytes: 16-byte body with 1-byte kernels for the unaligned head/tail.
int fill_bytes_zero(void* d_ptr, size_t n_bytes, hipStream_t s) {
  if (n_bytes == 0) return MCX_OK;
  uint8_t* p = static_cast<uint8_t*>(d_ptr);
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  size_t head = (16 - (a & 15)) & 15;
  if (head > n_bytes) head = n_bytes;
  if (head) {
    hipLaunchKernelGGL(fill_bytes_kernel, dim3(1), dim3(64), 0, s, p, head);
    MCX_HIP(hipGetLastError());
  }
  const size_t body = (n_bytes - head) / 16;
  if (body) {
    hipLaunchKernelGGL(fill_u128_kernel, dim3(grid_for(body)), dim3(256), 0, s,
                       reinterpret_cast<uint4*>(p + head), body);
    MCX_HIP(hipGetLastError());
  }
  const size_t tail = n_bytes - head - body * 16;
  if (tail) {
    hipLaunchKernelGGL(fill_bytes_kernel, dim3(1), dim3(64), 0, s, p + head + body * 16, tail);
    MCX_HIP(hipGetLastError());
  }
  return MCX_OK;
}

int take_error(void* d_ws, hipStream_t s) {
  int rc = fill_u32(static_cast<char*>(d_ws) + offsetof(WsHeader, error), 0u, 1, s);
  if (rc) return rc;
  MCX_HIP(hipStreamSynchronize(s));
  return MCX_ERR_INTERNAL;
}

}  // namespace mcx

using namespace mcx;

extern "C" int mcx_version(void) { return 10000; }

extern "C" const char* mcx_strerror(int status) {
  switch (status) {
    case MCX_OK: return "ok";
    case MCX_ERR_ARG: return "invalid argument";
    case MCX_ERR_CAPACITY: return "output capacity too small (count holds the required size)";
    case MCX_ERR_WORKSPACE: return "workspace too small or missing";
    case MCX_ERR_HIP: return t_last_hip != hipSuccess ? hipGetErrorString(t_last_hip) : "HIP error";
    case MCX_ERR_INTERNAL: return "device-side consistency check failed";
    case MCX_ERR_UNSUPPORTED: return "unsupported size or combination";
    case MCX_ERR_FORMAT: return "malformed input data";
    default: return "unknown status";
  }
}

extern "C" int mcx_get_tables(uint64_t* case_table, uint8_t* num_verts) {
  static const unsigned long long kCase[256] = MCX_CASE_TABLE_INIT;
  static const unsigned char kNv[256] = MCX_NUM_VERTS_INIT;
  if (case_table) memcpy(case_table, kCase, sizeof(kCase));
  if (num_verts) memcpy(num_verts, kNv, sizeof(kNv));
  return MCX_OK;
}

extern "C" int mcx_workspace_init(void* d_ws, size_t ws_bytes, void* stream) {
  if (!d_ws || ws_bytes < kWsHeader + 16) return MCX_ERR_WORKSPACE;
  return fill_bytes_zero(d_ws, ws_bytes, as_stream(stream));
}

extern "C" int mcx_workspace_status(void* d_ws, size_t ws_bytes, void* stream) {
  if (!d_ws || ws_bytes < kWsHeader) return MCX_ERR_WORKSPACE;
  hipStream_t s = as_stream(stream);
  uint32_t err = 0;
  MCX_HIP(hipMemcpyAsync(&err, static_cast<char*>(d_ws) + offsetof(WsHeader, error), sizeof(err),
                         hipMemcpyDeviceToHost, s));
  MCX_HIP(hipStreamSynchronize(s));
  return err ? take_error(d_ws, s) : MCX_OK;
}
