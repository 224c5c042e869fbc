// Device check of the exact fp32 forms in mcx_internal.h (built by
// marching_cube_amd/build.py, run by tests/test_exact_div_gpu.py):
//  * tsdf_unit(k) == (float)k / 32766.0f (IEEE division) for all 65536 k;
//  * tsdf_fraction(k) == (float)k / 32767.0f for all 65536 k (the packed
//    integration's stored tsdf);
//  * div_weight_short2(0 - fa, fb - fa + 1e-15f) == the IEEE quotient for all
//    2^31 pairs of opposite classification (ka < 0 <= kb and kb < 0 <= ka),
//    the operands of the marching-cubes interpolation weight
//    (cu_marching_cubes_v1.cu:259) for short2 volumes.
// Compiled with the library's flags (-fhip-fp32-correctly-rounded-divide-sqrt,
// -ffp-contract=off), so `/` is the IEEE division the reference computes.
// Prints "unit_mismatch U weight_mismatch W pairs P fraction_mismatch F";
// exits 0 when run.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "mcx_internal.h"

namespace {

__global__ void check_unit(unsigned long long* cnt) {
  const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x) - 32768;
  if (k > 32767) return;
  const float ieee = (float)k / 32766.0f;
  if (__float_as_uint(mcx::tsdf_unit(k)) != __float_as_uint(ieee)) atomicAdd(&cnt[0], 1ull);
  const float ieee7 = (float)k / 32767.0f;
  if (__float_as_uint(mcx::tsdf_fraction(k)) != __float_as_uint(ieee7)) atomicAdd(&cnt[3], 1ull);
}

__global__ void check_weight(unsigned long long* cnt) {
  const unsigned long long n_all = 1ull << 31;
  for (unsigned long long t = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
       t < n_all; t += (unsigned long long)gridDim.x * blockDim.x) {
    const int order = (int)(t >> 30);
    const int a = (int)((t >> 15) & 0x7fff), b = (int)(t & 0x7fff);
    const int kneg = -1 - a, kpos = b;
    const int ka = order ? kpos : kneg, kb = order ? kneg : kpos;
    const float fa = mcx::tsdf_unit(ka), fb = mcx::tsdf_unit(kb);
    const float num = 0.0f - fa;
    const float den = fb - fa + 1e-15f;
    const float ieee = num / den;
    if (__float_as_uint(mcx::div_weight_short2(num, den)) != __float_as_uint(ieee))
      atomicAdd(&cnt[1], 1ull);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&cnt[2], n_all);
}

}  // namespace

int main() {
  unsigned long long* d = nullptr;
  if (hipMalloc(&d, 4 * sizeof(unsigned long long)) != hipSuccess) return 2;
  if (hipMemset(d, 0, 4 * sizeof(unsigned long long)) != hipSuccess) return 2;
  hipLaunchKernelGGL(check_unit, dim3(256), dim3(256), 0, 0, d);
  hipLaunchKernelGGL(check_weight, dim3(8192), dim3(256), 0, 0, d);
  if (hipDeviceSynchronize() != hipSuccess) return 3;
  unsigned long long h[4] = {};
  if (hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return 3;
  std::printf("unit_mismatch %llu weight_mismatch %llu pairs %llu fraction_mismatch %llu\n", h[0],
              h[1], h[2], h[3]);
  return hipFree(d) == hipSuccess ? 0 : 4;
}
