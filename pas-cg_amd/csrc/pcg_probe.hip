// Measurement probe: a streaming copy of known size, the calibration kernel
// for HBM traffic counters and the repo's own achievable-bandwidth figure.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/pcg.h"

namespace {

// one 16-byte load and one 16-byte store per element, grid-stride
__global__ void k_stream_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, int64_t n) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

}  // namespace

extern "C" int pcg_stream_copy(int64_t bytes, int32_t iters, int32_t device, double *gbs,
                               double *ms_per_copy) {
  if (bytes < (1 << 20) || iters <= 0 || !gbs) return PCG_E_ARG;
  if (hipSetDevice(device) != hipSuccess) return PCG_E_HIP;
  int64_t n = bytes / 16;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * 16) != hipSuccess) return PCG_E_NOMEM;
  if (hipMalloc(&b, n * 16) != hipSuccess) {
    (void)hipFree(a);
    return PCG_E_NOMEM;
  }
  hipEvent_t e0, e1;
  int rc = PCG_OK;
  float ms = 0.f;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = PCG_E_HIP;
  if (rc == PCG_OK && hipMemset(a, 1, n * 16) != hipSuccess) rc = PCG_E_HIP;
  if (rc == PCG_OK) {
    dim3 grid(256 * 8 * 4), block(256);  // 4 workgroups of 256 per SIMD-quad on 256 CUs
    hipLaunchKernelGGL(k_stream_copy, grid, block, 0, 0, (const float4 *)a, (float4 *)b, n);
    (void)hipEventRecord(e0, 0);
    for (int k = 0; k < iters; ++k)
      hipLaunchKernelGGL(k_stream_copy, grid, block, 0, 0, (const float4 *)a, (float4 *)b, n);
    (void)hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = PCG_E_HIP;
  }
  if (rc == PCG_OK) {
    double per = ms / iters;
    *gbs = 2.0 * (double)n * 16.0 / (per * 1e-3) / 1e9;
    if (ms_per_copy) *ms_per_copy = per;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}
