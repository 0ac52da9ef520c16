// Device microbenchmarks used to calibrate roofline figures: a streaming copy of known size
// (HBM bandwidth, and the byte count against which profiler memory counters are scaled) and an
// FP32 fused-multiply-add chain (achievable vector FP32 rate).
#include <hip/hip_runtime.h>

#include "../../include/stormm_amd.h"

namespace {

__global__ void __launch_bounds__(256) k_stream_copy(const float4 *__restrict__ src,
                                                     float4 *__restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    dst[i] = src[i];
  }
}

// Each thread runs 8 independent FMA chains of `iters` steps; the result is stored so the
// compiler keeps the arithmetic.
__global__ void __launch_bounds__(256) k_fma_chain(float *out, int iters, float a, float b) {
  float v[8];
  for (int k = 0; k < 8; k++) {
    v[k] = static_cast<float>(threadIdx.x + k) * 1.0e-3f;
  }
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
      v[k] = __builtin_fmaf(v[k], a, b);
    }
  }
  float s = 0.0f;
  for (int k = 0; k < 8; k++) {
    s += v[k];
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

extern "C" int stormm_copy_bandwidth(int64_t bytes, int32_t repeats, double *gbps) {
  if (bytes < 1024 || bytes > (int64_t(8) << 30) || repeats <= 0 || gbps == nullptr) {
    return STORMM_ERR_INVALID_ARGUMENT;
  }
  const size_t n = static_cast<size_t>(bytes) / sizeof(float4);
  float4 *src = nullptr, *dst = nullptr;
  if (hipMalloc(&src, n * sizeof(float4)) != hipSuccess) {
    return STORMM_ERR_HIP;
  }
  if (hipMalloc(&dst, n * sizeof(float4)) != hipSuccess) {
    (void)hipFree(src);
    return STORMM_ERR_HIP;
  }
  (void)hipMemset(src, 0, n * sizeof(float4));
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const int grid = 256 * 16;
  hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, 0, src, dst, n);
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < repeats; r++) {
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, 0, src, dst, n);
  }
  (void)hipEventRecord(e1, 0);
  (void)hipEventSynchronize(e1);
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  *gbps = (2.0 * static_cast<double>(n * sizeof(float4)) * repeats) / (ms * 1.0e-3) / 1.0e9;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(src);
  (void)hipFree(dst);
  return STORMM_OK;
}

extern "C" int stormm_fma_throughput(int32_t iters, double *tflops) {
  if (iters <= 0 || iters > (1 << 22) || tflops == nullptr) {
    return STORMM_ERR_INVALID_ARGUMENT;
  }
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    return STORMM_ERR_HIP;
  }
  const int grid = prop.multiProcessorCount * 32;
  float *out = nullptr;
  if (hipMalloc(&out, static_cast<size_t>(grid) * 256 * sizeof(float)) != hipSuccess) {
    return STORMM_ERR_HIP;
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipLaunchKernelGGL(k_fma_chain, dim3(grid), dim3(256), 0, 0, out, 16, 0.999f, 1.0e-4f);
  (void)hipEventRecord(e0, 0);
  hipLaunchKernelGGL(k_fma_chain, dim3(grid), dim3(256), 0, 0, out, iters, 0.999f, 1.0e-4f);
  (void)hipEventRecord(e1, 0);
  (void)hipEventSynchronize(e1);
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  const double flops = 2.0 * 8.0 * static_cast<double>(iters) * grid * 256.0;
  *tflops = flops / (ms * 1.0e-3) / 1.0e12;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(out);
  return STORMM_OK;
}
