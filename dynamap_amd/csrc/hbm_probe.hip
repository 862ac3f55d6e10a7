// HBM traffic reference for the counter scale: a plain fill (writes n bytes)
// and a plain copy (reads n bytes, writes n bytes) of a buffer of n bytes,
// each launched three times.  Run under rocprofv3 --pmc FETCH_SIZE /
// WRITE_SIZE beside a kernel of interest: the counters of these kernels
// against their known byte counts give the factor that converts the
// counters to bytes.  Usage: hbm_probe [MiB]  (default 1024)
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

__global__ __launch_bounds__(256) void k_probe_fill(float4 *__restrict__ dst, size_t n4) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256)
    dst[i] = make_float4(1.f, 2.f, 3.f, 4.f);
}

__global__ __launch_bounds__(256) void k_probe_copy(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                   size_t n4) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

#define CK(x)                                                                   \
  do {                                                                          \
    hipError_t e = (x);                                                         \
    if (e != hipSuccess) {                                                      \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e));                    \
      return 1;                                                                 \
    }                                                                           \
  } while (0)

int main(int argc, char **argv) {
  const size_t mib = argc > 1 ? (size_t)atol(argv[1]) : 1024;
  if (mib == 0 || mib > 65536) { fprintf(stderr, "size must be 1..65536 MiB\n"); return 2; }
  const size_t bytes = mib << 20, n4 = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  CK(hipMalloc(&a, bytes));
  CK(hipMalloc(&b, bytes));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  const unsigned grid = 256 * 32;
  for (int rep = 0; rep < 3; ++rep) {
    float ms_fill = 0.f, ms_copy = 0.f;
    CK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(k_probe_fill, dim3(grid), dim3(256), 0, 0, a, n4);
    CK(hipEventRecord(e1, 0));
    CK(hipEventSynchronize(e1));
    CK(hipEventElapsedTime(&ms_fill, e0, e1));
    CK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(k_probe_copy, dim3(grid), dim3(256), 0, 0, a, b, n4);
    CK(hipEventRecord(e1, 0));
    CK(hipEventSynchronize(e1));
    CK(hipEventElapsedTime(&ms_copy, e0, e1));
    printf("bytes %zu fill %.3f ms (%.1f GB/s) copy %.3f ms (%.1f GB/s read+write)\n", bytes, ms_fill,
           bytes / ms_fill / 1e6, ms_copy, 2.0 * bytes / ms_copy / 1e6);
  }
  CK(hipFree(a));
  CK(hipFree(b));
  return 0;
}
