// hbm_copy: calibration for rocprofv3's FETCH_SIZE / WRITE_SIZE counters.
// A plain copy of N floats reads exactly N*4 bytes and writes N*4 bytes per
// launch; comparing the counters' per-dispatch values with those sizes gives
// the factors that turn counter readings into HBM bytes.
#include <hip/hip_runtime.h>
#include <cstdio>
__global__ void copy_f32(const float* __restrict__ a, float* __restrict__ b, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}
int main() {
  const size_t n = 64ull << 20;  // 256 MiB per array
  float *a, *b;
  if (hipMalloc(&a, n * 4) != hipSuccess || hipMalloc(&b, n * 4) != hipSuccess) return 1;
  if (hipMemset(a, 0, n * 4) != hipSuccess || hipMemset(b, 0, n * 4) != hipSuccess) return 1;
  for (int it = 0; it < 5; it++) hipLaunchKernelGGL(copy_f32, dim3(8192), dim3(256), 0, 0, a, b, n);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
  (void)hipEventRecord(e0);
  for (int it = 0; it < 10; it++) hipLaunchKernelGGL(copy_f32, dim3(8192), dim3(256), 0, 0, a, b, n);
  (void)hipEventRecord(e1);
  float ms = 0.f;
  if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
  printf("copy %zu MiB x2 per launch: %.1f GB/s\n", (n * 4) >> 20, 10 * 2.0 * n * 4 / (ms * 1e-3) / 1e9);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
