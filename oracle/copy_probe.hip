// copy_probe.hip -- a plain device-to-device copy of known size, launched 5
// times.  Profiling the memory counters (FETCH_SIZE, WRITE_SIZE) of this
// program gives the factor between what those counters report and the bytes
// actually moved (1 GiB read and 1 GiB written per launch), which scales the
// same counters read on the force kernels.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void k_copy(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

__global__ void k_fill(float4 *__restrict__ a, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    a[i] = make_float4(1.0f, 2.0f, 3.0f, 4.0f);
}

int main() {
  const size_t bytes = 1ull << 30;
  const size_t n = bytes / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc((void **)&a, bytes) != hipSuccess || hipMalloc((void **)&b, bytes) != hipSuccess) {
    std::fprintf(stderr, "copy_probe: allocation failed\n");
    return 1;
  }
  hipLaunchKernelGGL(k_fill, dim3(8192), dim3(256), 0, 0, a, n);
  for (int r = 0; r < 5; r++) hipLaunchKernelGGL(k_copy, dim3(8192), dim3(256), 0, 0, a, b, n);
  if (hipDeviceSynchronize() != hipSuccess) {
    std::fprintf(stderr, "copy_probe: kernel failed\n");
    return 1;
  }
  std::printf("copy_probe: 5 copies of %zu bytes\n", bytes);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
