// Calibration probe for HBM traffic counters: copies a buffer of known size
// with a plain kernel, so that FETCH_SIZE / WRITE_SIZE read for the (T)
// kernels can be scaled by what the same counters read for exactly N bytes.
// Usage: copy_probe [MiB]   (default 256)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy_kernel(const double4 *__restrict__ src,
                            double4 *__restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int main(int argc, char **argv) {
  const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 256;
  const size_t bytes = mib << 20, n = bytes / sizeof(double4);
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    std::fprintf(stderr, "allocation failed\n");
    return 1;
  }
  (void)hipMemset(a, 1, bytes);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  float best = 1e30f;
  for (int rep = 0; rep < 3; rep++) {
    (void)hipEventRecord(e0, 0);
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, 0,
                       (const double4 *)a, (double4 *)b, n);
    (void)hipEventRecord(e1, 0);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms < best) best = ms;
  }
  std::printf("copy_probe: %zu bytes read, %zu bytes written per dispatch, "
              "best %.3f ms (%.1f GB/s read+write)\n",
              bytes, bytes, best, 2.0 * bytes / best / 1e6);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
