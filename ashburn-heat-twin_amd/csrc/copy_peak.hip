// Measured HBM copy peak and the calibration reference for rocprofv3 FETCH_SIZE / WRITE_SIZE counters.
// Copies `bytes` (default 2 GiB) from one device buffer to another with 16-byte loads and stores, `reps` times,
// and prints the event-timed bandwidth (read + write bytes per second) as one JSON line.
// Usage: aht_copy_peak [bytes] [reps]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void copy_f4(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) b[t] = a[t];
}

#define CK(x)                                                                     \
  do {                                                                            \
    hipError_t e = (x);                                                           \
    if (e != hipSuccess) {                                                        \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e));                      \
      return 1;                                                                   \
    }                                                                             \
  } while (0)

int main(int argc, char **argv) {
  const size_t bytes = argc > 1 ? strtoull(argv[1], nullptr, 10) : (size_t)2 << 30;
  const int reps = argc > 2 ? atoi(argv[2]) : 20;
  const size_t n = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  CK(hipMalloc(&a, n * 16));
  CK(hipMalloc(&b, n * 16));
  CK(hipMemset(a, 0, n * 16));
  CK(hipMemset(b, 0, n * 16));
  int cus = 0;
  CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  const int blocks = cus * 32;
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  hipLaunchKernelGGL(copy_f4, dim3(blocks), dim3(256), 0, 0, a, b, n);
  CK(hipDeviceSynchronize());
  CK(hipEventRecord(e0, 0));
  for (int r = 0; r < reps; r++) hipLaunchKernelGGL(copy_f4, dim3(blocks), dim3(256), 0, 0, a, b, n);
  CK(hipEventRecord(e1, 0));
  CK(hipEventSynchronize(e1));
  float ms = 0;
  CK(hipEventElapsedTime(&ms, e0, e1));
  const double per = ms / reps;
  printf("{\"kernel\": \"copy_f4\", \"bytes_copied\": %zu, \"reps\": %d, \"ms_per_launch\": %.5f, "
         "\"GBps_read_plus_write\": %.1f, \"cus\": %d}\n",
         n * 16, reps, per, 2.0 * n * 16 / (per * 1e-3) / 1e9, cus);
  CK(hipFree(a));
  CK(hipFree(b));
  return 0;
}
