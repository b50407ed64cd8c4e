// Measured HBM copy peak and the calibration reference for rocprofv3 FETCH_SIZE / WRITE_SIZE counters.
// Copies `bytes` (default 4 GiB) between two device buffers with 16-byte loads and stores in several launch shapes,
// `reps` launches each, and prints one JSON line per variant with the event-timed read+write bandwidth, then one
// line with the best (the measured copy peak).
// Usage: aht_copy_peak [bytes] [reps] [variant]   (variant: all | grid | flat | unroll4 | unroll4_xcd | unroll4_nt)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

// grid-stride loop over the buffer, 32 blocks per CU
__global__ __launch_bounds__(256) void copy_grid(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) b[t] = a[t];
}

// one 16-byte element per thread
__global__ __launch_bounds__(256) void copy_flat(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = a[t];
}

typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t swz(uint32_t blk, uint32_t G) {
  const uint32_t q = G / 8, r = G % 8, x = blk % 8, i = blk / 8;
  return x * q + (x < r ? x : r) + i;
}

// four elements per thread (a block moves 16 KiB), optionally XCD-aware block order and non-temporal stores
template <bool XCD, bool NT>
__global__ __launch_bounds__(256) void copy_unroll4(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n) {
  const uint32_t blk = XCD ? swz(blockIdx.x, gridDim.x) : blockIdx.x;
  const size_t base = (size_t)blk * 1024 + threadIdx.x;
  float4 v[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const size_t t = base + (size_t)u * 256;
    if (t < n) v[u] = a[t];
  }
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const size_t t = base + (size_t)u * 256;
    if (t < n) {
      if (NT) __builtin_nontemporal_store(*(const f4v *)&v[u], (f4v *)&b[t]);
      else b[t] = v[u];
    }
  }
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
  const size_t bytes = argc > 1 ? strtoull(argv[1], nullptr, 10) : (size_t)4 << 30;
  const int reps = argc > 2 ? atoi(argv[2]) : 20;
  const char *which = argc > 3 ? argv[3] : "all";
  const size_t n = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  CK(hipMalloc(&a, n * 16));
  CK(hipMalloc(&b, n * 16));
  CK(hipMemset(a, 0, n * 16));
  CK(hipMemset(b, 0, n * 16));
  int cus = 0;
  CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  const char *names[5] = {"grid", "flat", "unroll4", "unroll4_xcd", "unroll4_nt"};
  double best = 0;
  const char *best_name = "";
  for (int v = 0; v < 5; v++) {
    if (strcmp(which, "all") && strcmp(which, names[v])) continue;
    auto launch = [&]() {
      switch (v) {
        case 0: hipLaunchKernelGGL(copy_grid, dim3(cus * 32), dim3(256), 0, 0, a, b, n); break;
        case 1: hipLaunchKernelGGL(copy_flat, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, a, b, n); break;
        case 2: hipLaunchKernelGGL((copy_unroll4<false, false>), dim3((unsigned)((n + 1023) / 1024)), dim3(256), 0, 0, a, b, n); break;
        case 3: hipLaunchKernelGGL((copy_unroll4<true, false>), dim3((unsigned)((n + 1023) / 1024)), dim3(256), 0, 0, a, b, n); break;
        case 4: hipLaunchKernelGGL((copy_unroll4<true, true>), dim3((unsigned)((n + 1023) / 1024)), dim3(256), 0, 0, a, b, n); break;
      }
    };
    launch();
    CK(hipDeviceSynchronize());
    CK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; r++) launch();
    CK(hipEventRecord(e1, 0));
    CK(hipEventSynchronize(e1));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, e0, e1));
    const double per = ms / reps, gbs = 2.0 * n * 16 / (per * 1e-3) / 1e9;
    printf("{\"kernel\": \"copy_%s\", \"bytes_copied\": %zu, \"reps\": %d, \"ms_per_launch\": %.5f, "
           "\"GBps_read_plus_write\": %.1f, \"cus\": %d}\n", names[v], n * 16, reps, per, gbs, cus);
    if (gbs > best) { best = gbs; best_name = names[v]; }
  }
  printf("{\"copy_peak_GBps\": %.1f, \"variant\": \"%s\"}\n", best, best_name);
  CK(hipFree(a));
  CK(hipFree(b));
  return 0;
}
