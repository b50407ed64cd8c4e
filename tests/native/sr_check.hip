// Lane-by-lane check of gfx950's stochastic-rounding conversions against the models storage.h states for them
// (f32_to_f16_sr / f32_to_bf16_sr call these instructions in the step's stores):
//   v_cvt_sr_bf16_f32 x, r : (bits(x) + (r >> 16)) >> 16, for every finite x;
//   v_cvt_sr_f16_f32  x, r : bits(x) + (r >> 19), then truncation to binary16, for x in the binary16 normal range;
//                            unbiased in the subnormal range (mean error over many seeds ~ 0).
// 4 M random values (magnitudes 2^-30 .. 2^13, random seeds) and 256 subnormal values x 4,096 seeds. Prints the counts
// and exits non-zero on any mismatch or a subnormal bias above 0.05 units of 2^-24.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef __bf16 b2 __attribute__((ext_vector_type(2)));

__global__ void convert(const float *x, const uint32_t *r, uint16_t *o16, uint16_t *ob, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const h2 v = __builtin_amdgcn_cvt_sr_f16_f32(__builtin_bit_cast(h2, x[i]), x[i], r[i], false);
  o16[i] = __builtin_bit_cast(uint16_t, (_Float16)v[0]);
  const b2 w = __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(b2, x[i]), x[i], r[i], false);
  ob[i] = __builtin_bit_cast(uint16_t, (__bf16)w[0]);
}

static uint32_t bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static float from_bits(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
static double h2d(uint16_t h) {   // binary16 -> value
  const int s = h >> 15, e = (h >> 10) & 31, m = h & 1023;
  const double v = e == 0 ? std::ldexp((double)m, -24) : (e == 31 ? INFINITY : std::ldexp((double)(m | 1024), e - 25));
  return s ? -v : v;
}
static uint16_t f2h_rtz(float y) {   // truncation toward zero, normal binary16 range
  const uint32_t u = bits(y);
  const uint16_t s = (u >> 16) & 0x8000;
  const int e = (int)((u >> 23) & 255) - 127 + 15;
  return s | (uint16_t)(e << 10) | (uint16_t)((u >> 13) & 1023);
}

#define CK(x) do { if ((x) != hipSuccess) { std::printf("HIP error at line %d\n", __LINE__); return 2; } } while (0)

int main() {
  const int n = 1 << 22, nsub = 1 << 20, off = n - nsub;
  std::vector<float> x(n);
  std::vector<uint32_t> r(n);
  std::mt19937 g(11);
  for (int i = 0; i < n; i++) {
    const int e = (int)(g() % 44) - 30;
    const float m = 1.0f + (float)(g() & 0x7fffff) / 8388608.0f;
    x[i] = std::ldexp(m, e) * ((g() & 1) ? -1.0f : 1.0f);
    r[i] = g();
  }
  for (int i = 0; i < nsub; i++) x[off + i] = std::ldexp(1.0f + (float)((i >> 12) % 256) / 256.0f, -20 - (i >> 12) / 64);
  float *dx; uint32_t *dr; uint16_t *o16, *ob;
  CK(hipMalloc(&dx, n * 4)); CK(hipMalloc(&dr, n * 4)); CK(hipMalloc(&o16, n * 2)); CK(hipMalloc(&ob, n * 2));
  CK(hipMemcpy(dx, x.data(), n * 4, hipMemcpyHostToDevice));
  CK(hipMemcpy(dr, r.data(), n * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(convert, dim3(n / 256), dim3(256), 0, 0, dx, dr, o16, ob, n);
  CK(hipGetLastError());
  std::vector<uint16_t> h16(n), hb(n);
  CK(hipMemcpy(h16.data(), o16, n * 2, hipMemcpyDeviceToHost));
  CK(hipMemcpy(hb.data(), ob, n * 2, hipMemcpyDeviceToHost));
  long nb = 0, bad_b = 0, nh = 0, bad_h = 0;
  for (int i = 0; i < off; i++) {
    nb++;
    if ((uint16_t)((bits(x[i]) + (r[i] >> 16)) >> 16) != hb[i]) bad_b++;
    const float ax = std::fabs(x[i]);
    if (ax >= 0x1p-14f && ax < 65504.0f) {
      const float y = from_bits(bits(x[i]) + (r[i] >> 19));
      if (std::fabs(y) < 65520.0f) {
        nh++;
        if (f2h_rtz(y) != h16[i]) bad_h++;
      }
    }
  }
  double worst = 0;
  for (int gi = 0; gi < nsub / 4096; gi++) {
    double s = 0;
    for (int j = 0; j < 4096; j++) {
      const int i = off + gi * 4096 + j;
      s += (h2d(h16[i]) - (double)x[i]) / 0x1p-24;
    }
    s /= 4096;
    if (std::fabs(s) > std::fabs(worst)) worst = s;
  }
  std::printf("{\"bf16_checked\": %ld, \"bf16_mismatch\": %ld, \"f16_normal_checked\": %ld, \"f16_normal_mismatch\": %ld, "
              "\"f16_subnormal_worst_mean_error_units\": %.4f}\n", nb, bad_b, nh, bad_h, worst);
  return (bad_b || bad_h || std::fabs(worst) > 0.05) ? 1 : 0;
}
