// Population storage formats. The arithmetic type `real` is what the kernels compute in; `elem` is what sits in HBM.
//
//   F32  : f and g stored raw in f32 (the reference's layout: read_state/write_state exchange it bit for bit).
//   F16  : f stored as (f_q - w_q) * 2^10, g as g_q * 2^10, IEEE binary16; f32 arithmetic. Storing the deviation
//          from the rest weight keeps the mantissa for the part that changes; the power-of-two scale (exact) keeps
//          deviations down to ~6e-8 out of the subnormal range (binary16 normals start at 6.1e-5) and leaves
//          |f - w| <= 64 and |g_q| <= 64 representable.
//   BF16 : the same deviations in bfloat16 (round to nearest even), f32 arithmetic.
//   F64  : f and g raw in f64 with f64 arithmetic: the FP64 twin used as the reference of reduced-precision errors.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// cache-policy operand of the population buffer loads and stores (gfx950 encoding: 1 = sc0, 2 = nt, 16 = sc1);
// a build option, 0 = default policy
#ifndef AHT_LD_CPOL
#define AHT_LD_CPOL 0
#endif
#ifndef AHT_ST_CPOL
#define AHT_ST_CPOL 0
#endif

namespace aht {

constexpr float DEV_SCALE = 1024.0f;
constexpr float DEV_INV = 1.0f / 1024.0f;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, -1, 0x00020000);
}

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16(float x) {
  const uint32_t u = __float_as_uint(x);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);   // round to nearest even (finite inputs)
}

struct StoreF32 {
  using elem = float;
  using real = float;
  static constexpr int bytes = 4;
  static constexpr int id = 0;
  __device__ static real dec_f(elem e, float) { return e; }
  __device__ static elem enc_f(real v, float) { return v; }
  __device__ static real dec_g(elem e) { return e; }
  __device__ static elem enc_g(real v) { return v; }
  __device__ static elem ld_raw(const void *base, uint32_t boff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsrc(base), boff, 0, AHT_LD_CPOL));
  }
  __device__ static void st_raw(void *base, uint32_t boff, elem v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), rsrc(base), boff, 0, AHT_ST_CPOL);
  }
};

struct StoreF16 {
  using elem = _Float16;
  using real = float;
  static constexpr int bytes = 2;
  static constexpr int id = 1;
  __device__ static real dec_f(elem e, float w) { return (float)e * DEV_INV + w; }
  __device__ static elem enc_f(real v, float w) { return (elem)((v - w) * DEV_SCALE); }
  __device__ static real dec_g(elem e) { return (float)e * DEV_INV; }
  __device__ static elem enc_g(real v) { return (elem)(v * DEV_SCALE); }
  __device__ static elem ld_raw(const void *base, uint32_t boff) {
    return __builtin_bit_cast(_Float16, (uint16_t)__builtin_amdgcn_raw_buffer_load_b16(rsrc(base), boff, 0, AHT_LD_CPOL));
  }
  __device__ static void st_raw(void *base, uint32_t boff, elem v) {
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, v), rsrc(base), boff, 0, AHT_ST_CPOL);
  }
};

struct bf16_t {
  uint16_t bits;
};

struct StoreBF16 {
  using elem = bf16_t;
  using real = float;
  static constexpr int bytes = 2;
  static constexpr int id = 2;
  __device__ static real dec_f(elem e, float w) { return bf16_to_f32(e.bits) * DEV_INV + w; }
  __device__ static elem enc_f(real v, float w) { return bf16_t{f32_to_bf16((v - w) * DEV_SCALE)}; }
  __device__ static real dec_g(elem e) { return bf16_to_f32(e.bits) * DEV_INV; }
  __device__ static elem enc_g(real v) { return bf16_t{f32_to_bf16(v * DEV_SCALE)}; }
  __device__ static elem ld_raw(const void *base, uint32_t boff) {
    return bf16_t{(uint16_t)__builtin_amdgcn_raw_buffer_load_b16(rsrc(base), boff, 0, AHT_LD_CPOL)};
  }
  __device__ static void st_raw(void *base, uint32_t boff, elem v) {
    __builtin_amdgcn_raw_buffer_store_b16(v.bits, rsrc(base), boff, 0, AHT_ST_CPOL);
  }
};

struct StoreF64 {
  using elem = double;
  using real = double;
  static constexpr int bytes = 8;
  static constexpr int id = 3;
  __device__ static real dec_f(elem e, float) { return e; }
  __device__ static elem enc_f(real v, float) { return v; }
  __device__ static real dec_g(elem e) { return e; }
  __device__ static elem enc_g(real v) { return v; }
  __device__ static elem ld_raw(const void *base, uint32_t boff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc(base), boff, 0, AHT_LD_CPOL));
  }
  __device__ static void st_raw(void *base, uint32_t boff, elem v) {
    using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(rsrc(base), boff, 0, AHT_LD_CPOL));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, v), rsrc(base), boff, 0, AHT_ST_CPOL);
  }
};

}  // namespace aht
