// Population storage formats. The arithmetic type `real` is what the kernels compute in; `elem` is what sits in HBM.
//
//   F32  : f and g stored raw in f32 (the reference's layout: read_state/write_state exchange it bit for bit).
//   F16  : f stored as (f_q - w_q) * 2^10 (deviation from the rest weight, rho_0 = 1) in IEEE binary16; the D3Q7
//          lattice as the cell's theta in f32 plus the deviations (g_q - w_q theta) * 2^10 in binary16, i.e. theta_0
//          taken per cell. Keeping theta itself in f32 keeps the slowly growing heat content out of binary16 rounding
//          (stored as g_q * 2^10, the rounding of g alone biased the 10-minute mean warming by 1.5e-3 at 1200 x 1200
//          x 96, while the f deviations contributed 8e-5). f32 arithmetic. The power-of-two scale (exact) keeps
//          deviations down to ~6e-8 out of the subnormal range and leaves |deviation| <= 64 representable.
//   BF16 : the same layout with bfloat16 deviations, f32 arithmetic, stochastic rounding in the step (below).
//   F16S : f as F16, the D3Q7 lattice as g_q * 2^10 in binary16 with no theta plane (104 B per update instead of
//          112), both with stochastic rounding in the step.
//   F64  : f and g raw in f64 with f64 arithmetic: the FP64 twin used as the reference of reduced-precision errors.
//
// Stochastic rounding (SR = true): the step's stores round each value up or down to a neighbouring 16-bit value with
// probabilities given by its position between them, from random bits that depend only on the cell, the population
// and the step index (so a cell gives the same bits in any launch shape and in a batch). Unbiased rounding keeps the
// slowly growing heat content and the small per-step increments of the populations from stalling at a fixed
// 16-bit value, which round-to-nearest does: it is what limits bf16 (8-bit mantissa) and g stored without its theta
// reference. Initial states, re-initialised cells and written states round to nearest.
//
// D3Q7 lattice of a theta-referenced format (theta_ref): [7][N] deviations, then (at byte offset theta_offset(N),
// 256-aligned) [N] f32 theta = sum_q g_q of the stored post-collision populations.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Population loads and stores go through raw buffer descriptors whose base is the population plane with the pull
// offset folded in (built from SGPRs per wavefront), so the only per-lane address is the cell's byte offset.
// Cache-policy operand of those loads and stores (gfx950 encoding: 1 = sc0, 2 = nt, 16 = sc1), a build option. Stores
// are non-temporal by default: the next lattice is written once and read one step later by other workgroups, so
// keeping it out of the way of the pulled lines in L2 is worth ~6 % of the step time (measured at 600 x 600 x 64:
// 0.970 -> 0.908 ms f32, 0.548 -> 0.511 ms f16). Non-temporal loads are slower (the pulls share lines).
#ifndef AHT_LD_CPOL
#define AHT_LD_CPOL 0
#endif
#ifndef AHT_ST_CPOL
#define AHT_ST_CPOL 2
#endif

namespace aht {

constexpr float DEV_SCALE = 1024.0f;
constexpr float DEV_INV = 1.0f / 1024.0f;
constexpr int PAIR_QB = 26;   // populations per block of a blocked pair: 19 D3Q19, then 7 D3Q7

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, -1, 0x00020000);
}

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16(float x) {
  const uint32_t u = __float_as_uint(x);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);   // round to nearest even (finite inputs)
}

// Stochastic rounding with gfx950's conversion instructions, which take the random bits as a 32-bit operand:
// v_cvt_sr_bf16_f32 adds the top 16 bits of r to the magnitude bits of x and truncates to bfloat16; v_cvt_sr_f16_f32
// adds the top 13 bits of r below the binary16 mantissa and truncates, rounding stochastically into the binary16
// subnormal range too, and gives infinity beyond the binary16 range (both checked lane by lane against these models
// on the device). The bit patterns of one sign are ordered like their values, so the carry rounds up with
// probability equal to the dropped fraction. x itself is passed as the other half of the result vector (unused), so
// no register is cleared for it.
typedef _Float16 f16x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint16_t f32_to_bf16_sr(float x, uint32_t r) {
  const bf16x2_t v = __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(bf16x2_t, x), x, r, false);
  return __builtin_bit_cast(uint16_t, (__bf16)v[0]);
}
__device__ __forceinline__ _Float16 f32_to_f16_sr(float x, uint32_t r) {
  return __builtin_amdgcn_cvt_sr_f16_f32(__builtin_bit_cast(f16x2_t, x), x, r, false)[0];
}

struct StoreF32 {
  using elem = float;
  using real = float;
  static constexpr int bytes = 4;
  static constexpr int id = 0;
  static constexpr int LB = 0;              // population layout: 0 = [q][cell]; see lat_index
  static constexpr bool theta_ref = false;
  static constexpr bool SR = false;         // stochastic rounding in the step's stores
  __device__ static elem enc_f_sr(real v, float w, uint32_t) { return v; }
  __device__ static elem enc_g_sr(real v, float, real, uint32_t) { return v; }
  __device__ static real dec_f(elem e, float) { return e; }
  __device__ static elem enc_f(real v, float) { return v; }
  __device__ static real dec_g(elem e, float, real) { return e; }
  __device__ static elem enc_g(real v, float, real) { return v; }
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
  static constexpr int LB = 0;
  static constexpr bool theta_ref = true;
  static constexpr bool SR = false;
  __device__ static real dec_f(elem e, float w) { return (float)e * DEV_INV + w; }
  __device__ static elem enc_f(real v, float w) { return (elem)((v - w) * DEV_SCALE); }
  __device__ static real dec_g(elem e, float w, real th) { return (float)e * DEV_INV + w * th; }
  __device__ static elem enc_g(real v, float w, real th) { return (elem)((v - w * th) * DEV_SCALE); }
  __device__ static elem enc_f_sr(real v, float w, uint32_t) { return enc_f(v, w); }
  __device__ static elem enc_g_sr(real v, float w, real th, uint32_t) { return enc_g(v, w, th); }
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
  static constexpr int LB = 0;
  static constexpr bool theta_ref = true;
  static constexpr bool SR = true;
  __device__ static real dec_f(elem e, float w) { return bf16_to_f32(e.bits) * DEV_INV + w; }
  __device__ static elem enc_f(real v, float w) { return bf16_t{f32_to_bf16((v - w) * DEV_SCALE)}; }
  __device__ static real dec_g(elem e, float w, real th) { return bf16_to_f32(e.bits) * DEV_INV + w * th; }
  __device__ static elem enc_g(real v, float w, real th) { return bf16_t{f32_to_bf16((v - w * th) * DEV_SCALE)}; }
  __device__ static elem enc_f_sr(real v, float w, uint32_t r) { return bf16_t{f32_to_bf16_sr((v - w) * DEV_SCALE, r)}; }
  __device__ static elem enc_g_sr(real v, float w, real th, uint32_t r) {
    return bf16_t{f32_to_bf16_sr((v - w * th) * DEV_SCALE, r)};
  }
  __device__ static elem ld_raw(const void *base, uint32_t boff) {
    return bf16_t{(uint16_t)__builtin_amdgcn_raw_buffer_load_b16(rsrc(base), boff, 0, AHT_LD_CPOL)};
  }
  __device__ static void st_raw(void *base, uint32_t boff, elem v) {
    __builtin_amdgcn_raw_buffer_store_b16(v.bits, rsrc(base), boff, 0, AHT_ST_CPOL);
  }
};

// f deviations as StoreF16, the D3Q7 populations as g_q * 2^10 in binary16 (no theta plane), stochastic rounding
struct StoreF16S : StoreF16 {
  static constexpr int id = 4;
  static constexpr bool theta_ref = false;
  static constexpr bool SR = true;
  __device__ static real dec_g(elem e, float, real) { return (float)e * DEV_INV; }
  __device__ static elem enc_g(real v, float, real) { return (elem)(v * DEV_SCALE); }
  __device__ static elem enc_f_sr(real v, float w, uint32_t r) { return f32_to_f16_sr((v - w) * DEV_SCALE, r); }
  __device__ static elem enc_g_sr(real v, float, real, uint32_t r) { return f32_to_f16_sr(v * DEV_SCALE, r); }
};

struct StoreF64 {
  using elem = double;
  using real = double;
  static constexpr int bytes = 8;
  static constexpr int id = 3;
  static constexpr int LB = 0;
  static constexpr bool theta_ref = false;
  static constexpr bool SR = false;
  __device__ static elem enc_f_sr(real v, float, uint32_t) { return v; }
  __device__ static elem enc_g_sr(real v, float, real, uint32_t) { return v; }
  __device__ static real dec_f(elem e, float) { return e; }
  __device__ static elem enc_f(real v, float) { return v; }
  __device__ static real dec_g(elem e, float, real) { return e; }
  __device__ static elem enc_g(real v, float, real) { return v; }
  __device__ static elem ld_raw(const void *base, uint32_t boff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc(base), boff, 0, AHT_LD_CPOL));
  }
  __device__ static void st_raw(void *base, uint32_t boff, elem v) {
    using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(rsrc(base), boff, 0, AHT_LD_CPOL));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, v), rsrc(base), boff, 0, AHT_ST_CPOL);
  }
};

// f32 populations in the blocked layout: the D3Q19 and D3Q7 lattices of one parity share one array
// [cell >> 11][26][cell & 2047], 2048 consecutive cells of every population side by side, the 19 D3Q19 populations of
// a block followed by its 7 D3Q7 ones (the D3Q7 lattice's base is 19 populations into the first block). The 52 streams
// of a wavefront then fall into a few contiguous regions instead of 52 distant planes: at 600 x 600 x 64 the step
// runs 7 % faster than in [q][cell] (blocks of 1024 and 4096 cells: 0.8 % slower and equal). Used when the array
// stays below 4 GiB (32-bit buffer offsets).
struct StoreF32B : StoreF32 {
  static constexpr int LB = 11;
  static constexpr int QB = PAIR_QB;
};
// the same layout written through the cache: the two-step sweep writes row segments of its column tiles, which start
// and end inside 128-byte lines that the neighbouring tile writes; non-temporal stores send such partial lines to
// HBM one by one (1.24 ms per step at 600 x 600 x 64), cached stores let the two halves meet in L2 (0.89 ms)
struct StoreF32BC : StoreF32B {
  __device__ static void st_raw(void *base, uint32_t boff, elem v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), rsrc(base), boff, 0, 0);
  }
};

// the 16-bit formats in the same interleaved layout, in blocks of 256 cells: [cell >> 8][26][cell & 255] of 2-byte
// deviations, with the f32 theta plane [cell] after the array (the f16 step runs 4.9 % faster than in [q][cell] at
// 600 x 600 x 64)
struct StoreF16B : StoreF16 {
  static constexpr int LB = 8;
  static constexpr int QB = PAIR_QB;
};
struct StoreBF16B : StoreBF16 {
  static constexpr int LB = 8;
  static constexpr int QB = PAIR_QB;
};
struct StoreF16SB : StoreF16S {
  static constexpr int LB = 8;
  static constexpr int QB = PAIR_QB;
};

// element index of population q (of Q) at cell n in a lattice of N cells: [q][cell] when S::LB == 0, else
// [cell >> LB][S::QB][cell & (2^LB - 1)] from the lattice's base (the cell count rounded up to a whole block,
// lattice_cells)
template <class S>
__host__ __device__ __forceinline__ size_t lat_index(int Q, size_t N, int q, size_t n) {
  if constexpr (S::LB == 0) return (size_t)q * N + n;
  else return (((n >> S::LB) * (size_t)S::QB + (size_t)q) << S::LB) | (n & ((size_t(1) << S::LB) - 1));
}
// byte offset of population q's first element, and of cell n from there (32 bits: the lattice is below 4 GiB)
template <class S>
__host__ __device__ __forceinline__ size_t plane_off(int Q, size_t N, int q) {
  if constexpr (S::LB == 0) return (size_t)q * N * S::bytes;
  else return ((size_t)q << S::LB) * S::bytes;
}
template <class S, int Q>
__host__ __device__ __forceinline__ uint32_t cell_off(uint32_t n) {
  if constexpr (S::LB == 0) return n * (uint32_t)S::bytes;
  else return ((n >> S::LB) * ((uint32_t)S::QB << S::LB) + (n & ((1u << S::LB) - 1))) * (uint32_t)S::bytes;
}

// Pulled offsets in a blocked layout with one compare per population. Write the pull offset (in cells) as
// off = B 2^LB + o with 0 <= o < 2^LB (B may be negative). The source cell n - off lies in block (n >> LB) - B at
// position (n & M) - o when n's position in its block (n & M) >= o, and one block further back otherwise, so its byte
// offset is cell_off(n) - K(off), minus D = (QB - 1) 2^LB bytes in the second case, with the uniform
// K(off) = (B QB 2^LB + o) bytes. The stream's base takes -K(off) - D (pull_shift), and a lane picks cell_off(n) + D
// or cell_off(n) (pull_voff), both non-negative. Sources outside [0, N) (face cells at levels 1 and nz - 2, whose
// out-of-domain pulls are replaced afterwards) land in the guard bytes the allocation keeps around the lattice pairs
// (pull_guard_bytes).
__host__ __device__ __forceinline__ constexpr uint32_t pull_dbytes(int lb, int bytes) {
  return ((uint32_t)(PAIR_QB - 1) << lb) * (uint32_t)bytes;
}
__host__ __device__ __forceinline__ int64_t pull_shift(int lb, int bytes, int64_t off) {
  const int64_t B = off >> lb, o = off & (((int64_t)1 << lb) - 1);     // floor division and remainder
  return (B * PAIR_QB * ((int64_t)1 << lb) + o) * bytes + pull_dbytes(lb, bytes);
}
__host__ __device__ __forceinline__ uint32_t pull_rem(int lb, int64_t off) {
  return (uint32_t)(off & (((int64_t)1 << lb) - 1));
}
// bytes kept free before the first and after the last blocked lattice pair of an allocation: the fast path's pulls
// reach at most nx cells outside [0, N) (levels 1 and nz - 2), so whole blocks covering nx + 1 cells
__host__ __device__ __forceinline__ size_t pull_guard_bytes(int lb, int bytes, int nx) {
  if (!lb) return 0;
  const size_t blocks = ((size_t)nx + 1 + ((size_t)1 << lb) - 1) >> lb;
  return (blocks * PAIR_QB << lb) * (size_t)bytes;
}
template <class S, int Q>
__host__ __device__ __forceinline__ uint32_t pull_voff(uint32_t n, uint32_t o) {
  static_assert(S::LB != 0 && S::QB == PAIR_QB, "blocked pair layouts only");
  const uint32_t own = cell_off<S, Q>(n);
  return (n & ((1u << S::LB) - 1)) < o ? own : own + pull_dbytes(S::LB, S::bytes);
}

// f32 theta of a theta-referenced D3Q7 lattice at a byte offset from a base
__device__ __forceinline__ float ld_theta(const void *base, uint32_t boff) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsrc(base), boff, 0, AHT_LD_CPOL));
}
__device__ __forceinline__ void st_theta(void *base, uint32_t boff, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), rsrc(base), boff, 0, AHT_ST_CPOL);
}

}  // namespace aht
