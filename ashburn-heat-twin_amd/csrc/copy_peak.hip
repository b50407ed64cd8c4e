// Measured HBM copy peak and the calibration reference for rocprofv3 FETCH_SIZE / WRITE_SIZE counters.
// Copies `bytes` (default 4 GiB) between two device buffers with 16-byte loads and stores in several launch shapes,
// `reps` launches each, and prints one JSON line per variant with the event-timed read+write bandwidth, then one
// line with the best (the measured copy peak).
// The lattice variants move 26 population arrays [q][cell] (D3Q19 + D3Q7) of a 600 x 600 x 64 grid with one thread
// per cell, the access patterns of a step without its arithmetic: `lattice` copies every array at the cell itself,
// `lattice_pull` reads at the pull source cell - c_q (clamped), as the two-lattice pull step does; `lattice_aa_even`
// and `lattice_aa_odd` work in place on one lattice as the two steps of the AA pattern do (even: read slot q and
// write slot opp(q) of the cell itself; odd: read slot opp(q) at cell - c_q, write slot q at cell + c_q, interior
// cells); `lattice_pull_nt` is the pull with non-temporal stores (as the step stores), `lattice16_pull_nt` the same
// for the 16-bit formats: 26 arrays of 2-byte elements plus an f32 theta array pulled at the 7 D3Q7 offsets. They
// give the rates attainable with 52+ concurrent streams, a tighter bound for the step than the plain copy.
// `lattice_aosoa_pull_nt` is the f32 pull pattern on a blocked layout [cell / 64][q][cell % 64] (26 populations of
// 64 consecutive cells side by side), the bound of a layout in which a wavefront's streams are one contiguous region.
// `lattice_pull_nt_x2` / `_x4` move 2 / 4 consecutive cells per thread (8- / 16-byte accesses, the pulled loads
// dword-aligned), the bound of a step that widens its per-lane accesses. `lattice_blocked_pull_nt` has a
// 19-population and a 7-population lattice, each [cell / 1024][q][cell % 1024]; `lattice_pair_pull_nt` is the f32
// step's own layout, both in one array [cell / 2048][19 + 7][cell % 2048]; `lattice16_pair_pull_nt` the same for
// 2-byte elements in blocks of 256 cells, with the f32 theta array of the 16-bit formats pulled beside it.
// `lattice16s_pair_pull_nt` is the f16s step's pattern (the 16-bit pair without theta array);
// `lattice16_pair_theta_in_block_pull_nt` keeps the f32 theta values inside each 256-cell block (28 two-byte units);
// `lattice_colmajor_pull_nt` the f32 pair with column blocks outermost and levels inside them; `lattice_pair_aa_*` and
// `lattice_pair_esotwist*` the in-place patterns in the pair layout. `lattice_pair_{noshift,zonly,xonly,yonly}_nt`
// and `lattice16s_pair_{noshift,zonly,xonly,yonly}_nt` take the pair patterns apart: every stream read at the cell
// itself, or pulled along z, x or y only; `lattice_pair_pull_nt_w{2,4,16,64}` and `lattice16s_pair_pull_nt_w{...}`
// give the pull pattern the step's XCD-aware tile order (runs of W consecutive 256-cell tiles per XCD).
// Usage: aht_copy_peak [bytes] [reps] [variant] [waves] (waves: cap of wavefronts per SIMD for the two pair patterns)
//   (variant: all | grid | flat | unroll4 | unroll4_xcd | unroll4_nt | lattice | lattice_pull | lattice_aa_even
//    | lattice_aa_odd | lattice_pull_nt | lattice16_pull_nt | lattice_aosoa_pull_nt | lattice_pull_nt_x2
//    | lattice_pull_nt_x4 | lattice_blocked_pull_nt | lattice_pair_pull_nt | lattice16_pair_pull_nt
//    | lattice16s_pair_pull_nt | lattice_colmajor_pull_nt | lattice_pair_aa_even | lattice_pair_aa_odd
//    | lattice_pair_esotwist | lattice_pair_aa_even_nt | lattice_pair_aa_odd_nt | lattice_pair_esotwist_nt
//    | lattice[16s]_pair_{noshift,zonly,xonly,yonly}_nt | lattice[16s]_pair_pull_nt_w{2,4,16,64}
//    | lattice16_pair_theta_in_block_pull_nt)
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

// D3Q19 then D3Q7 velocities, as the lattice tables of the step
__constant__ int LCX[26] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0};
__constant__ int LCY[26] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 1, -1, 0, 0};
__constant__ int LCZ[26] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0, 0, 1, -1};

template <bool PULL, bool NT>
__global__ __launch_bounds__(256) void copy_lattice(const float *__restrict__ a, float *__restrict__ b, uint32_t N,
                                                    int nx, int nxy) {
  const uint32_t n = blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  float v[26];
#pragma unroll
  for (int q = 0; q < 26; q++) {
    int64_t src = n;
    if (PULL) {
      src -= LCX[q] + LCY[q] * nx + LCZ[q] * nxy;
      src = src < 0 ? 0 : (src >= (int64_t)N ? (int64_t)N - 1 : src);
    }
    v[q] = a[(size_t)q * N + (size_t)src];
  }
#pragma unroll
  for (int q = 0; q < 26; q++) {
    if (NT) __builtin_nontemporal_store(v[q], &b[(size_t)q * N + n]);
    else b[(size_t)q * N + n] = v[q];
  }
}

// 16-bit lattice: 26 arrays of 2-byte elements [q][cell] then an f32 theta array, pulled; non-temporal stores
__global__ __launch_bounds__(256) void copy_lattice16(const uint16_t *__restrict__ a, uint16_t *__restrict__ b,
                                                      uint32_t N, int nx, int nxy) {
  const uint32_t n = blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  const float *ta = (const float *)(a + (size_t)26 * N);
  float *tb = (float *)(b + (size_t)26 * N);
  uint16_t v[26];
  float th = 0.0f;
#pragma unroll
  for (int q = 0; q < 26; q++) {
    int64_t src = (int64_t)n - (LCX[q] + LCY[q] * nx + LCZ[q] * nxy);
    src = src < 0 ? 0 : (src >= (int64_t)N ? (int64_t)N - 1 : src);
    v[q] = a[(size_t)q * N + (size_t)src];
    if (q >= 19) th += ta[src];
  }
#pragma unroll
  for (int q = 0; q < 26; q++) __builtin_nontemporal_store(v[q], &b[(size_t)q * N + n]);
  __builtin_nontemporal_store(th, &tb[n]);
}

__constant__ int LOPP[26] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17, 19, 21, 20, 23, 22, 25, 24};

// in place, AA pattern; ODD moves every population one cell along c_q (interior cells; boundary cells only copy)
template <bool ODD>
__global__ __launch_bounds__(256) void copy_lattice_aa(float *__restrict__ a, uint32_t N, int nx, int ny, int nz) {
  const uint32_t n = blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  const int nxy = nx * ny, k = (int)(n / nxy), c = (int)(n - (uint32_t)k * nxy), j = c / nx, i = c - j * nx;
  const bool in = ODD && i > 0 && i < nx - 1 && j > 0 && j < ny - 1 && k > 0 && k < nz - 1;
  float v[26];
#pragma unroll
  for (int q = 0; q < 26; q++) {
    const int d = in ? LCX[q] + LCY[q] * nx + LCZ[q] * nxy : 0;
    v[q] = a[(size_t)(ODD ? LOPP[q] : q) * N + (size_t)((int64_t)n - d)];
  }
#pragma unroll
  for (int q = 0; q < 26; q++) {
    const int d = in ? LCX[q] + LCY[q] * nx + LCZ[q] * nxy : 0;
    a[(size_t)(ODD ? q : LOPP[q]) * N + (size_t)((int64_t)n + d)] = v[q];
  }
}

// in-place single-lattice patterns in the f32 step's interleaved pair layout [cell >> LB][26][cell & (2^LB - 1)]:
// AA (even step: own cell, read slot q, write slot opp(q); odd step: read slot opp(q) at n - c_q, write slot q at
// n + c_q) and the esoteric twist (every step: population q read from and written back to the cell n + e_q, with
// e_q = max(-c_q, 0) per component, i.e. the eight cells of the cube above n, the slot switching to opp(q)); interior
// cells only take the offsets, so every access stays inside the lattice
template <int LB, int KIND, bool NT>   // KIND 0: AA even, 1: AA odd, 2: esoteric twist
__global__ __launch_bounds__(256) void copy_lattice_inplace_pair(float *__restrict__ a, uint32_t N, int nx, int ny,
                                                                 int nz) {
  const uint32_t n = blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  constexpr uint32_t M = (1u << LB) - 1;
  const int nxy = nx * ny, k = (int)(n / nxy), c = (int)(n - (uint32_t)k * nxy), j = c / nx, i = c - j * nx;
  const bool in = KIND != 0 && i > 0 && i < nx - 1 && j > 0 && j < ny - 1 && k > 0 && k < nz - 1;
  auto at = [&](int q, int64_t cell) { return ((((size_t)cell >> LB) * 26 + q) << LB) | ((size_t)cell & M); };
  auto off = [&](int q) {
    if (!in) return 0;
    if (KIND == 1) return LCX[q] + LCY[q] * nx + LCZ[q] * nxy;
    return (LCX[q] < 0 ? 1 : 0) + (LCY[q] < 0 ? nx : 0) + (LCZ[q] < 0 ? nxy : 0);
  };
  float v[26];
#pragma unroll
  for (int q = 0; q < 26; q++) {
    const int d = off(q);
    v[q] = a[at(KIND == 1 ? LOPP[q] : q, KIND == 1 ? (int64_t)n - d : (int64_t)n + d)];
  }
#pragma unroll
  for (int q = 0; q < 26; q++) {
    float *p = &a[at(KIND == 1 ? q : LOPP[q], (int64_t)n + off(q))];
    if (NT) __builtin_nontemporal_store(v[q], p);
    else *p = v[q];
  }
}

// the f32 step's layout: D3Q19 lattice then D3Q7 lattice, each [cell >> LB][q][cell & (2^LB - 1)] (N a multiple of
// 2^LB), pulled, non-temporal stores
template <int LB>
__global__ __launch_bounds__(256) void copy_lattice_blocked(const float *__restrict__ a, float *__restrict__ b,
                                                            uint32_t N, int nx, int nxy) {
  const uint32_t n = blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  constexpr uint32_t M = (1u << LB) - 1;
  float v[26];
#pragma unroll
  for (int q = 0; q < 26; q++) {
    int64_t src = (int64_t)n - (LCX[q] + LCY[q] * nx + LCZ[q] * nxy);
    src = src < 0 ? 0 : (src >= (int64_t)N ? (int64_t)N - 1 : src);
    const size_t blk = (size_t)src >> LB, lo = (size_t)src & M;
    v[q] = q < 19 ? a[((blk * 19 + q) << LB) | lo] : a[(size_t)19 * N + (((blk * 7 + (q - 19)) << LB) | lo)];
  }
  const size_t blk = n >> LB, lo = n & M;
#pragma unroll
  for (int q = 0; q < 26; q++) {
    float *d = q < 19 ? &b[((blk * 19 + q) << LB) | lo] : &b[(size_t)19 * N + (((blk * 7 + (q - 19)) << LB) | lo)];
    __builtin_nontemporal_store(v[q], d);
  }
}

// the f32 step's layout: the D3Q19 and D3Q7 populations in one array [cell >> LB][26][cell & (2^LB - 1)] (N a
// multiple of 2^LB), pulled, non-temporal stores
template <int LB>
__global__ __launch_bounds__(256) void copy_lattice_pair(const float *__restrict__ a, float *__restrict__ b, uint32_t N,
                                                         int nx, int nxy) {
  const uint32_t n = blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  constexpr uint32_t M = (1u << LB) - 1;
  float v[26];
#pragma unroll
  for (int q = 0; q < 26; q++) {
    int64_t src = (int64_t)n - (LCX[q] + LCY[q] * nx + LCZ[q] * nxy);
    src = src < 0 ? 0 : (src >= (int64_t)N ? (int64_t)N - 1 : src);
    v[q] = a[((((size_t)src >> LB) * 26 + q) << LB) | ((size_t)src & M)];
  }
#pragma unroll
  for (int q = 0; q < 26; q++)
    __builtin_nontemporal_store(v[q], &b[((((size_t)n >> LB) * 26 + q) << LB) | ((size_t)n & M)]);
}

// the pair patterns taken apart (T = float, LB = 11: the f32 step's; T = uint16_t, LB = 8: the f16s step's). MODE:
// 0 pull, 1 every stream at the cell itself, 2 pulled along z only, 3 along x only, 4 along y only. Tile t of 256 cells
// is taken by block b in the step's XCD-aware order (tile_of_block with window `win`; 1 = linear).
__device__ __forceinline__ uint32_t tile_of(uint32_t b, uint32_t G, uint32_t win) {
  if (win <= 1) return b;
  const uint32_t x = b & 7u, i = b >> 3, lw = (uint32_t)__builtin_ctz(win), span_mask = (win << 3) - 1u;
  if (b >= (G & ~span_mask)) return b;
  return ((i >> lw) << (lw + 3)) + (x << lw) + (i & (win - 1u));
}
template <class T, int LB, int MODE>
__global__ __launch_bounds__(256) void copy_pair_parts(const T *__restrict__ a, T *__restrict__ b, uint32_t N, int nx,
                                                       int nxy, uint32_t win) {
  const uint32_t n = tile_of(blockIdx.x, gridDim.x, win) * 256 + threadIdx.x;
  if (n >= N) return;
  constexpr uint32_t M = (1u << LB) - 1;
  T v[26];
#pragma unroll
  for (int q = 0; q < 26; q++) {
    int cx = LCX[q], cy = LCY[q], cz = LCZ[q];
    if (MODE == 1) cx = cy = cz = 0;
    if (MODE == 2) cx = cy = 0;
    if (MODE == 3) cy = cz = 0;
    if (MODE == 4) cx = cz = 0;
    int64_t src = (int64_t)n - (cx + cy * nx + cz * nxy);
    src = src < 0 ? 0 : (src >= (int64_t)N ? (int64_t)N - 1 : src);
    v[q] = a[((((size_t)src >> LB) * 26 + q) << LB) | ((size_t)src & M)];
  }
#pragma unroll
  for (int q = 0; q < 26; q++)
    __builtin_nontemporal_store(v[q], &b[((((size_t)n >> LB) * 26 + q) << LB) | ((size_t)n & M)]);
}

// 16-bit populations interleaved as the f32 pair, [cell >> LB][26][cell & (2^LB - 1)], and (THETA) an f32 theta
// array [cell] pulled at the 7 D3Q7 offsets, non-temporal stores; without THETA the pattern of the f16s step
template <int LB, bool THETA>
__global__ __launch_bounds__(256) void copy_lattice16_pair(const uint16_t *__restrict__ a, uint16_t *__restrict__ b,
                                                           uint32_t N, int nx, int nxy) {
  const uint32_t n = blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  constexpr uint32_t M = (1u << LB) - 1;
  const float *ta = (const float *)(a + (size_t)26 * N);
  float *tb = (float *)(b + (size_t)26 * N);
  uint16_t v[26];
  float th = 0.0f;
#pragma unroll
  for (int q = 0; q < 26; q++) {
    int64_t src = (int64_t)n - (LCX[q] + LCY[q] * nx + LCZ[q] * nxy);
    src = src < 0 ? 0 : (src >= (int64_t)N ? (int64_t)N - 1 : src);
    v[q] = a[((((size_t)src >> LB) * 26 + q) << LB) | ((size_t)src & M)];
    if (THETA && q >= 19) th += ta[src];
  }
#pragma unroll
  for (int q = 0; q < 26; q++)
    __builtin_nontemporal_store(v[q], &b[((((size_t)n >> LB) * 26 + q) << LB) | ((size_t)n & M)]);
  if (THETA) __builtin_nontemporal_store(th, &tb[n]);
}

// the 16-bit pair with the f32 theta values inside each block: [cell >> 8][28][256] 2-byte units, units 0-25 the 26
// populations, units 26-27 the 256 theta values of the block's cells (as f32), theta pulled at the 7 D3Q7 offsets
__global__ __launch_bounds__(256) void copy_lattice16_pair_theta_in_block(const uint16_t *__restrict__ a,
                                                                         uint16_t *__restrict__ b, uint32_t N, int nx,
                                                                         int nxy) {
  const uint32_t n = blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  const float *ta = (const float *)a;
  float *tb = (float *)b;
  uint16_t v[26];
  float th = 0.0f;
#pragma unroll
  for (int q = 0; q < 26; q++) {
    int64_t src = (int64_t)n - (LCX[q] + LCY[q] * nx + LCZ[q] * nxy);
    src = src < 0 ? 0 : (src >= (int64_t)N ? (int64_t)N - 1 : src);
    v[q] = a[((((size_t)src >> 8) * 28 + q) << 8) | ((size_t)src & 255)];
    if (q >= 19) th += ta[((size_t)src >> 8) * 3584 + 3328 + ((size_t)src & 255)];
  }
#pragma unroll
  for (int q = 0; q < 26; q++)
    __builtin_nontemporal_store(v[q], &b[((((size_t)n >> 8) * 28 + q) << 8) | ((size_t)n & 255)]);
  __builtin_nontemporal_store(th, &tb[((size_t)n >> 8) * 3584 + 3328 + ((size_t)n & 255)]);
}

// the f32 pair with column blocks outermost: element (q, level k, column c) at
// ((((c >> LB) * nz + k) * 26 + q) << LB) | (c & (2^LB - 1)), so the levels of one column block are adjacent in memory;
// thread blocks walk a column block through its levels (2^LB / 256 thread blocks per level) before the next one
template <int LB>
__global__ __launch_bounds__(256) void copy_lattice_colmajor(const float *__restrict__ a, float *__restrict__ b, int nx,
                                                             int nxy, int nz) {
  constexpr uint32_t M = (1u << LB) - 1, SUB = (1u << LB) / 256;
  const uint32_t t = blockIdx.x, sub = t % SUB, kk = (t / SUB) % nz, cb = t / (SUB * nz);
  const int c = (int)((cb << LB) + sub * 256 + threadIdx.x), k = (int)kk;
  if (c >= nxy) return;
  float v[26];
#pragma unroll
  for (int q = 0; q < 26; q++) {
    int sc = c - (LCX[q] + LCY[q] * nx), sk = k - LCZ[q];
    sc = sc < 0 ? 0 : (sc >= nxy ? nxy - 1 : sc);
    sk = sk < 0 ? 0 : (sk >= nz ? nz - 1 : sk);
    v[q] = a[(((((size_t)sc >> LB) * nz + sk) * 26 + q) << LB) | ((size_t)sc & M)];
  }
#pragma unroll
  for (int q = 0; q < 26; q++)
    __builtin_nontemporal_store(v[q], &b[(((((size_t)c >> LB) * nz + k) * 26 + q) << LB) | ((size_t)c & M)]);
}

// blocked layout: element (q, n) at ((n / 64) * 26 + q) * 64 + n % 64
__global__ __launch_bounds__(256) void copy_lattice_aosoa(const float *__restrict__ a, float *__restrict__ b, uint32_t N,
                                                          int nx, int nxy) {
  const uint32_t n = blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  float v[26];
#pragma unroll
  for (int q = 0; q < 26; q++) {
    int64_t src = (int64_t)n - (LCX[q] + LCY[q] * nx + LCZ[q] * nxy);
    src = src < 0 ? 0 : (src >= (int64_t)N ? (int64_t)N - 1 : src);
    v[q] = a[((size_t)(src >> 6) * 26 + q) * 64 + (src & 63)];
  }
#pragma unroll
  for (int q = 0; q < 26; q++) __builtin_nontemporal_store(v[q], &b[((size_t)(n >> 6) * 26 + q) * 64 + (n & 63)]);
}

// W consecutive cells per thread (N a multiple of W): W-wide loads at the pulled position (clamped), W-wide
// non-temporal stores
template <int W>
__global__ __launch_bounds__(256) void copy_lattice_wide(const float *__restrict__ a, float *__restrict__ b, uint32_t N,
                                                         int nx, int nxy) {
  typedef float vw __attribute__((ext_vector_type(W)));
  const uint32_t n = (blockIdx.x * 256 + threadIdx.x) * W;
  if (n >= N) return;
  vw v[26];
#pragma unroll
  for (int q = 0; q < 26; q++) {
    int64_t src = (int64_t)n - (LCX[q] + LCY[q] * nx + LCZ[q] * nxy);
    src = src < 0 ? 0 : (src > (int64_t)N - W ? (int64_t)N - W : src);
    v[q] = *(const vw *)&a[(size_t)q * N + (size_t)src];
  }
#pragma unroll
  for (int q = 0; q < 26; q++) __builtin_nontemporal_store(v[q], (vw *)&b[(size_t)q * N + n]);
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
  // optional cap of wavefronts per SIMD for the pair patterns (15, 17): LDS reserved per 256-thread block so that at
  // most `waves` blocks fit on a CU (the copies use no LDS themselves); 0 = none
  const int waves = argc > 4 ? atoi(argv[4]) : 0;
  const size_t n = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  CK(hipMalloc(&a, n * 16));
  CK(hipMalloc(&b, n * 16));
  CK(hipMemset(a, 0, n * 16));
  CK(hipMemset(b, 0, n * 16));
  int cus = 0, lds_cu = 0;
  CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  CK(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, 0));
  const unsigned cap_lds = waves > 0 ? (unsigned)(lds_cu / waves) / 1024u * 1024u : 0u;   // 256 threads: waves blocks
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  const int NV = 42;
  const char *names[NV] = {"grid", "flat", "unroll4", "unroll4_xcd", "unroll4_nt", "lattice", "lattice_pull",
                           "lattice_aa_even", "lattice_aa_odd", "lattice_pull_nt", "lattice16_pull_nt",
                           "lattice_aosoa_pull_nt", "lattice_pull_nt_x2", "lattice_pull_nt_x4",
                           "lattice_blocked_pull_nt", "lattice_pair_pull_nt", "lattice16_pair_pull_nt",
                           "lattice16s_pair_pull_nt", "lattice_colmajor_pull_nt", "lattice_pair_aa_even",
                           "lattice_pair_aa_odd", "lattice_pair_esotwist", "lattice_pair_aa_even_nt",
                           "lattice_pair_aa_odd_nt", "lattice_pair_esotwist_nt",
                           "lattice_pair_noshift_nt", "lattice_pair_zonly_nt", "lattice_pair_xonly_nt",
                           "lattice_pair_yonly_nt", "lattice_pair_pull_nt_w2", "lattice_pair_pull_nt_w4",
                           "lattice_pair_pull_nt_w16", "lattice_pair_pull_nt_w64",
                           "lattice16s_pair_noshift_nt", "lattice16s_pair_zonly_nt", "lattice16s_pair_xonly_nt",
                           "lattice16s_pair_yonly_nt", "lattice16s_pair_pull_nt_w2", "lattice16s_pair_pull_nt_w4",
                           "lattice16s_pair_pull_nt_w16", "lattice16s_pair_pull_nt_w64",
                           "lattice16_pair_theta_in_block_pull_nt"};
  const uint32_t wins[4] = {2, 4, 16, 64};
  const int lnx = 600, lnxy = 600 * 600;
  const uint32_t lN = 600u * 600u * 64u;   // lattice variants: 26 * lN floats each way (2.4 GB)
  double best = 0;
  const char *best_name = "";
  const uint32_t ncb = (lnxy + 2047) / 2048;     // column blocks of the column-major variant (2.44 GB per buffer)
  for (int v = 0; v < NV; v++) {
    if (strcmp(which, "all") && strcmp(which, names[v])) continue;
    if (v >= 5 && n * 16 < (size_t)26 * ncb * 2048 * 64 * 4) continue;   // the lattices need 2.44 GB per buffer
    auto launch = [&]() {
      switch (v) {
        case 0: hipLaunchKernelGGL(copy_grid, dim3(cus * 32), dim3(256), 0, 0, a, b, n); break;
        case 1: hipLaunchKernelGGL(copy_flat, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, a, b, n); break;
        case 2: hipLaunchKernelGGL((copy_unroll4<false, false>), dim3((unsigned)((n + 1023) / 1024)), dim3(256), 0, 0, a, b, n); break;
        case 3: hipLaunchKernelGGL((copy_unroll4<true, false>), dim3((unsigned)((n + 1023) / 1024)), dim3(256), 0, 0, a, b, n); break;
        case 4: hipLaunchKernelGGL((copy_unroll4<true, true>), dim3((unsigned)((n + 1023) / 1024)), dim3(256), 0, 0, a, b, n); break;
        case 5: hipLaunchKernelGGL((copy_lattice<false, false>), dim3((lN + 255) / 256), dim3(256), 0, 0, (const float *)a, (float *)b, lN, lnx, lnxy); break;
        case 6: hipLaunchKernelGGL((copy_lattice<true, false>), dim3((lN + 255) / 256), dim3(256), 0, 0, (const float *)a, (float *)b, lN, lnx, lnxy); break;
        case 7: hipLaunchKernelGGL((copy_lattice_aa<false>), dim3((lN + 255) / 256), dim3(256), 0, 0, (float *)a, lN, lnx, lnx, 64); break;
        case 8: hipLaunchKernelGGL((copy_lattice_aa<true>), dim3((lN + 255) / 256), dim3(256), 0, 0, (float *)a, lN, lnx, lnx, 64); break;
        case 9: hipLaunchKernelGGL((copy_lattice<true, true>), dim3((lN + 255) / 256), dim3(256), 0, 0, (const float *)a, (float *)b, lN, lnx, lnxy); break;
        case 10: hipLaunchKernelGGL(copy_lattice16, dim3((lN + 255) / 256), dim3(256), 0, 0, (const uint16_t *)a, (uint16_t *)b, lN, lnx, lnxy); break;
        case 11: hipLaunchKernelGGL(copy_lattice_aosoa, dim3((lN + 255) / 256), dim3(256), 0, 0, (const float *)a, (float *)b, lN, lnx, lnxy); break;
        case 12: hipLaunchKernelGGL(copy_lattice_wide<2>, dim3((lN / 2 + 255) / 256), dim3(256), 0, 0, (const float *)a, (float *)b, lN, lnx, lnxy); break;
        case 13: hipLaunchKernelGGL(copy_lattice_wide<4>, dim3((lN / 4 + 255) / 256), dim3(256), 0, 0, (const float *)a, (float *)b, lN, lnx, lnxy); break;
        case 14: hipLaunchKernelGGL(copy_lattice_blocked<10>, dim3((lN + 255) / 256), dim3(256), 0, 0, (const float *)a, (float *)b, lN, lnx, lnxy); break;
        case 15: hipLaunchKernelGGL(copy_lattice_pair<11>, dim3((lN + 255) / 256), dim3(256), cap_lds, 0, (const float *)a, (float *)b, lN, lnx, lnxy); break;
        case 16: hipLaunchKernelGGL((copy_lattice16_pair<8, true>), dim3((lN + 255) / 256), dim3(256), 0, 0, (const uint16_t *)a, (uint16_t *)b, lN, lnx, lnxy); break;
        case 17: hipLaunchKernelGGL((copy_lattice16_pair<8, false>), dim3((lN + 255) / 256), dim3(256), cap_lds, 0, (const uint16_t *)a, (uint16_t *)b, lN, lnx, lnxy); break;
        case 18: hipLaunchKernelGGL(copy_lattice_colmajor<11>, dim3(ncb * 64 * 8), dim3(256), 0, 0, (const float *)a, (float *)b, lnx, lnxy, 64); break;
        case 19: hipLaunchKernelGGL((copy_lattice_inplace_pair<11, 0, false>), dim3((lN + 255) / 256), dim3(256), 0, 0, (float *)a, lN, lnx, lnx, 64); break;
        case 20: hipLaunchKernelGGL((copy_lattice_inplace_pair<11, 1, false>), dim3((lN + 255) / 256), dim3(256), 0, 0, (float *)a, lN, lnx, lnx, 64); break;
        case 21: hipLaunchKernelGGL((copy_lattice_inplace_pair<11, 2, false>), dim3((lN + 255) / 256), dim3(256), 0, 0, (float *)a, lN, lnx, lnx, 64); break;
        case 22: hipLaunchKernelGGL((copy_lattice_inplace_pair<11, 0, true>), dim3((lN + 255) / 256), dim3(256), 0, 0, (float *)a, lN, lnx, lnx, 64); break;
        case 23: hipLaunchKernelGGL((copy_lattice_inplace_pair<11, 1, true>), dim3((lN + 255) / 256), dim3(256), 0, 0, (float *)a, lN, lnx, lnx, 64); break;
        case 24: hipLaunchKernelGGL((copy_lattice_inplace_pair<11, 2, true>), dim3((lN + 255) / 256), dim3(256), 0, 0, (float *)a, lN, lnx, lnx, 64); break;
        case 25: hipLaunchKernelGGL((copy_pair_parts<float, 11, 1>), dim3(lN / 256), dim3(256), 0, 0, (const float *)a, (float *)b, lN, lnx, lnxy, 1u); break;
        case 26: hipLaunchKernelGGL((copy_pair_parts<float, 11, 2>), dim3(lN / 256), dim3(256), 0, 0, (const float *)a, (float *)b, lN, lnx, lnxy, 1u); break;
        case 27: hipLaunchKernelGGL((copy_pair_parts<float, 11, 3>), dim3(lN / 256), dim3(256), 0, 0, (const float *)a, (float *)b, lN, lnx, lnxy, 1u); break;
        case 28: hipLaunchKernelGGL((copy_pair_parts<float, 11, 4>), dim3(lN / 256), dim3(256), 0, 0, (const float *)a, (float *)b, lN, lnx, lnxy, 1u); break;
        case 29: case 30: case 31: case 32:
          hipLaunchKernelGGL((copy_pair_parts<float, 11, 0>), dim3(lN / 256), dim3(256), 0, 0, (const float *)a, (float *)b, lN, lnx, lnxy, wins[v - 29]); break;
        case 33: hipLaunchKernelGGL((copy_pair_parts<uint16_t, 8, 1>), dim3(lN / 256), dim3(256), 0, 0, (const uint16_t *)a, (uint16_t *)b, lN, lnx, lnxy, 1u); break;
        case 34: hipLaunchKernelGGL((copy_pair_parts<uint16_t, 8, 2>), dim3(lN / 256), dim3(256), 0, 0, (const uint16_t *)a, (uint16_t *)b, lN, lnx, lnxy, 1u); break;
        case 35: hipLaunchKernelGGL((copy_pair_parts<uint16_t, 8, 3>), dim3(lN / 256), dim3(256), 0, 0, (const uint16_t *)a, (uint16_t *)b, lN, lnx, lnxy, 1u); break;
        case 36: hipLaunchKernelGGL((copy_pair_parts<uint16_t, 8, 4>), dim3(lN / 256), dim3(256), 0, 0, (const uint16_t *)a, (uint16_t *)b, lN, lnx, lnxy, 1u); break;
        case 37: case 38: case 39: case 40:
          hipLaunchKernelGGL((copy_pair_parts<uint16_t, 8, 0>), dim3(lN / 256), dim3(256), 0, 0, (const uint16_t *)a, (uint16_t *)b, lN, lnx, lnxy, wins[v - 37]); break;
        case 41: hipLaunchKernelGGL(copy_lattice16_pair_theta_in_block, dim3(lN / 256), dim3(256), 0, 0, (const uint16_t *)a, (uint16_t *)b, lN, lnx, lnxy); break;
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
    const size_t moved = (v == 10 || v == 16 || v == 41) ? (size_t)28 * lN * 2
                         : ((v == 17 || v >= 33) ? (size_t)26 * lN * 2 : (v >= 5 ? (size_t)26 * lN * 4 : n * 16));
    const double per = ms / reps, gbs = 2.0 * moved / (per * 1e-3) / 1e9;
    printf("{\"kernel\": \"copy_%s\", \"bytes_copied\": %zu, \"reps\": %d, \"ms_per_launch\": %.5f, "
           "\"GBps_read_plus_write\": %.1f, \"cus\": %d, \"waves_cap\": %d}\n", names[v], moved, reps, per, gbs, cus,
           (v == 15 || v == 17) ? waves : 0);
    if (v < 5 && gbs > best) { best = gbs; best_name = names[v]; }
  }
  printf("{\"copy_peak_GBps\": %.1f, \"variant\": \"%s\"}\n", best, best_name);
  CK(hipFree(a));
  CK(hipFree(b));
  return 0;
}
