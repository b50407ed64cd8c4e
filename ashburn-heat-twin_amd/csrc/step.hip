// CDNA4 (gfx950) kernels of the coupled D3Q19 + D3Q7 lattice-Boltzmann step, f32 two-lattice pull scheme.
//
// Thread mapping: one thread per cell. The x-y plane is flattened into columns c = j*nx + i so that the 64 lanes of a
// wavefront touch 64 consecutive cells of every [q][cell] stream (256 contiguous bytes; the stores and the
// unshifted loads are 256-byte aligned whenever nx*ny is a multiple of 64). blockIdx.y is the level k, starting at the
// lowest fluid level of the grid, so all-solid planes are never launched. Solid cells above that plane exit at once.
//
// Two streaming paths feed one collision:
//   * fast path: interior cells (not on a lateral face, below the lid) at or above the column's kfast level, the first
//     level where every source cell in the 3 x 3 x 3 neighbourhood is fluid (host-computed). All 26 pulls are plain
//     loads through buffer descriptors: one per population with the pull offset folded into the scalar base, so the
//     only per-lane address is the cell's byte offset. Drag and the rooftop source cannot apply there (k > ktop).
//   * boundary path: every branch of js/live/kernel.js:159-243 per population (lid first, then lateral faces, then the
//     interior pull with half-way bounce-back from solid sources), with the 3 x 3 neighbour ktop values loaded once.
#include "kernels.h"

namespace aht {

struct Outflow {
  float rn, uxn, uyn, uzn, usqn, uxc, uyc, usqc;
};

// raw buffer descriptor over a population array (gfx9 dword3); offsets are the 32-bit byte offsets of cells
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, -1, 0x00020000);
}
// the b32 builtins move raw bits: bit casts, never numeric conversions
__device__ __forceinline__ float ldb(const float *base, uint32_t boff) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsrc(base), boff, 0, 0));
}
__device__ __forceinline__ void stb(float *base, uint32_t boff, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), rsrc(base), boff, 0, 0);
}
// base address of population q shifted back by `off` cells (the pull offset), as a plain integer
__device__ __forceinline__ const float *shifted(const float *a, size_t N, int q, int off) {
  return (const float *)((uintptr_t)a + (uintptr_t)((int64_t)q * (int64_t)N - off) * 4u);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ void split_column(const Grid &g, int c, int &i, int &j) {
  int jj = (int)((float)c * g.inv_nx);
  int ii = c - jj * g.nx;
  if (ii < 0) { jj -= 1; ii += g.nx; }
  else if (ii >= g.nx) { jj += 1; ii -= g.nx; }
  i = ii; j = jj;
}

// packed column word: kt | ktop << 10 | kfast << 20 (nz < 1024)
__device__ __forceinline__ int col_kt(uint32_t x) { return (int)(x & 1023u); }
__device__ __forceinline__ int col_ktop(uint32_t x) { return (int)((x >> 10) & 1023u); }
__device__ __forceinline__ int col_kfast(uint32_t x) { return (int)(x >> 20); }

// outflow-face prelude from the own cell's pre-stream populations (js/live/kernel.js:122-148)
__device__ __forceinline__ Outflow outflow_prelude(const Grid &g, const float *__restrict__ fA, size_t N, uint32_t n,
                                                   int i, int j) {
  const Params &P = g.P;
  Outflow o;
  o.rn = 1.0f; o.uxn = 0.0f; o.uyn = 0.0f; o.uzn = 0.0f;
  const bool onb = (i == 0 && P.inflow_w == 0) || (i == g.nx - 1 && P.inflow_e == 0) ||
                   (j == 0 && P.inflow_s == 0) || (j == g.ny - 1 && P.inflow_n == 0);
  if (onb) {
    float a[19];
#pragma unroll
    for (int q = 0; q < 19; q++) a[q] = fA[(size_t)q * N + n];
    float r = a[0], sx = 0.f, sy = 0.f, sz = 0.f;
#pragma unroll
    for (int q = 1; q < 19; q++) {
      r += a[q];
      sx += (float)CX[q] * a[q];
      sy += (float)CY[q] * a[q];
      sz += (float)CZ[q] * a[q];
    }
    const float ir = rcp(r);
    o.rn = r; o.uxn = sx * ir; o.uyn = sy * ir; o.uzn = sz * ir;
  }
  o.usqn = o.uxn * o.uxn + o.uyn * o.uyn + o.uzn * o.uzn;
  float uxc = o.uxn, uyc = o.uyn;
  if (P.oclip != 0) {                       // backflow-clipped outlet: no inward normal velocity
    if (i == 0 && P.inflow_w == 0 && uxc > 0.0f) uxc = 0.0f;
    if (i == g.nx - 1 && P.inflow_e == 0 && uxc < 0.0f) uxc = 0.0f;
    if (j == 0 && P.inflow_s == 0 && uyc > 0.0f) uyc = 0.0f;
    if (j == g.ny - 1 && P.inflow_n == 0 && uyc < 0.0f) uyc = 0.0f;
  }
  o.uxc = uxc; o.uyc = uyc;
  o.usqc = uxc * uxc + uyc * uyc + o.uzn * o.uzn;
  return o;
}

// post-stream population q of cell (i, j, k), every boundary branch (js/live/kernel.js:159-217)
template <int q>
__device__ __forceinline__ float pull19(const Grid &g, const float *__restrict__ fA, size_t N, uint32_t n, int i, int j,
                                        int k, const int (&kn)[3][3], float uI, float vI, const Outflow &ob) {
  constexpr int cx = CX[q], cy = CY[q], cz = CZ[q];
  const Params &P = g.P;
  const int NX = g.nx, NY = g.ny;
  if (cz == -1 && k == g.nz - 1) {
    // free-slip lid: the mirrored population of the lateral neighbour at the same level
    constexpr int m = MIRZ[q];
    const int si = i - cx, sj = j - cy;
    if (si < 0 || si >= NX || sj < 0 || sj >= NY) return fA[(size_t)m * N + n];
    if (k < kn[1 - cx][1 - cy]) return fA[(size_t)m * N + n];
    return fA[(size_t)m * N + (n - (uint32_t)(cx + cy * NX))];
  }
  const bool outx = (cx == 1 && i == 0) || (cx == -1 && i == NX - 1);
  const bool outy = (cy == 1 && j == 0) || (cy == -1 && j == NY - 1);
  if (outx || outy) {
    const bool inflow = (cx == 1 && i == 0 && P.inflow_w != 0) || (cx == -1 && i == NX - 1 && P.inflow_e != 0) ||
                        (cy == 1 && j == 0 && P.inflow_s != 0) || (cy == -1 && j == NY - 1 && P.inflow_n != 0);
    const float w = w19<q>();
    if (inflow) return w * epoly(cdot<cx, cy, 0>(uI, vI, 0.0f), uI * uI + vI * vI);   // equilibrium, rho = 1
    // pressure outlet: non-equilibrium extrapolation towards rho = 1, optionally backflow-clipped
    const float cu = cdot<cx, cy, cz>(ob.uxn, ob.uyn, ob.uzn);
    const float own = fA[(size_t)q * N + n];
    if (P.oclip != 0) {
      const float cuc = cdot<cx, cy, cz>(ob.uxc, ob.uyc, ob.uzn);
      return own + w * (epoly(cuc, ob.usqc) - ob.rn * epoly(cu, ob.usqn));
    }
    return own + w * (1.0f - ob.rn) * epoly(cu, ob.usqn);
  }
  // interior: half-way bounce-back when the source cell is solid
  const int sk = k - cz;
  if (sk < kn[1 - cx][1 - cy]) return fA[(size_t)OPP[q] * N + n];
  const int off = cx + cy * NX + cz * (int)g.NXY;
  return fA[(size_t)q * N + (uint32_t)((int)n - off)];
}

// post-stream theta population q (js/live/kernel.js:219-243)
template <int q>
__device__ __forceinline__ float pull7(const Grid &g, const float *__restrict__ gA, size_t N, uint32_t n, int i, int j,
                                       int k, const int (&kn)[3][3]) {
  constexpr int cx = CX7[q], cy = CY7[q], cz = CZ7[q];
  const Params &P = g.P;
  if (cz == -1 && k == g.nz - 1) return gA[(size_t)q * N + n];   // zero-gradient top
  const bool outx = (cx == 1 && i == 0) || (cx == -1 && i == g.nx - 1);
  const bool outy = (cy == 1 && j == 0) || (cy == -1 && j == g.ny - 1);
  if (outx || outy) {
    const bool inflow = (cx == 1 && i == 0 && P.inflow_w != 0) || (cx == -1 && i == g.nx - 1 && P.inflow_e != 0) ||
                        (cy == 1 && j == 0 && P.inflow_s != 0) || (cy == -1 && j == g.ny - 1 && P.inflow_n != 0);
    return inflow ? 0.0f : gA[(size_t)q * N + n];
  }
  const int sk = k - cz;
  if (sk < kn[1 - cx][1 - cy]) return gA[(size_t)OPP7[q] * N + n];
  const int off = cx + cy * g.nx + cz * (int)g.NXY;
  return gA[(size_t)q * N + (uint32_t)((int)n - off)];
}

template <int q>
struct Pull19 {
  __device__ __forceinline__ static void run(float (&f)[19], const Grid &g, const float *__restrict__ fA, size_t N,
                                             uint32_t n, int i, int j, int k, const int (&kn)[3][3], float uI, float vI,
                                             const Outflow &ob) {
    f[q] = pull19<q>(g, fA, N, n, i, j, k, kn, uI, vI, ob);
    Pull19<q + 1>::run(f, g, fA, N, n, i, j, k, kn, uI, vI, ob);
  }
};
template <>
struct Pull19<19> {
  __device__ __forceinline__ static void run(float (&)[19], const Grid &, const float *__restrict__, size_t, uint32_t,
                                             int, int, int, const int (&)[3][3], float, float, const Outflow &) {}
};

template <int q>
struct Pull7 {
  __device__ __forceinline__ static void run(float (&gg)[7], const Grid &g, const float *__restrict__ gA, size_t N,
                                             uint32_t n, int i, int j, int k, const int (&kn)[3][3]) {
    gg[q] = pull7<q>(g, gA, N, n, i, j, k, kn);
    Pull7<q + 1>::run(gg, g, gA, N, n, i, j, k, kn);
  }
};
template <>
struct Pull7<7> {
  __device__ __forceinline__ static void run(float (&)[7], const Grid &, const float *__restrict__, size_t, uint32_t,
                                             int, int, int, const int (&)[3][3]) {}
};

// fast path: all 26 sources are fluid interior cells; plain shifted loads through per-population descriptors
template <int q>
struct FastPull {
  __device__ __forceinline__ static void run(float (&f)[19], float (&gg)[7], const Grid &g,
                                             const float *__restrict__ fA, const float *__restrict__ gA, uint32_t boff) {
    const int off = CX[q] + CY[q] * g.nx + CZ[q] * (int)g.NXY;
    f[q] = ldb(shifted(fA, g.N, q, off), boff);
    if (q < 7) {
      const int o7 = CX7[q < 7 ? q : 0] + CY7[q < 7 ? q : 0] * g.nx + CZ7[q < 7 ? q : 0] * (int)g.NXY;
      gg[q < 7 ? q : 0] = ldb(shifted(gA, g.N, q, o7), boff);
    }
    FastPull<q + 1>::run(f, gg, g, fA, gA, boff);
  }
};
template <>
struct FastPull<19> {
  __device__ __forceinline__ static void run(float (&)[19], float (&)[7], const Grid &, const float *__restrict__,
                                             const float *__restrict__, uint32_t) {}
};

template <int q>
struct StorePairs {
  __device__ __forceinline__ static void run(float *__restrict__ fB, size_t N, uint32_t boff, const Moments &m,
                                             const Forces &F, const Relax &R) {
    float a, b;
    collide_pair<q>(m, F, R, a, b);
    stb(fB + (size_t)q * N, boff, a);
    stb(fB + (size_t)(q + 1) * N, boff, b);
    StorePairs<q + 2>::run(fB, N, boff, m, F, R);
  }
};
template <>
struct StorePairs<19> {
  __device__ __forceinline__ static void run(float *__restrict__, size_t, uint32_t, const Moments &, const Forces &,
                                             const Relax &) {}
};

__device__ __forceinline__ void load_neighbours(const Grid &g, int i, int j, int kg, int (&kn)[3][3]) {
  const int im = max(i - 1, 0), ip = min(i + 1, g.nx - 1), jm = max(j - 1, 0), jp = min(j + 1, g.ny - 1);
  const int xs[3] = {im, i, ip}, ys[3] = {jm, j, jp};
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) kn[a][b] = (a == 1 && b == 1) ? kg : (int)g.ktop[ys[b] * g.nx + xs[a]];
}

// streaming + moments + forces of one fluid cell: the fast or the boundary path
__device__ __forceinline__ void stream_cell(const Grid &g, const float *__restrict__ fA, const float *__restrict__ gA,
                                            uint32_t n, int i, int j, int k, int kg, int ktc, bool fast, float uI,
                                            float vI, float (&f)[19], float (&gg)[7]) {
  if (fast) {
    FastPull<0>::run(f, gg, g, fA, gA, n * 4u);
  } else {
    int kn[3][3];
    load_neighbours(g, i, j, kg, kn);
    const Outflow ob = outflow_prelude(g, fA, g.N, n, i, j);
    Pull19<0>::run(f, g, fA, g.N, n, i, j, k, kn, uI, vI, ob);
    Pull7<0>::run(gg, g, gA, g.N, n, i, j, k, kn);
  }
}

// the fused step (genStepWGSL, js/live/kernel.js:395-431) with FP64 accumulation (ACC) and debug outputs (OUT)
template <bool ACC, bool OUT>
__global__ __launch_bounds__(256) void step_f32(Grid g, const float *__restrict__ fA, float *__restrict__ fB,
                                                const float *__restrict__ gA, float *__restrict__ gB, Acc acc,
                                                float *__restrict__ out4) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = g.kmin + blockIdx.y;
  const bool live = c < (int)g.NXY;
  const uint4 cv = g.colv[live ? c : 0];
  const int ktc = col_kt(cv.x), kg = col_ktop(cv.x);
  const bool fluid = live && k >= kg;
  if (!ACC && !fluid) return;
  double th_cell = 0.0;
  if (fluid) {
    const size_t N = g.N;
    const uint32_t n = (uint32_t)k * g.NXY + (uint32_t)c;
    int i, j;
    split_column(g, c, i, j);
    const float4 pr = g.prof[max(k - ktc, 0)];
    const bool fast = k >= col_kfast(cv.x) && k < g.nz - 1 && i > 0 && i < g.nx - 1 && j > 0 && j < g.ny - 1;
    float f[19], gg[7];
    stream_cell(g, fA, gA, n, i, j, k, kg, ktc, fast, pr.x, pr.y, f, gg);
    const Moments m = moments(f, gg);
    const Forces F = forces<true, true>(g.P, g.nz, k, kg, ktc, __uint_as_float(cv.z), pr.x, pr.y, m);
    const Relax R = relax(g.P, g.nx, g.ny, i, j, pr.z, m, F);
    const uint32_t boff = n * 4u;
    stb(fB, boff, collide_rest(R));
    StorePairs<1>::run(fB, N, boff, m, F, R);
    // D3Q7: g^post = a g + b (1 + 4 c.u) + w S (1 - sg)
    const float S = (k == kg) ? __uint_as_float(cv.w) * g.P.srcscale : 0.0f;
    const Scalar sc = scalar_coeffs(m.th, R.omg, S, F.sg);
    const float b1 = sc.b + sc.cs1, b4 = 4.0f * sc.b;
    stb(gB, boff, sc.a * gg[0] + sc.cs0);
    stb(gB + N, boff, sc.a * gg[1] + (b1 + b4 * F.ux));
    stb(gB + 2 * N, boff, sc.a * gg[2] + (b1 - b4 * F.ux));
    stb(gB + 3 * N, boff, sc.a * gg[3] + (b1 + b4 * F.uy));
    stb(gB + 4 * N, boff, sc.a * gg[4] + (b1 - b4 * F.uy));
    stb(gB + 5 * N, boff, sc.a * gg[5] + (b1 + b4 * F.uz));
    stb(gB + 6 * N, boff, sc.a * gg[6] + (b1 - b4 * F.uz));
    if (OUT) {
      out4[n] = m.th;
      out4[N + n] = F.ux;
      out4[2 * N + n] = F.uy;
      out4[3 * N + n] = F.uz;
    }
    if (ACC) {
      th_cell = (double)m.th;
      if (k == kg) {
        acc.surf[c] += (double)m.th;
        acc.surf[g.NXY + c] += (double)F.ux;
        acc.surf[2 * (size_t)g.NXY + c] += (double)F.uy;
      }
      const int nx = g.nx, ny = g.ny, nz = g.nz;
      const size_t lw = (size_t)nz * ny, ls = (size_t)nz * nx;
      if (i == 0) {
        double *fw = acc.face;
        fw[(size_t)k * ny + j] += (double)m.th;
        fw[lw + (size_t)k * ny + j] += (double)F.ux;
      }
      if (i == nx - 1) {
        double *fe = acc.face + 2 * lw;
        fe[(size_t)k * ny + j] += (double)m.th;
        fe[lw + (size_t)k * ny + j] += (double)F.ux;
      }
      if (j == 0) {
        double *fs = acc.face + 4 * lw;
        fs[(size_t)k * nx + i] += (double)m.th;
        fs[ls + (size_t)k * nx + i] += (double)F.uy;
      }
      if (j == ny - 1) {
        double *fn = acc.face + 4 * lw + 2 * ls;
        fn[(size_t)k * nx + i] += (double)m.th;
        fn[ls + (size_t)k * nx + i] += (double)F.uy;
      }
    }
  }
  if (ACC) {
    // heat content: one slot per wavefront, each updated by that wavefront only (deterministic, no atomics)
    const double s = wave_sum(th_cell);
    if ((threadIdx.x & 63) == 0) {
      const uint32_t slot = ((uint32_t)blockIdx.y * gridDim.x + blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
      acc.content[slot] += s;
    }
  }
}

// near-surface read (genSurfWGSL, js/live/kernel.js:435-471): out3 = [3][NXY] north-up theta, u*conv, v*conv
__global__ __launch_bounds__(256) void surface_f32(Grid g, const float *__restrict__ fA, const float *__restrict__ gA,
                                                   float conv, float *__restrict__ out3) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int)g.NXY) return;
  int i, j;
  split_column(g, c, i, j);
  const uint4 cv = g.colv[c];
  const int ktc = col_kt(cv.x), kg = col_ktop(cv.x), k = kg;
  const uint32_t o = (uint32_t)(g.ny - 1 - j) * g.nx + i;
  if (k >= g.nz) {
    out3[o] = 0.0f; out3[g.NXY + o] = 0.0f; out3[2 * g.NXY + o] = 0.0f;
    return;
  }
  const size_t N = g.N;
  const uint32_t n = (uint32_t)k * g.NXY + (uint32_t)c;
  const int zagl = max(k - ktc, 0);
  const float4 pr = g.prof[zagl];
  int kn[3][3];
  load_neighbours(g, i, j, kg, kn);
  const Outflow ob = outflow_prelude(g, fA, N, n, i, j);
  float f[19], gg[7];
  Pull19<0>::run(f, g, fA, N, n, i, j, k, kn, pr.x, pr.y, ob);
  Pull7<0>::run(gg, g, gA, N, n, i, j, k, kn);
  const Moments m = moments(f, gg);
  const Forces F = forces<true, true>(g.P, g.nz, k, kg, ktc, __uint_as_float(cv.z), pr.x, pr.y, m);
  out3[o] = m.th;
  out3[g.NXY + o] = F.ux * conv;
  out3[2 * g.NXY + o] = F.uy * conv;
}

// equilibrium of the inflow profile at the local height above terrain, both lattices (genInitWGSL :524-553)
// evaluated operation by operation as the reference (no contraction), so the state is bit-identical to the closed form
__global__ __launch_bounds__(256) void init_f32(Grid g, float *__restrict__ f0, float *__restrict__ f1) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (c >= (int)g.NXY) return;
  const int ktc = col_kt(g.colv[c].x);
  const int zagl = min(max(k - ktc, 0), g.nz - 1);
  const float4 pr = g.prof[zagl];
  const float U = pr.x, V = pr.y, usq = U * U + V * V;
  const size_t N = g.N;
  const uint32_t n = (uint32_t)k * g.NXY + (uint32_t)c;
#pragma unroll
  for (int q = 0; q < 19; q++) {
    const float cu = (float)CX[q] * U + (float)CY[q] * V;
    const float w = q == 0 ? (float)(1.0 / 3.0) : (q < 7 ? (float)(1.0 / 18.0) : (float)(1.0 / 36.0));
    const float fq = w * (((1.0f + 3.0f * cu) + 4.5f * cu * cu) - 1.5f * usq);
    f0[(size_t)q * N + n] = fq;
    f1[(size_t)q * N + n] = fq;
  }
}

// cells that turned from solid to fluid: rest state f = w_q, g = 0 (genReinitWGSL :557-575)
__global__ __launch_bounds__(256) void reinit_f32(Grid g, const uint32_t *__restrict__ oldk, float *__restrict__ f,
                                                  float *__restrict__ gq) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int)g.NXY) return;
  const int k1 = min((int)oldk[c], g.nz);
  const size_t N = g.N;
  for (int k = (int)g.ktop[c]; k < k1; k++) {
    const uint32_t n = (uint32_t)k * g.NXY + (uint32_t)c;
#pragma unroll
    for (int q = 0; q < 19; q++)
      f[(size_t)q * N + n] = q == 0 ? (float)(1.0 / 3.0) : (q < 7 ? (float)(1.0 / 18.0) : (float)(1.0 / 36.0));
#pragma unroll
    for (int q = 0; q < 7; q++) gq[(size_t)q * N + n] = 0.0f;
  }
}

// u8 theta volume, north-up [k][row][col], 4 cells per u32, and max theta (genVolWGSL :475-520)
__global__ __launch_bounds__(256) void volume_f32(Grid g, const float *__restrict__ gA, int nzv, float volscale,
                                                  uint32_t *__restrict__ vout, uint32_t *__restrict__ tmax_bits) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nvol = nzv * (int)g.NXY;
  const int nw = (nvol + 3) / 4;
  float tmax = 0.0f;
  if (t < nw) {
    uint32_t packed = 0u;
    for (int e = 0; e < 4; e++) {
      const int p = 4 * t + e;
      if (p < nvol) {
        const int k = p / (int)g.NXY, rem = p % (int)g.NXY;
        const int i = rem % g.nx, j = g.ny - 1 - rem / g.nx;
        const int c = j * g.nx + i;
        const int kg = (int)g.ktop[c];
        float th = 0.0f;
        if (k >= kg) {
          int kn[3][3];
          load_neighbours(g, i, j, kg, kn);
          float gg[7];
          Pull7<0>::run(gg, g, gA, g.N, (uint32_t)k * g.NXY + (uint32_t)c, i, j, k, kn);
          th = gg[0] + gg[1] + gg[2] + gg[3] + gg[4] + gg[5] + gg[6];
        }
        const uint32_t qv = (uint32_t)fminf(fmaxf(rintf(th * volscale), 0.0f), 255.0f);
        packed |= qv << (8u * (uint32_t)e);
        tmax = fmaxf(tmax, th);
      }
    }
    vout[t] = packed;
  }
  // max over the block, then one atomic per block (max is order-independent, so the result is deterministic)
  __shared__ float red[256];
  red[threadIdx.x] = fmaxf(tmax, 0.0f);
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(tmax_bits, __float_as_uint(red[0]));
}

// diagnostics over the stored state (genStatsWGSL :580-633), FP64 per-block partials:
// part[8*b + (0..6)] = sum theta, sum ke, sum (rho-1), fluid count, max theta, max |u|, non-finite count
__global__ __launch_bounds__(256) void stats_f32(Grid g, const float *__restrict__ fA, const float *__restrict__ gA,
                                                 double *__restrict__ part) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  if (n < g.N) {
    const uint32_t c = n % g.NXY, k = n / g.NXY;
    if ((int)k >= (int)g.ktop[c]) {
      const size_t N = g.N;
      float a[19];
#pragma unroll
      for (int q = 0; q < 19; q++) a[q] = fA[(size_t)q * N + n];
      float rho = a[0];
#pragma unroll
      for (int q = 1; q < 19; q++) rho = rho + a[q];
      float jx = 0.f, jy = 0.f, jz = 0.f;
#pragma unroll
      for (int q = 0; q < 19; q++) {
        if (CX[q] == 1) jx = jx + a[q]; else if (CX[q] == -1) jx = jx - a[q];
        if (CY[q] == 1) jy = jy + a[q]; else if (CY[q] == -1) jy = jy - a[q];
        if (CZ[q] == 1) jz = jz + a[q]; else if (CZ[q] == -1) jz = jz - a[q];
      }
      float th = gA[n];
#pragma unroll
      for (int q = 1; q < 7; q++) th = th + gA[(size_t)q * N + n];
      const uint32_t br = __float_as_uint(rho) & 0x7f800000u, bt = __float_as_uint(th) & 0x7f800000u;
      if (br == 0x7f800000u || bt == 0x7f800000u) {
        v[6] = 1.0;
      } else {
        const float jj = jx * jx + jy * jy + jz * jz;
        v[0] = th;
        v[1] = 0.5f * jj / rho;
        v[2] = rho - 1.0f;
        v[3] = 1.0;
        v[4] = fmaxf(th, 0.0f);
        v[5] = sqrtf(jj) / rho;
      }
    }
  }
  __shared__ double red[7][4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int s = 0; s < 7; s++) {
    double x = v[s];
    if (s == 4 || s == 5) {
      for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
    } else {
      x = wave_sum(x);
    }
    if (lane == 0) red[s][wv] = x;
  }
  __syncthreads();
  if (threadIdx.x < 7) {
    const int s = threadIdx.x;
    const int nwv = blockDim.x >> 6;
    double x = red[s][0];
    for (int w = 1; w < nwv; w++) x = (s == 4 || s == 5) ? fmax(x, red[s][w]) : x + red[s][w];
    part[8 * (size_t)blockIdx.x + s] = x;
  }
}

// one block: out8[s] = fixed-order reduction of part[8*b + s] over nb blocks
__global__ __launch_bounds__(256) void stats_finish(const double *__restrict__ part, uint32_t nb, double *__restrict__ out8) {
  __shared__ double red[256];
  for (int s = 0; s < 7; s++) {
    const bool mx = (s == 4 || s == 5);
    double x = 0.0;
    for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
      const double y = part[8 * (size_t)b + s];
      x = mx ? fmax(x, y) : x + y;
    }
    red[threadIdx.x] = x;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) red[threadIdx.x] = mx ? fmax(red[threadIdx.x], red[threadIdx.x + h]) : red[threadIdx.x] + red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) out8[s] = red[0];
    __syncthreads();
  }
}

// one block: *out = fixed-order sum of x[0..n)
__global__ __launch_bounds__(256) void sum_f64(const double *__restrict__ x, uint32_t n, double *__restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (uint32_t b = threadIdx.x; b < n; b += blockDim.x) s += x[b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// ---- launchers ------------------------------------------------------------------------------------------------
size_t face_len(int face, int nx, int ny, int nz) { return face < 2 ? 2 * (size_t)nz * ny : 2 * (size_t)nz * nx; }
size_t face_offset(int face, int nx, int ny, int nz) {
  size_t o = 0;
  for (int f = 0; f < face; f++) o += face_len(f, nx, ny, nz);
  return o;
}

uint32_t content_slots(const Grid &g, int block) {
  const uint32_t bx = (g.NXY + block - 1) / block;
  return bx * (uint32_t)(g.nz - g.kmin) * (uint32_t)(block / 64);
}

hipError_t launch_step(const Grid &g, const float *fA, float *fB, const float *gA, float *gB, const Acc *acc,
                       float *out4, int block, hipStream_t st) {
  if (g.kmin >= g.nz) return hipSuccess;
  const dim3 grid((g.NXY + block - 1) / block, g.nz - g.kmin);
  const Acc a = acc ? *acc : Acc{nullptr, nullptr, nullptr};
  if (acc && out4) hipLaunchKernelGGL((step_f32<true, true>), grid, dim3(block), 0, st, g, fA, fB, gA, gB, a, out4);
  else if (acc) hipLaunchKernelGGL((step_f32<true, false>), grid, dim3(block), 0, st, g, fA, fB, gA, gB, a, out4);
  else if (out4) hipLaunchKernelGGL((step_f32<false, true>), grid, dim3(block), 0, st, g, fA, fB, gA, gB, a, out4);
  else hipLaunchKernelGGL((step_f32<false, false>), grid, dim3(block), 0, st, g, fA, fB, gA, gB, a, out4);
  return hipGetLastError();
}

hipError_t launch_init(const Grid &g, float *f0, float *f1, hipStream_t st) {
  const dim3 grid((g.NXY + 255) / 256, g.nz);
  hipLaunchKernelGGL(init_f32, grid, dim3(256), 0, st, g, f0, f1);
  return hipGetLastError();
}

hipError_t launch_reinit(const Grid &g, const uint32_t *old_ktop, float *f, float *gq, hipStream_t st) {
  hipLaunchKernelGGL(reinit_f32, dim3((g.NXY + 255) / 256), dim3(256), 0, st, g, old_ktop, f, gq);
  return hipGetLastError();
}

hipError_t launch_surface(const Grid &g, const float *fA, const float *gA, float conv, float *out3, hipStream_t st) {
  hipLaunchKernelGGL(surface_f32, dim3((g.NXY + 255) / 256), dim3(256), 0, st, g, fA, gA, conv, out3);
  return hipGetLastError();
}

hipError_t launch_volume(const Grid &g, const float *gA, int nzv, float volscale, uint32_t *vout, uint32_t *tmax_bits,
                         hipStream_t st) {
  const uint32_t nw = ((uint32_t)nzv * g.NXY + 3) / 4;
  hipLaunchKernelGGL(volume_f32, dim3((nw + 255) / 256), dim3(256), 0, st, g, gA, nzv, volscale, vout, tmax_bits);
  return hipGetLastError();
}

uint32_t stats_partials(const Grid &g) { return (g.N + 255) / 256; }

hipError_t launch_stats(const Grid &g, const float *fA, const float *gA, double *part, double *out8, hipStream_t st) {
  const uint32_t nb = stats_partials(g);
  hipLaunchKernelGGL(stats_f32, dim3(nb), dim3(256), 0, st, g, fA, gA, part);
  hipLaunchKernelGGL(stats_finish, dim3(1), dim3(256), 0, st, part, nb, out8);
  return hipGetLastError();
}

hipError_t launch_sum(const double *x, uint32_t n, double *out, hipStream_t st) {
  hipLaunchKernelGGL(sum_f64, dim3(1), dim3(256), 0, st, x, n, out);
  return hipGetLastError();
}

}  // namespace aht
