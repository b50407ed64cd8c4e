// CDNA4 (gfx950) kernels of the coupled D3Q19 + D3Q7 lattice-Boltzmann step, f32 two-lattice pull scheme.
//
// Thread mapping: one thread per cell. The x-y plane is flattened into columns c = j*nx + i so that the 64 lanes of a
// wavefront touch 64 consecutive cells of every [q][cell] stream (256 contiguous bytes; the stores and the
// unshifted loads are 256-byte aligned whenever nx*ny is a multiple of 64). blockIdx.y is the level k, starting at the
// lowest fluid level of the grid, so all-solid planes are never launched. Solid cells above that plane exit at once.
//
// Boundary logic follows js/live/kernel.js:159-243: per population, the lid branch first, then the lateral faces,
// then the interior pull with half-way bounce-back from solid sources. The 3 x 3 neighbour ktop values (the cacheNbr
// path of the reference, js/live/kernel.js:102-110) are loaded once per thread.
#include "kernels.h"

namespace aht {

struct Outflow {
  float rn, uxn, uyn, uzn, usqn, uxc, uyc, usqc;
};

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
    float r = a[0];
#pragma unroll
    for (int q = 1; q < 19; q++) r = r + a[q];
    float sx = 0.f, sy = 0.f, sz = 0.f;
#pragma unroll
    for (int q = 0; q < 19; q++) {
      if (CX[q] == 1) sx = sx + a[q]; else if (CX[q] == -1) sx = sx - a[q];
      if (CY[q] == 1) sy = sy + a[q]; else if (CY[q] == -1) sy = sy - a[q];
      if (CZ[q] == 1) sz = sz + a[q]; else if (CZ[q] == -1) sz = sz - a[q];
    }
    o.rn = r; o.uxn = sx / r; o.uyn = sy / r; o.uzn = sz / r;
  }
  o.usqn = o.uxn * o.uxn + o.uyn * o.uyn + o.uzn * o.uzn;
  float uxc = o.uxn, uyc = o.uyn;
  if (P.oclip != 0) {
    if (i == 0 && P.inflow_w == 0 && uxc > 0.0f) uxc = 0.0f;
    if (i == g.nx - 1 && P.inflow_e == 0 && uxc < 0.0f) uxc = 0.0f;
    if (j == 0 && P.inflow_s == 0 && uyc > 0.0f) uyc = 0.0f;
    if (j == g.ny - 1 && P.inflow_n == 0 && uyc < 0.0f) uyc = 0.0f;
  }
  o.uxc = uxc; o.uyc = uyc;
  o.usqc = uxc * uxc + uyc * uyc + o.uzn * o.uzn;
  return o;
}

// post-stream population q of cell (i, j, k) (js/live/kernel.js:159-217)
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
    if (inflow) {
      // inflow face: equilibrium of rho = 1, U = (uI, vI, 0)
      float cu = 0.0f;
      if (cx) cu = cu + (float)cx * uI;
      if (cy) cu = cu + (float)cy * vI;
      return w * epoly(cu, uI * uI + vI * vI);
    }
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
  constexpr int cx = (q == 1) - (q == 2), cy = (q == 3) - (q == 4), cz = (q == 5) - (q == 6);
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

template <int q>
struct Collide19 {
  __device__ __forceinline__ static void run(float *__restrict__ fB, size_t N, uint32_t n, const Moments &m,
                                             const Forces &F, const Relax &R) {
    fB[(size_t)q * N + n] = collide19<q>(m, F, R);
    Collide19<q + 1>::run(fB, N, n, m, F, R);
  }
};
template <>
struct Collide19<19> {
  __device__ __forceinline__ static void run(float *__restrict__, size_t, uint32_t, const Moments &, const Forces &,
                                             const Relax &) {}
};

template <int q>
struct Collide7 {
  __device__ __forceinline__ static void run(float *__restrict__ gB, size_t N, uint32_t n, const float (&gg)[7], float th,
                                             float omg, float S, float sg, const Forces &F) {
    gB[(size_t)q * N + n] = collide7<q>(gg[q], th, omg, S, sg, F);
    Collide7<q + 1>::run(gB, N, n, gg, th, omg, S, sg, F);
  }
};
template <>
struct Collide7<7> {
  __device__ __forceinline__ static void run(float *__restrict__, size_t, uint32_t, const float (&)[7], float, float,
                                             float, float, const Forces &) {}
};

__device__ __forceinline__ void load_neighbours(const Grid &g, int i, int j, int kg, int (&kn)[3][3]) {
  const int im = max(i - 1, 0), ip = min(i + 1, g.nx - 1), jm = max(j - 1, 0), jp = min(j + 1, g.ny - 1);
  const int xs[3] = {im, i, ip}, ys[3] = {jm, j, jp};
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) kn[a][b] = (a == 1 && b == 1) ? kg : (int)g.ktop[ys[b] * g.nx + xs[a]];
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
  const int ktc = (int)cv.x, kg = (int)cv.y;
  const bool fluid = live && k >= kg;
  if (!ACC && !fluid) return;
  double th_cell = 0.0;
  if (fluid) {
    const size_t N = g.N;
    const uint32_t n = (uint32_t)k * g.NXY + (uint32_t)c;
    int i, j;
    split_column(g, c, i, j);
    const int zagl = max(k - ktc, 0);
    const float4 pr = g.prof[zagl];
    const float uI = pr.x, vI = pr.y, kb = pr.z;
    int kn[3][3];
    load_neighbours(g, i, j, kg, kn);
    const Outflow ob = outflow_prelude(g, fA, N, n, i, j);
    float f[19], gg[7];
    Pull19<0>::run(f, g, fA, N, n, i, j, k, kn, uI, vI, ob);
    Pull7<0>::run(gg, g, gA, N, n, i, j, k, kn);
    const Moments m = moments<true>(f, gg);
    const Forces F = forces(g.P, g.nz, k, kg, ktc, __uint_as_float(cv.z), uI, vI, m);
    const Relax R = relax(g.P, g.nx, g.ny, i, j, kb, m, F);
    Collide19<0>::run(fB, N, n, m, F, R);
    const float S = (k == kg) ? __uint_as_float(cv.w) * g.P.srcscale : 0.0f;
    const float omg = 1.0f / R.taug;
    Collide7<0>::run(gB, N, n, gg, m.th, omg, S, F.sg, F);
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
  const int ktc = (int)cv.x, kg = (int)cv.y, k = kg;
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
  const Moments m = moments<false>(f, gg);
  const Forces F = forces(g.P, g.nz, k, kg, ktc, __uint_as_float(cv.z), pr.x, pr.y, m);
  out3[o] = m.th;
  out3[g.NXY + o] = F.ux * conv;
  out3[2 * g.NXY + o] = F.uy * conv;
}

// equilibrium of the inflow profile at the local height above terrain, both lattices (genInitWGSL :524-553)
__global__ __launch_bounds__(256) void init_f32(Grid g, float *__restrict__ f0, float *__restrict__ f1) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (c >= (int)g.NXY) return;
  const int ktc = (int)g.colv[c].x;
  const int zagl = min(max(k - ktc, 0), g.nz - 1);
  const float4 pr = g.prof[zagl];
  const float U = pr.x, V = pr.y, usq = U * U + V * V;
  const size_t N = g.N;
  const uint32_t n = (uint32_t)k * g.NXY + (uint32_t)c;
#pragma unroll
  for (int q = 0; q < 19; q++) {
    const float cu = (float)CX[q] * U + (float)CY[q] * V;
    const float w = q == 0 ? (float)(1.0 / 3.0) : (q < 7 ? (float)(1.0 / 18.0) : (float)(1.0 / 36.0));
    const float fq = w * epoly(cu, usq);
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
