// Device-side lattice tables and per-cell physics of the coupled D3Q19 + D3Q7 step (gfx950).
//
// The arithmetic follows the reference kernel generator js/live/kernel.js section by section (stream :159-243,
// moments :264-277, forces :279-306, relaxation :308-338, collision :340-386) in the same operation order, so that
// with FMA contraction off the f32 results equal the CPU oracle's. Layout: populations [q][cell],
// cell = (k*ny + j)*nx + i (js/live/kernel.js:17).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aht {

// D3Q19 / D3Q7 velocity sets, weights, opposite and z-mirror maps (js/live/kernel.js:19-28)
__device__ constexpr int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
__device__ constexpr int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
__device__ constexpr int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
__device__ constexpr int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
__device__ constexpr int MIRZ[19] = {0, 1, 2, 3, 4, 6, 5, 7, 8, 9, 10, 13, 14, 11, 12, 17, 18, 15, 16};
__device__ constexpr int OPP7[7] = {0, 2, 1, 4, 3, 6, 5};

template <int q> __device__ __forceinline__ constexpr float w19() {
  return q == 0 ? (float)(1.0 / 3.0) : (q < 7 ? (float)(1.0 / 18.0) : (float)(1.0 / 36.0));
}
template <int q> __device__ __forceinline__ constexpr float w7() { return q == 0 ? 0.25f : 0.125f; }
// fmt(4.5 * w_q): the regularisation prefactor, exact in f32 (js/live/kernel.js:361)
template <int q> __device__ __forceinline__ constexpr float w45() { return q == 0 ? 1.5f : (q < 7 ? 0.25f : 0.125f); }

// kernel parameters: the reference's uniform Params (js/live/kernel.js:44-50) plus grid and buffers
struct Params {
  float tau0, cs2smag, gbeta, prt, kfloor, momfac, srcscale, sigmax;
  int inflow_w, inflow_e, inflow_s, inflow_n, nsp, nsw, oclip, pad;
};

// equilibrium polynomial 1 + 3cu + 4.5cu^2 - 1.5u^2 in the generator's order ((1 + 3cu) + (4.5cu)cu) - 1.5usq
__device__ __forceinline__ float epoly(float cu, float usq) {
  return ((1.0f + 3.0f * cu) + 4.5f * cu * cu) - 1.5f * usq;
}

// c.u with only the nonzero components, x then y then z (the generator's term list)
template <int cx, int cy, int cz>
__device__ __forceinline__ float cdot(float x, float y, float z) {
  float s = 0.0f;
  if (cx) s = s + (float)cx * x;
  if (cy) s = s + (float)cy * y;
  if (cz) s = s + (float)cz * z;
  return s;
}

// post-stream moments: rho, j, theta and the second moments (js/live/kernel.js:264-277)
struct Moments {
  float rho, jx, jy, jz, th, Pxx, Pyy, Pzz, Pxy, Pxz, Pyz;
};

template <bool SECOND>
__device__ __forceinline__ Moments moments(const float (&f)[19], const float (&g)[7]) {
  Moments m;
  float r = f[0];
#pragma unroll
  for (int q = 1; q < 19; q++) r = r + f[q];
  m.rho = r;
  float sx = 0.f, sy = 0.f, sz = 0.f;
#pragma unroll
  for (int q = 0; q < 19; q++) {
    if (CX[q] == 1) sx = sx + f[q]; else if (CX[q] == -1) sx = sx - f[q];
    if (CY[q] == 1) sy = sy + f[q]; else if (CY[q] == -1) sy = sy - f[q];
    if (CZ[q] == 1) sz = sz + f[q]; else if (CZ[q] == -1) sz = sz - f[q];
  }
  m.jx = sx; m.jy = sy; m.jz = sz;
  float t = g[0];
#pragma unroll
  for (int q = 1; q < 7; q++) t = t + g[q];
  m.th = t;
  if (SECOND) {
    float xx = 0.f, yy = 0.f, zz = 0.f, xy = 0.f, xz = 0.f, yz = 0.f;
#pragma unroll
    for (int q = 0; q < 19; q++) {
      if (CX[q] != 0) xx = xx + f[q];
      if (CY[q] != 0) yy = yy + f[q];
      if (CZ[q] != 0) zz = zz + f[q];
      const int vxy = CX[q] * CY[q], vxz = CX[q] * CZ[q], vyz = CY[q] * CZ[q];
      if (vxy == 1) xy = xy + f[q]; else if (vxy == -1) xy = xy - f[q];
      if (vxz == 1) xz = xz + f[q]; else if (vxz == -1) xz = xz - f[q];
      if (vyz == 1) yz = yz + f[q]; else if (vyz == -1) yz = yz - f[q];
    }
    m.Pxx = xx; m.Pyy = yy; m.Pzz = zz; m.Pxy = xy; m.Pxz = xz; m.Pyz = yz;
  }
  return m;
}

struct Forces {
  float Fx, Fy, Fz, sg, ux, uy, uz;
};

// buoyancy, limited log-law drag, top sponge, force-corrected velocity (js/live/kernel.js:279-306)
__device__ __forceinline__ Forces forces(const Params &P, int NZ, int k, int kg, int ktc, float cd, float uI, float vI,
                                         const Moments &m) {
  Forces o;
  const float rho = m.rho;
  float Fx = 0.0f, Fy = 0.0f;
  float Fz = rho * P.gbeta * m.th;
  if (k == kg && kg == ktc) {
    const float ux0 = m.jx / rho, uy0 = m.jy / rho;
    const float sp = sqrtf(ux0 * ux0 + uy0 * uy0);
    Fx = -rho * cd * sp * ux0;
    Fy = -rho * cd * sp * uy0;
    const float lim = 0.5f * rho;
    if (cd * sp > 0.5f) {
      Fx = -lim * ux0;
      Fy = -lim * uy0;
    }
  }
  float sg = 0.0f;
  if (k >= NZ - P.nsp) {
    const float aa = ((float)(k - (NZ - P.nsp)) + 1.0f) / (float)P.nsp;
    sg = P.sigmax * aa * aa;
    Fx += rho * sg * (uI - m.jx / rho);
    Fy += rho * sg * (vI - m.jy / rho);
    Fz += -rho * sg * (m.jz / rho);
  }
  o.Fx = Fx; o.Fy = Fy; o.Fz = Fz; o.sg = sg;
  o.ux = (m.jx + 0.5f * Fx) / rho;
  o.uy = (m.jy + 0.5f * Fy) / rho;
  o.uz = (m.jz + 0.5f * Fz) / rho;
  return o;
}

// relaxation times (js/live/kernel.js:308-338)
struct Relax {
  float om, fpre, taug, trN, Nxx, Nyy, Nzz, Nxy, Nxz, Nyz, usq;
};

__device__ __forceinline__ Relax relax(const Params &P, int NX, int NY, int i, int j, float kb, const Moments &m,
                                       const Forces &F) {
  Relax r;
  const float rho = m.rho, ux = F.ux, uy = F.uy, uz = F.uz;
  r.usq = ux * ux + uy * uy + uz * uz;
  const float third = (float)(1.0 / 3.0);
  const float r3 = rho * third;
  r.Nxx = m.Pxx - r3 - rho * ux * ux;
  r.Nyy = m.Pyy - r3 - rho * uy * uy;
  r.Nzz = m.Pzz - r3 - rho * uz * uz;
  r.Nxy = m.Pxy - rho * ux * uy;
  r.Nxz = m.Pxz - rho * ux * uz;
  r.Nyz = m.Pyz - rho * uy * uz;
  const float Pn = sqrtf(r.Nxx * r.Nxx + r.Nyy * r.Nyy + r.Nzz * r.Nzz +
                         2.0f * (r.Nxy * r.Nxy + r.Nxz * r.Nxz + r.Nyz * r.Nyz));
  float tb = P.tau0 + 3.0f * P.momfac * kb;
  int dmin = 1000;
  if (P.inflow_w == 0 && i < dmin) dmin = i;
  if (P.inflow_e == 0 && NX - 1 - i < dmin) dmin = NX - 1 - i;
  if (P.inflow_s == 0 && j < dmin) dmin = j;
  if (P.inflow_n == 0 && NY - 1 - j < dmin) dmin = NY - 1 - j;
  if (dmin < P.nsw) {
    const float rr = ((float)P.nsw - (float)dmin) / (float)P.nsw;
    tb += 0.15f * rr * rr;
  }
  const float tau = 0.5f * (tb + sqrtf(tb * tb + P.cs2smag * Pn / rho));
  const float nusgs = (tau - tb) * third;
  const float kap = nusgs / P.prt + kb + P.kfloor;
  r.taug = 0.5f + 4.0f * kap;
  r.om = 1.0f - 1.0f / tau;
  r.fpre = 1.0f - 0.5f / tau;
  r.trN = (r.Nxx + r.Nyy + r.Nzz) * third;
  return r;
}

// regularised collision with Guo forcing of population q (js/live/kernel.js:340-367)
template <int q>
__device__ __forceinline__ float collide19(const Moments &m, const Forces &F, const Relax &R) {
  constexpr int cx = CX[q], cy = CY[q], cz = CZ[q];
  const float w = w19<q>();
  const float cu = cdot<cx, cy, cz>(F.ux, F.uy, F.uz);
  const float feq = w * m.rho * epoly(cu, R.usq);
  // (c c):N in the generator's term order: diagonals xx, yy, zz, then cross xy, xz, yz
  float qn = 0.0f;
  bool first = true;
  if (cx != 0) { qn = first ? R.Nxx : qn + R.Nxx; first = false; }
  if (cy != 0) { qn = first ? R.Nyy : qn + R.Nyy; first = false; }
  if (cz != 0) { qn = first ? R.Nzz : qn + R.Nzz; first = false; }
  if (cx != 0 && cy != 0) { const float t = (float)(2 * cx * cy) * R.Nxy; qn = first ? t : qn + t; first = false; }
  if (cx != 0 && cz != 0) { const float t = (float)(2 * cx * cz) * R.Nxz; qn = first ? t : qn + t; first = false; }
  if (cy != 0 && cz != 0) { const float t = (float)(2 * cy * cz) * R.Nyz; qn = first ? t : qn + t; first = false; }
  const float fr = w45<q>() * (qn - R.trN);
  const float cF = cdot<cx, cy, cz>(F.Fx, F.Fy, F.Fz);
  const float Fq = R.fpre * w * (3.0f * (cF - (F.ux * F.Fx + F.uy * F.Fy + F.uz * F.Fz)) + 9.0f * cu * cF);
  return feq + R.om * fr + Fq;
}

// advection-diffusion collision of theta population q with the rooftop source and top sponge (:369-386)
template <int q>
__device__ __forceinline__ float collide7(float gq, float th, float omg, float S, float sg, const Forces &F) {
  constexpr int cx = (q == 1) - (q == 2), cy = (q == 3) - (q == 4), cz = (q == 5) - (q == 6);
  const float w = w7<q>();
  const float geq = w * th * (1.0f + 4.0f * cdot<cx, cy, cz>(F.ux, F.uy, F.uz));
  return (gq + omg * (geq - gq) + w * S) * (1.0f - sg);
}

}  // namespace aht
