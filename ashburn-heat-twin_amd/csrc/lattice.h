// Device-side lattice tables and per-cell physics of the coupled D3Q19 + D3Q7 step (gfx950).
//
// The physics is the reference kernel generator's (js/live/kernel.js): moments :264-277, buoyancy / log-law drag /
// top sponge :279-306, Smagorinsky + K-profile relaxation with the outflow viscous sponge :308-338, regularised BGK
// with Guo forcing :340-367, D3Q7 advection-diffusion with rooftop source and sponge damping :369-386.
// The arithmetic is arranged for the CDNA4 VALU instead of the generator's term-by-term order:
//   * opposite populations (q, opp(q)) are handled as pairs: the moments use their sums and differences, and the
//     collision evaluates the part even in c once and the odd part once, then writes both populations;
//   * 1/rho, 1/tau and 1/tau_g are each computed once (v_rcp_f32) and multiplied, square roots use v_sqrt_f32;
//   * the file is compiled with FMA contraction.
// The result equals the reference expressions to f32 rounding (checked against the CPU oracle, which keeps the
// generator's order, to relative L2 <= 1e-6 per step). Layout: populations [q][cell],
// cell = (k*ny + j)*nx + i (js/live/kernel.js:17).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aht {

// D3Q19 / D3Q7 velocity sets, opposite and z-mirror maps (js/live/kernel.js:19-28). Pairs (q, q+1) for odd q are
// opposite directions: 1/2 x, 3/4 y, 5/6 z, 7/8 (1,1,0), 9/10 (1,-1,0), 11/12 (1,0,1), 13/14 (1,0,-1), 15/16 (0,1,1),
// 17/18 (0,1,-1).
__device__ constexpr int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
__device__ constexpr int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
__device__ constexpr int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
__device__ constexpr int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
__device__ constexpr int MIRZ[19] = {0, 1, 2, 3, 4, 6, 5, 7, 8, 9, 10, 13, 14, 11, 12, 17, 18, 15, 16};
__device__ constexpr int OPP7[7] = {0, 2, 1, 4, 3, 6, 5};
__device__ constexpr int CX7[7] = {0, 1, -1, 0, 0, 0, 0};
__device__ constexpr int CY7[7] = {0, 0, 0, 1, -1, 0, 0};
__device__ constexpr int CZ7[7] = {0, 0, 0, 0, 0, 1, -1};

// weights as the generator prints them (fmt(): f32-rounded, js/live/kernel.js:31-33)
template <int q> __device__ __forceinline__ constexpr float w19() {
  return q == 0 ? (float)(1.0 / 3.0) : (q < 7 ? (float)(1.0 / 18.0) : (float)(1.0 / 36.0));
}
template <int q> __device__ __forceinline__ constexpr float w7() { return q == 0 ? 0.25f : 0.125f; }

// the reference's uniform Params block (js/live/kernel.js:44-50)
struct Params {
  float tau0, cs2smag, gbeta, prt, kfloor, momfac, srcscale, sigmax;
  int inflow_w, inflow_e, inflow_s, inflow_n, nsp, nsw, oclip, pad;
};

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// equilibrium polynomial 1 + 3cu + 4.5cu^2 - 1.5u^2 (boundary reconstruction)
__device__ __forceinline__ float epoly(float cu, float usq) { return 1.0f + cu * (3.0f + 4.5f * cu) - 1.5f * usq; }

// c.u over the nonzero components
template <int cx, int cy, int cz>
__device__ __forceinline__ float cdot(float x, float y, float z) {
  float s = 0.0f;
  if (cx) s = (cx > 0) ? s + x : s - x;
  if (cy) s = (cy > 0) ? s + y : s - y;
  if (cz) s = (cz > 0) ? s + z : s - z;
  return s;
}

// post-stream moments (js/live/kernel.js:264-277) from pair sums s_q = f_q + f_opp and differences d_q = f_q - f_opp
struct Moments {
  float rho, jx, jy, jz, th, Pxx, Pyy, Pzz, Pxy, Pxz, Pyz;
};

__device__ __forceinline__ Moments moments(const float (&f)[19], const float (&g)[7]) {
  float s[19], d[19];
#pragma unroll
  for (int q = 1; q < 19; q += 2) { s[q] = f[q] + f[q + 1]; d[q] = f[q] - f[q + 1]; }
  Moments m;
  m.rho = f[0] + ((s[1] + s[3]) + (s[5] + s[7])) + ((s[9] + s[11]) + (s[13] + s[15])) + s[17];
  m.jx = (d[1] + d[7]) + (d[9] + d[11]) + d[13];
  m.jy = (d[3] + d[7]) - d[9] + (d[15] + d[17]);
  m.jz = (d[5] + d[11]) - d[13] + (d[15] - d[17]);
  m.Pxx = (s[1] + s[7]) + (s[9] + s[11]) + s[13];
  m.Pyy = (s[3] + s[7]) + (s[9] + s[15]) + s[17];
  m.Pzz = (s[5] + s[11]) + (s[13] + s[15]) + s[17];
  m.Pxy = s[7] - s[9];
  m.Pxz = s[11] - s[13];
  m.Pyz = s[15] - s[17];
  m.th = ((g[0] + g[1]) + (g[2] + g[3])) + ((g[4] + g[5]) + g[6]);
  return m;
}

struct Forces {
  float Fx, Fy, Fz, sg, ux, uy, uz, inv_rho;
};

// buoyancy, limited log-law drag at the first cell of building-free columns, top sponge, force-corrected velocity
// (js/live/kernel.js:279-306). DRAG / SPONGE = false drop branches the caller has proven inactive.
template <bool DRAG, bool SPONGE>
__device__ __forceinline__ Forces forces(const Params &P, int NZ, int k, int kg, int ktc, float cd, float uI, float vI,
                                         const Moments &m) {
  Forces o;
  const float rho = m.rho, ir = rcp(rho);
  o.inv_rho = ir;
  float Fx = 0.0f, Fy = 0.0f;
  float Fz = rho * P.gbeta * m.th;
  if (DRAG && k == kg && kg == ktc) {
    const float ux0 = m.jx * ir, uy0 = m.jy * ir;
    const float sp = fsqrt(ux0 * ux0 + uy0 * uy0);
    const bool lim = cd * sp > 0.5f;                     // implicit limit: never reverse the flow within a step
    const float c = lim ? -0.5f * rho : -rho * cd * sp;
    Fx = c * ux0;
    Fy = c * uy0;
  }
  float sg = 0.0f;
  if (SPONGE && k >= NZ - P.nsp) {
    const float aa = ((float)(k - (NZ - P.nsp)) + 1.0f) / (float)P.nsp;
    sg = P.sigmax * aa * aa;
    const float rs = rho * sg;
    Fx += rs * (uI - m.jx * ir);
    Fy += rs * (vI - m.jy * ir);
    Fz -= rs * (m.jz * ir);
  }
  o.Fx = Fx; o.Fy = Fy; o.Fz = Fz; o.sg = sg;
  o.ux = (m.jx + 0.5f * Fx) * ir;
  o.uy = (m.jy + 0.5f * Fy) * ir;
  o.uz = (m.jz + 0.5f * Fz) * ir;
  return o;
}

// relaxation (js/live/kernel.js:308-338)
struct Relax {
  float om45, fp3, fp9, omg, Nxx, Nyy, Nzz, Nxy, Nxz, Nyz, C0;
};

__device__ __forceinline__ Relax relax(const Params &P, int NX, int NY, int i, int j, float kb, const Moments &m,
                                       const Forces &F) {
  Relax r;
  const float rho = m.rho, ux = F.ux, uy = F.uy, uz = F.uz;
  const float third = (float)(1.0 / 3.0);
  const float r3 = rho * third;
  const float rux = rho * ux, ruy = rho * uy, ruz = rho * uz;
  r.Nxx = m.Pxx - r3 - rux * ux;
  r.Nyy = m.Pyy - r3 - ruy * uy;
  r.Nzz = m.Pzz - r3 - ruz * uz;
  r.Nxy = m.Pxy - rux * uy;
  r.Nxz = m.Pxz - rux * uz;
  r.Nyz = m.Pyz - ruy * uz;
  const float Pn = fsqrt(r.Nxx * r.Nxx + r.Nyy * r.Nyy + r.Nzz * r.Nzz +
                         2.0f * (r.Nxy * r.Nxy + r.Nxz * r.Nxz + r.Nyz * r.Nyz));
  float tb = P.tau0 + 3.0f * P.momfac * kb;
  // viscous sponge next to outflow faces
  int dmin = 1000;
  if (P.inflow_w == 0) dmin = min(dmin, i);
  if (P.inflow_e == 0) dmin = min(dmin, NX - 1 - i);
  if (P.inflow_s == 0) dmin = min(dmin, j);
  if (P.inflow_n == 0) dmin = min(dmin, NY - 1 - j);
  if (dmin < P.nsw) {
    const float rr = ((float)P.nsw - (float)dmin) / (float)P.nsw;
    tb += 0.15f * rr * rr;
  }
  const float tau = 0.5f * (tb + fsqrt(tb * tb + P.cs2smag * Pn * F.inv_rho));
  const float nusgs = (tau - tb) * third;
  const float kap = nusgs / P.prt + kb + P.kfloor;      // P.prt is uniform: one scalar division per wave
  const float itau = rcp(tau);
  const float om = 1.0f - itau, fpre = 1.0f - 0.5f * itau;
  r.omg = rcp(0.5f + 4.0f * kap);
  r.om45 = 4.5f * om;
  r.fp3 = 3.0f * fpre;
  r.fp9 = 9.0f * fpre;
  const float usq = ux * ux + uy * uy + uz * uz;
  const float trN = (r.Nxx + r.Nyy + r.Nzz) * third;
  const float uF = ux * F.Fx + uy * F.Fy + uz * F.Fz;
  // the part of w_q^-1 f_q^post common to every q: rho (1 - 1.5 u^2) - 4.5 om trN - 3 fpre u.F
  r.C0 = rho * (1.0f - 1.5f * usq) - r.om45 * trN - r.fp3 * uF;
  return r;
}

// regularised collision with Guo forcing of the pair (q, q+1), q odd (js/live/kernel.js:340-367):
// f_q^post / w_q = C0 + 3 rho cu + 4.5 rho cu^2 + 4.5 om (c c):N + fpre (3 cF + 9 cu cF), split into the part
// even in c (E) and the odd part (O); the opposite population gets E - O.
template <int q>
__device__ __forceinline__ void collide_pair(const Moments &m, const Forces &F, const Relax &R, float &fq, float &fo) {
  constexpr int cx = CX[q], cy = CY[q], cz = CZ[q];
  const float cu = cdot<cx, cy, cz>(F.ux, F.uy, F.uz);
  const float cF = cdot<cx, cy, cz>(F.Fx, F.Fy, F.Fz);
  float qn = 0.0f;                                      // (c c):N
  if (cx) qn += R.Nxx;
  if (cy) qn += R.Nyy;
  if (cz) qn += R.Nzz;
  if (cx && cy) qn += (float)(2 * cx * cy) * R.Nxy;
  if (cx && cz) qn += (float)(2 * cx * cz) * R.Nxz;
  if (cy && cz) qn += (float)(2 * cy * cz) * R.Nyz;
  const float E = R.C0 + cu * (4.5f * m.rho * cu + R.fp9 * cF) + R.om45 * qn;
  const float O = 3.0f * m.rho * cu + R.fp3 * cF;
  const float w = w19<q>();
  fq = w * (E + O);
  fo = w * (E - O);
}

__device__ __forceinline__ float collide_rest(const Relax &R) { return w19<0>() * R.C0; }

// D3Q7 collision with the rooftop source and the top-sponge damping (js/live/kernel.js:369-386):
// g^post = ((1 - omg) g + omg w th (1 + 4 c.u) + w S) (1 - sg)
struct Scalar {
  float a, b, cs0, cs1;   // a = (1-omg)(1-sg), b = omg w1 th (1-sg) (w1 = 1/8), source terms w S (1-sg)
};

__device__ __forceinline__ Scalar scalar_coeffs(float th, float omg, float S, float sg) {
  Scalar c;
  const float ds = 1.0f - sg;
  c.a = (1.0f - omg) * ds;
  c.b = omg * 0.125f * th * ds;
  c.cs0 = (omg * 0.25f * th + 0.25f * S) * ds;
  c.cs1 = 0.125f * S * ds;
  return c;
}

}  // namespace aht
