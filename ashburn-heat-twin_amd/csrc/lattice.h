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
// Everything is templated on the arithmetic type R: float for the production kernels (f32, f16 and bf16 storage),
// double for the FP64 twin that serves as the reference of the reduced-precision error statements.
// The f32 result equals the reference expressions to f32 rounding (checked against the CPU oracle, which keeps the
// generator's order, to relative L2 <= 1e-6 per step). Cells: cell = (k*ny + j)*nx + i (js/live/kernel.js:17); the
// population layouts in HBM are in storage.h.
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

// reciprocals of the uniform parameters that the step divides by (host-computed once per parameter set)
struct ParamRecip {
  float prt, nsp, nsw;
};

// x / d for a parameter d with 1/d given: a multiply in f32 (within an ulp of the division, and no per-lane IEEE
// division sequence), the division itself in the FP64 twin (kept at the f64 oracle's rounding)
template <typename R>
__device__ __forceinline__ R div_param(R x, float d, float inv_d) {
  if constexpr (sizeof(R) == 8) return x / R(d);
  else return x * inv_d;
}

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ double fsqrt(double x) { return sqrt(x); }

// equilibrium polynomial 1 + 3cu + 4.5cu^2 - 1.5u^2 (boundary reconstruction)
template <typename R>
__device__ __forceinline__ R epoly(R cu, R usq) { return R(1) + cu * (R(3) + R(4.5) * cu) - R(1.5) * usq; }

// c.u over the nonzero components
template <int cx, int cy, int cz, typename R>
__device__ __forceinline__ R cdot(R x, R y, R z) {
  R s = R(0);
  if (cx) s = (cx > 0) ? s + x : s - x;
  if (cy) s = (cy > 0) ? s + y : s - y;
  if (cz) s = (cz > 0) ? s + z : s - z;
  return s;
}

// post-stream moments (js/live/kernel.js:264-277) from pair sums s_q = f_q + f_opp and differences d_q = f_q - f_opp
template <typename R>
struct Moments {
  R rho, jx, jy, jz, th, Pxx, Pyy, Pzz, Pxy, Pxz, Pyz;
};

template <typename R>
__device__ __forceinline__ Moments<R> moments(const R (&f)[19], const R (&g)[7]) {
  R s[19], d[19];
#pragma unroll
  for (int q = 1; q < 19; q += 2) { s[q] = f[q] + f[q + 1]; d[q] = f[q] - f[q + 1]; }
  Moments<R> m;
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

template <typename R>
struct Forces {
  R Fx, Fy, Fz, sg, ux, uy, uz, inv_rho;
};

// buoyancy, limited log-law drag at the first cell of building-free columns, top sponge, force-corrected velocity
// (js/live/kernel.js:279-306). DRAG / SPONGE = false drop branches the caller has proven inactive.
template <bool DRAG, bool SPONGE, typename R>
__device__ __forceinline__ Forces<R> forces(const Params &P, const ParamRecip &rp, int NZ, int k, int kg, int ktc, R cd,
                                            R uI, R vI, const Moments<R> &m) {
  Forces<R> o;
  const R rho = m.rho, ir = rcp(rho);
  o.inv_rho = ir;
  R Fx = R(0), Fy = R(0);
  R Fz = rho * R(P.gbeta) * m.th;
  if (DRAG && k == kg && kg == ktc) {
    const R ux0 = m.jx * ir, uy0 = m.jy * ir;
    const R sp = fsqrt(ux0 * ux0 + uy0 * uy0);
    const bool lim = cd * sp > R(0.5);                   // implicit limit: never reverse the flow within a step
    const R c = lim ? R(-0.5) * rho : -rho * cd * sp;
    Fx = c * ux0;
    Fy = c * uy0;
  }
  R sg = R(0);
  if (SPONGE && k >= NZ - P.nsp) {
    const R aa = div_param(R(k - (NZ - P.nsp)) + R(1), (float)P.nsp, rp.nsp);
    sg = R(P.sigmax) * aa * aa;
    const R rs = rho * sg;
    Fx += rs * (uI - m.jx * ir);
    Fy += rs * (vI - m.jy * ir);
    Fz -= rs * (m.jz * ir);
  }
  o.Fx = Fx; o.Fy = Fy; o.Fz = Fz; o.sg = sg;
  o.ux = (m.jx + R(0.5) * Fx) * ir;
  o.uy = (m.jy + R(0.5) * Fy) * ir;
  o.uz = (m.jz + R(0.5) * Fz) * ir;
  return o;
}

// relaxation (js/live/kernel.js:308-338)
template <typename R>
struct Relax {
  R fp3, fp9, omg, Nxx, Nyy, Nzz, Nxy, Nxz, Nyz, C0;
  R oa, od, oat, odt, o0t;   // om a_q for a_q = 4.5 w_q of the axis (0.25) and diagonal (0.125) classes, times trN
};

template <typename R>
__device__ __forceinline__ Relax<R> relax(const Params &P, const ParamRecip &rp, int NX, int NY, int i, int j, R kb,
                                          const Moments<R> &m, const Forces<R> &F) {
  Relax<R> r;
  const R rho = m.rho, ux = F.ux, uy = F.uy, uz = F.uz;
  const R third = R(1.0 / 3.0);
  const R r3 = rho * third;
  const R rux = rho * ux, ruy = rho * uy, ruz = rho * uz;
  r.Nxx = m.Pxx - r3 - rux * ux;
  r.Nyy = m.Pyy - r3 - ruy * uy;
  r.Nzz = m.Pzz - r3 - ruz * uz;
  r.Nxy = m.Pxy - rux * uy;
  r.Nxz = m.Pxz - rux * uz;
  r.Nyz = m.Pyz - ruy * uz;
  const R Pn = fsqrt(r.Nxx * r.Nxx + r.Nyy * r.Nyy + r.Nzz * r.Nzz +
                     R(2) * (r.Nxy * r.Nxy + r.Nxz * r.Nxz + r.Nyz * r.Nyz));
  R tb = R(P.tau0) + R(3) * R(P.momfac) * kb;
  // viscous sponge next to outflow faces
  int dmin = 1000;
  if (P.inflow_w == 0) dmin = min(dmin, i);
  if (P.inflow_e == 0) dmin = min(dmin, NX - 1 - i);
  if (P.inflow_s == 0) dmin = min(dmin, j);
  if (P.inflow_n == 0) dmin = min(dmin, NY - 1 - j);
  if (dmin < P.nsw) {
    const R rr = div_param(R(P.nsw) - R(dmin), (float)P.nsw, rp.nsw);
    tb += R(0.15) * rr * rr;
  }
  const R tau = R(0.5) * (tb + fsqrt(tb * tb + R(P.cs2smag) * Pn * F.inv_rho));
  const R nusgs = (tau - tb) * third;
  const R kap = div_param(nusgs, P.prt, rp.prt) + kb + R(P.kfloor);
  const R itau = rcp(tau);
  const R om = R(1) - itau, fpre = R(1) - R(0.5) * itau;
  r.omg = rcp(R(0.5) + R(4) * kap);
  r.fp3 = R(3) * fpre;
  r.fp9 = R(9) * fpre;
  const R usq = ux * ux + uy * uy + uz * uz;
  const R trN = (r.Nxx + r.Nyy + r.Nzz) * third;
  const R uF = ux * F.Fx + uy * F.Fy + uz * F.Fz;
  // the part of w_q^-1 (f_q^eq + F_q) common to every q: rho (1 - 1.5 u^2) - 3 fpre u.F
  r.C0 = rho * (R(1) - R(1.5) * usq) - r.fp3 * uF;
  // the regularised non-equilibrium part om a_q ((c c):N - trN) carries its own coefficient a_q, the reference's
  // literal 4.5 w_q (exactly 1.5, 0.25, 0.125), not 4.5 times the rounded weight
  r.oa = R(0.25) * om;
  r.od = R(0.125) * om;
  r.oat = r.oa * trN;
  r.odt = r.od * trN;
  r.o0t = R(1.5) * om * trN;
  return r;
}

// regularised collision with Guo forcing of the pair (q, q+1), q odd (js/live/kernel.js:340-367):
// f_q^post = w_q (C0 + 3 rho cu + 4.5 rho cu^2 + fpre (3 cF + 9 cu cF)) + om a_q ((c c):N - trN); the part of the
// bracket even in c (E) and the odd part (O) are formed once, the opposite population gets E - O.
template <int q, typename T>
__device__ __forceinline__ void collide_pair(const Moments<T> &m, const Forces<T> &F, const Relax<T> &R, T &fq, T &fo) {
  constexpr int cx = CX[q], cy = CY[q], cz = CZ[q];
  const T cu = cdot<cx, cy, cz>(F.ux, F.uy, F.uz);
  const T cF = cdot<cx, cy, cz>(F.Fx, F.Fy, F.Fz);
  T qn = T(0);                                          // (c c):N
  if (cx) qn += R.Nxx;
  if (cy) qn += R.Nyy;
  if (cz) qn += R.Nzz;
  if (cx && cy) qn += T(2 * cx * cy) * R.Nxy;
  if (cx && cz) qn += T(2 * cx * cz) * R.Nxz;
  if (cy && cz) qn += T(2 * cy * cz) * R.Nyz;
  const T E = R.C0 + cu * (T(4.5) * m.rho * cu + R.fp9 * cF);
  const T O = T(3) * m.rho * cu + R.fp3 * cF;
  const T t = q < 7 ? R.oa * qn - R.oat : R.od * qn - R.odt;
  const T w = T(w19<q>());
  fq = w * (E + O) + t;
  fo = w * (E - O) + t;
}

template <typename T>
__device__ __forceinline__ T collide_rest(const Relax<T> &R) { return T(w19<0>()) * R.C0 - R.o0t; }

// D3Q7 collision with the rooftop source and the top-sponge damping (js/live/kernel.js:369-386):
// g^post = ((1 - omg) g + omg w th (1 + 4 c.u) + w S) (1 - sg)
template <typename R>
struct Scalar {
  R a, b, cs0, cs1;   // a = (1-omg)(1-sg), b = omg w1 th (1-sg) (w1 = 1/8), source terms w S (1-sg)
};

template <typename R>
__device__ __forceinline__ Scalar<R> scalar_coeffs(R th, R omg, R S, R sg) {
  Scalar<R> c;
  const R ds = R(1) - sg;
  c.a = (R(1) - omg) * ds;
  c.b = omg * R(0.125) * th * ds;
  c.cs0 = (omg * R(0.25) * th + R(0.25) * S) * ds;
  c.cs1 = R(0.125) * S * ds;
  return c;
}

}  // namespace aht
