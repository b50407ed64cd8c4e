'use strict';
// Physical constants and the host-side parameter derivation of the coupled lattice-Boltzmann step.
// Follows js/live/solver.js:8-36 (constants, wind profile, K-profile), :103-119 (dt, nsp, nsw),
// :225-242 (uniform parameter block) and :245-270 (met clamping, profiles, inflow faces).
// ES2019 CommonJS (Node 12): no `??` or `?.`.

const RHO_CP = 1.16 * 1005.0;   // J m-3 K-1
const T0 = 303.0;               // K, Boussinesq reference temperature
const GRAV = 9.81;
const KAPPA = 0.4;
const CS = 0.16;                // Smagorinsky constant
const PRT = 0.7;                // turbulent Prandtl number
const NU0 = 0.2;                // m2/s molecular-plus-background viscosity
const KFLOOR = 0.5;             // m2/s diffusivity floor
const MOMFAC = 1.0;
const ULAT = 0.1;               // lattice Mach number at Umax
const ALPHA = 0.14;             // power-law exponent of the live solver
const ZCAP = 150.0;             // m, profile constant above
const SPONGE_TAU_S = 60.0;
const U10_MIN = 0.5;
const DIR_TOL = 30, CALM = 0.3, ADJ_TAU_S = 300;
const DEFAULT_MET = { u10: 3.09, dirFrom: 210, wstar: 1.67, blh: 903 };

function def(v, d) { return (v === undefined || v === null) ? d : v; }

/** power-law wind profile at z = (k + 0.5) dx, constant above zcap; m/s (js/live/solver.js:17-26) */
function windProfile(U10, dirFrom, alpha, nz, dx, zcap) {
  if (zcap === undefined) zcap = ZCAP;
  const th = dirFrom * Math.PI / 180, sx = -Math.sin(th), sy = -Math.cos(th);
  const u = new Float64Array(nz), v = new Float64Array(nz), sp = new Float64Array(nz);
  for (let k = 0; k < nz; k++) {
    const z = (k + 0.5) * dx;
    sp[k] = U10 * Math.pow(Math.min(z, zcap) / 10.0, alpha);
    u[k] = sp[k] * sx;
    v[k] = sp[k] * sy;
  }
  return { u, v, sp };
}

/** Holtslag-Boville K-profile, K = kappa w_m z (1 - z/h)^2 (js/live/solver.js:28-34) */
function kprofile(z, ustar, wstar, h) {
  const zz = Math.min(z, 0.1 * h);
  const wm = Math.cbrt(Math.pow(ustar, 3) + 0.6 * Math.pow(wstar, 3) * zz / h);
  const s = Math.min(1, Math.max(0, 1 - z / h));
  return KAPPA * wm * z * s * s;
}

const defaultUstar = u10 => KAPPA * u10 / Math.log(10 / 0.2);

/** time step and derived per-instance constants (js/live/solver.js:107-112) */
function timeStep(u10Max, nz, dx, alpha) {
  const a = def(alpha, ALPHA);
  const sp = windProfile(u10Max, 0, a, nz, dx).sp;
  let Umax = 0;
  for (let k = 0; k < nz; k++) Umax = Math.max(Umax, sp[k]);
  const dt = ULAT * dx / Umax;
  return { Umax, dt, conv: dx / dt };
}

/** clamp a met request against the previous one (js/live/solver.js:247-253) */
function clampMet(m, prev, u10Max) {
  m = m || {};
  prev = prev || DEFAULT_MET;
  const u10 = Math.min(Math.max(+def(m.u10, prev.u10), U10_MIN), u10Max);
  const dirFrom = (((+def(m.dirFrom, prev.dirFrom)) % 360) + 360) % 360;
  const wstar = Math.max(0, +def(m.wstar, prev.wstar));
  const blh = Math.min(Math.max(+def(m.blh, prev.blh), 50), 5000);
  const ustar = (m.ustar !== undefined && m.ustar !== null) ? Math.max(0, +m.ustar) : defaultUstar(u10);
  return { u10, dirFrom, wstar, blh, ustar };
}

/** lattice-unit profiles (u, v, kb) per level (js/live/solver.js:254-263) */
function latticeProfiles(met, nz, dx, dt, conv, alpha) {
  const wp = windProfile(met.u10, met.dirFrom, def(alpha, ALPHA), nz, dx);
  const ul = new Float32Array(nz), vl = new Float32Array(nz), kl = new Float32Array(nz), Kphys = new Float64Array(nz);
  const kmaxLat = 1 / 6;
  for (let k = 0; k < nz; k++) {
    ul[k] = wp.u[k] / conv;
    vl[k] = wp.v[k] / conv;
    Kphys[k] = kprofile((k + 0.5) * dx, met.ustar, met.wstar, met.blh);
    kl[k] = Math.min(Kphys[k] * dt / (dx * dx), kmaxLat);
  }
  return { ul, vl, kl, Kphys };
}

/** inflow faces by the eps rule on the direction the wind blows toward (js/live/solver.js:264-269) */
function inflowFaces(dirFrom, sideInflowDeg) {
  const th = dirFrom * Math.PI / 180, ux = -Math.sin(th), uy = -Math.cos(th), eps = 1e-3;
  const side = Math.sin(Math.max(0, def(sideInflowDeg, 0)) * Math.PI / 180);
  return {
    w: +(ux > eps || Math.abs(ux) < side), e: +(ux < -eps || Math.abs(ux) < side),
    s: +(uy > eps || Math.abs(uy) < side), n: +(uy < -eps || Math.abs(uy) < side),
  };
}

/** the scalar parameter block, f32-rounded as the device sees it (js/live/solver.js:225-238) */
function paramBlock(dx, dt, inflow, nsp, nsw, oclip) {
  const f = Math.fround;
  return {
    tau0: f(0.5 + 3 * NU0 * dt / (dx * dx)),
    cs2smag: f(18 * Math.SQRT2 * CS * CS),
    gbeta: f(GRAV / T0 * dt * dt / dx),
    prt: f(PRT),
    kfloor: f(KFLOOR * dt / (dx * dx)),
    momfac: f(MOMFAC),
    srcscale: 1,
    sigmax: f(Math.min(0.2, dt / SPONGE_TAU_S)),
    inflow_w: inflow.w, inflow_e: inflow.e, inflow_s: inflow.s, inflow_n: inflow.n,
    nsp, nsw, oclip,
  };
}

/** source strength in K per step per column: W * dt / (rho cp dx^3) (js/live/solver.js:359-372) */
function sourceScale(dt, dx) { return dt / (RHO_CP * dx * dx * dx); }

module.exports = {
  RHO_CP, T0, GRAV, KAPPA, CS, PRT, NU0, KFLOOR, MOMFAC, ULAT, ALPHA, ZCAP, SPONGE_TAU_S, U10_MIN,
  DIR_TOL, CALM, ADJ_TAU_S, DEFAULT_MET,
  def, windProfile, kprofile, defaultUstar, timeStep, clampMet, latticeProfiles, inflowFaces, paramBlock, sourceScale,
};
