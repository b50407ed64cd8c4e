'use strict';
// Scenario assembly shared by the solver host and the domain dumper: geometry (Ashburn static10 or a seeded
// synthetic grid), the per-instance time step, the column records (kt, ktop, cd, q; solver orientation) and the
// parameter block with its per-level profiles. Pure host code: no device calls.
// Column packing follows js/live/solver.js:359-376; geometry updates js/live/solver.js:323-356.

const path = require('path');
const P = require('./physics');
const D = require('./domain');
const S = require('./synth');

const DEFAULT_STATIC_DIR = path.join(__dirname, '..', '..', 'tests', 'golden', 'ashburn');

/** option validation with the reference's messages (js/live/solver.js:50-54) */
function validateOptions(o) {
  const dx = +o.dx;
  if (!o.synthetic && [20, 30, 40].indexOf(dx) < 0 && !o.allowAnyDx) throw new Error(`live solver: dx must be 20, 30 or 40 m (got ${o.dx})`);
  if (!(dx > 0)) throw new Error(`live solver: dx must be positive (got ${o.dx})`);
  const u10Max = +o.u10Max, heightM = +o.heightM;
  if (!(u10Max >= 1 && u10Max <= 20)) throw new Error(`live solver: u10Max must be within 1-20 m/s (got ${o.u10Max})`);
  if (!o.nz && !(heightM >= 300 && heightM <= 2000)) throw new Error(`live solver: heightM must be within 300-2000 m (got ${o.heightM})`);
  return { dx, u10Max, heightM };
}

/** geometry source: { agg, origin } */
function loadGeometry(o, dx) {
  if (o.synthetic) {
    const spec = Object.assign({}, o.synthetic, { dx });
    const agg = S.generate(spec);
    return { agg, origin: agg.origin, kind: 'synthetic' };
  }
  const st = D.loadStatic10(o.base || DEFAULT_STATIC_DIR);
  return { agg: D.aggregate(st, dx), origin: { x0: D.SCENE_X0, z0: D.SCENE_Z0 }, kind: 'ashburn' };
}

class Scenario {
  /** opts: createLiveSolver's options (dx, u10Max, heightM, nsp, nsw, outflowClip, existingFlux, met, sideInflowDeg,
   *  metReinit, dirTol, allowAnyDx, base) plus synthetic (generator spec), nz (fixed level count), alpha */
  constructor(opts) {
    const o = Object.assign({ dx: 30, u10Max: 8, heightM: 720 }, opts || {});
    const v = validateOptions(o);
    this.o = o;
    this.alpha = P.def(o.alpha, P.ALPHA);
    const g = loadGeometry(o, v.dx);
    this.agg = g.agg; this.origin = g.origin; this.kind = g.kind;
    const dom0 = D.buildDomain(this.agg, { heightM: v.heightM, nz: o.nz });
    this.dom = dom0;
    const { nx, ny, nz, dx } = dom0;
    this.nx = nx; this.ny = ny; this.nz = nz; this.dx = dx;
    const ts = P.timeStep(v.u10Max, nz, dx, this.alpha);
    this.dt = ts.dt; this.Umax = ts.Umax; this.conv = ts.conv; this.u10Max = v.u10Max;
    this.nsp = P.def(o.nsp, Math.max(1, Math.round(120 / dx)));
    this.nsw = P.def(o.nsw, Math.max(1, Math.round(160 / dx)));
    this.oclip = o.outflowClip === false ? 0 : 1;
    this.flux = P.def(o.existingFlux, 400);
    this.campuses = [];
    this.ras = D.rasterizeCampuses([], dx, nx, ny, this.origin);
    this.met = null; this.inflow = null; this.prof = null;
    let dcRoof = 0;
    for (let k = 0; k < this.agg.dcarea.length; k++) dcRoof += this.agg.dcarea[k];
    this.dcRoof = dcRoof;
    this.setGeometry();
    this.setMet(o.met || P.DEFAULT_MET);
  }

  /** campus list [{ id, mw, halls: [[x0, z0, x1, z1], ...] }] -> rasterised (js/live/solver.js:313-320) */
  setCampuses(list) {
    this.campuses = (list || []).map(c => ({ id: c.id, mw: +c.mw || 0, halls: (c.halls || []).map(h => h.slice(0, 4).map(Number)) }));
    this.ras = D.rasterizeCampuses(this.campuses, this.dx, this.nx, this.ny, this.origin);
    return this.setGeometry();
  }

  /** geometry with campuses (build_domain's campus rule), solver orientation; returns { changed, grow, oldKtop } */
  setGeometry() {
    const { nx, ny } = this;
    const dom = D.buildDomain(this.agg, { heightM: +this.o.heightM, nz: this.nz, newArea: this.ras.newArea, nzMin: this.nz });
    const oldKtop = this.ktopS;
    this.ktS = D.flipRows(dom.kt, nx, ny);
    this.ktopS = D.flipRows(dom.ktop, nx, ny);
    this.cdS = D.flipRows(dom.cd, nx, ny);
    this.grid = { ktop: dom.ktop, kt: dom.kt, nb: dom.nb, zbase: dom.zbase, dx: dom.dx, nx, ny, nz: this.nz };
    let changed = !oldKtop, grow = false;
    if (oldKtop) {
      for (let k = 0; k < oldKtop.length; k++) {
        if (this.ktopS[k] !== oldKtop[k]) changed = true;
        if (this.ktopS[k] < oldKtop[k]) grow = true;
      }
    }
    let nFluid = 0;
    for (let k = 0; k < this.ktopS.length; k++) nFluid += Math.max(0, this.nz - this.ktopS[k]);
    this.nFluid = nFluid;
    this.columns();
    return { changed, grow, oldKtop };
  }

  /** per-column source in K per step and the applied heat (js/live/solver.js:359-376) */
  columns() {
    const { nx, ny } = this, NXY = nx * ny;
    const kq = P.sourceScale(this.dt, this.dx);
    const dca = this.agg.dcarea, heat = this.ras.heatW;
    const q = new Float32Array(NXY), cd = new Float32Array(NXY);
    const kt = new Uint32Array(NXY), ktop = new Uint32Array(NXY);
    let totalW = 0;
    for (let jj = 0; jj < ny; jj++) {
      const row = ny - 1 - jj;
      for (let ii = 0; ii < nx; ii++) {
        const c = jj * nx + ii, cn = row * nx + ii;
        const W = dca[cn] * this.flux + heat[cn];
        totalW += W;
        kt[c] = this.ktS[c]; ktop[c] = this.ktopS[c]; cd[c] = this.cdS[c]; q[c] = W * kq;
      }
    }
    this.col = { kt, ktop, cd, q };
    this.heatMW = totalW / 1e6;
    return this.col;
  }

  setExistingFlux(wm2) {
    this.flux = Math.max(0, +wm2 || 0);
    return this.columns();
  }

  /** met -> profiles, inflow faces, parameter block (js/live/solver.js:245-270) */
  setMet(m) {
    const met = P.clampMet(m, this.met || P.DEFAULT_MET, this.u10Max);
    this.met = met;
    this.prof = P.latticeProfiles(met, this.nz, this.dx, this.dt, this.conv, this.alpha);
    this.inflow = P.inflowFaces(met.dirFrom, this.o.sideInflowDeg);
    this.params = P.paramBlock(this.dx, this.dt, this.inflow, this.nsp, this.nsw, this.oclip);
    return met;
  }

  /** [nz][4] f32 profile block: (u, v, kb, 0) in lattice units */
  profBlock() {
    const nz = this.nz, pr = new Float32Array(4 * nz);
    for (let k = 0; k < nz; k++) { pr[4 * k] = this.prof.ul[k]; pr[4 * k + 1] = this.prof.vl[k]; pr[4 * k + 2] = this.prof.kl[k]; }
    return pr;
  }

  info() {
    return {
      nx: this.nx, ny: this.ny, nz: this.nz, nzv: Math.min(this.nz, Math.round(420 / this.dx)), dx: this.dx, dt: this.dt,
      zbase: this.dom.zbase, baseElev: D.BASE_ELEV, x0: this.origin.x0, z0: this.origin.z0,
      heatExistingMW: this.dcRoof * this.flux / 1e6, dcRoofArea_m2: this.dcRoof,
      u10Max: this.u10Max, Umax: this.Umax, convU: this.conv, nsp: this.nsp, nsw: this.nsw,
      nCells: this.nx * this.ny * this.nz, nFluid: this.nFluid, top: this.dom.zbase + this.nz * this.dx, geometry: this.kind,
    };
  }
}

module.exports = { Scenario, DEFAULT_STATIC_DIR, validateOptions };
