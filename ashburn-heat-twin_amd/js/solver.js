'use strict';
// Node solver host: the reference's live-solver interface (js/live/solver.js, spec js/live/README.md:30-63) over the
// N-API addon of libaht.so. Geometry, parameters and columns come from scenario.js; every device operation goes
// through the addon, which fails loudly when the HIP library or a GPU is missing (there is no CPU path).
//
// Calls that touch the device are serialised on one promise chain per solver, so read() issued while steps are in
// flight sees the state after them (the WebGPU queue ordering of the reference) and the C-ABI is only ever used by
// one thread at a time.

const path = require('path');
const P = require('./physics');
const { Scenario } = require('./scenario');

let addon = null;
function loadAddon() {
  if (!addon) {
    const p = path.join(__dirname, '..', 'lib', 'aht_node.node');
    try {
      addon = require(p);
    } catch (e) {
      throw new Error(`aht: the N-API addon ${p} is not built or cannot load libaht.so (${e.message})`);
    }
  }
  return addon;
}

const STEP_ACC = 1, STEP_OUT = 2;
const STORAGE = { f32: 0, f16: 1, bf16: 2, f16s: 4 };

/** createLiveSolver (js/live/solver.js:44-101). opts: dx, u10Max, heightM, base (static10 directory),
 *  synthetic (generator spec), nz, alpha, nsp, nsw, outflowClip, existingFlux, met, sideInflowDeg, metReinit, dirTol,
 *  allowAnyDx, onProgress, device (HIP ordinal), storage ('f32' | 'f16' | 'bf16' | 'f16s'), block, autotune (tuning steps
 *  per launch-shape candidate, run once after the flow is initialised; 0 or absent = default shape) */
async function createLiveSolver(opts) {
  const o = Object.assign({ dx: 30, u10Max: 8, heightM: 720, onProgress: () => {}, device: 0 }, opts || {});
  const prog = (t, f) => { try { o.onProgress(t, f); } catch (e) { /* UI errors ignored */ } };
  const A = loadAddon();
  prog('Building the grid', 0.2);
  const sc = new Scenario(o);          // validates dx, u10Max, heightM with the reference's messages
  const s = new Solver(A, sc, o, prog);
  try {
    s._init();
    if (o.autotune) await s.autotune(+o.autotune);
  } catch (e) {
    s.destroy();
    throw (e instanceof Error ? e : new Error(String(e)));
  }
  prog('Ready', 1);
  return s;
}

// shortest signed turn from direction a to direction b, degrees in [-180, 180)
const signedTurn = (a, b) => ((b - a + 540) % 360) - 180;
const wrap360 = d => ((d % 360) + 360) % 360;

/** The inflow the interior flow has adjusted to, and the rule that re-initialises the flow on a large turn
 *  (js/live/solver.js:271-285, :296-302): the adjusted state relaxes towards the applied wind with time constant
 *  ADJ_TAU_S of simulated time; a new wind that turns more than dirTol from the adjusted or the applied direction,
 *  with neither calm, re-initialises the flow and becomes the adjusted state at once. */
class AdjustedInflow {
  constructor(tolDeg, enabled) {
    this.tol = tolDeg; this.enabled = enabled;
    this.applied = null;    // { dir, u } of the last met
    this.adjusted = null;   // { dir, u } the interior flow has adjusted to
  }

  /** a new met; true when the flow must be re-initialised */
  observe(dir, u) {
    const ad = this.adjusted, ap = this.applied;
    let reinit = false;
    if (this.enabled && ad && ap && u > P.CALM && Math.max(ad.u, ap.u) > P.CALM) {
      const turn = Math.max(Math.abs(signedTurn(ad.dir, dir)), Math.abs(signedTurn(ap.dir, dir)));
      reinit = turn > this.tol;
    }
    this.applied = { dir, u };
    if (!ad || reinit) this.adjusted = { dir, u };
    return reinit;
  }

  /** seconds of simulated time pass */
  relax(seconds) {
    const ad = this.adjusted, ap = this.applied;
    if (!ad || !ap) return;
    const a = 1 - Math.exp(-seconds / P.ADJ_TAU_S);
    ad.dir = wrap360(ad.dir + signedTurn(ad.dir, ap.dir) * a);
    ad.u += (ap.u - ad.u) * a;
  }

  /** the flow was reset to equilibrium with the applied wind */
  reset() { if (this.applied) this.adjusted = Object.assign({}, this.applied); }
}

class Solver {
  constructor(A, sc, o, prog) {
    this._A = A; this._sc = sc; this._o = o; this._prog = prog;
    this._h = null; this._destroyed = false;
    this._n = 0; this._steps = 0; this._q = Promise.resolve();
    this.flowReinits = 0;
    this._inflow = new AdjustedInflow(P.def(o.dirTol, P.DIR_TOL), o.metReinit !== false);
    this.info = null;
    this.campusInfo = [];
  }

  get simTime() { return this._n * this._sc.dt; }
  get grid() { return this._sc.grid; }

  _init() {
    const sc = this._sc, A = this._A;
    this._prog('Allocating GPU memory', 0.3);
    const storage = STORAGE[this._o.storage || 'f32'];
    if (storage === undefined) throw new Error(`aht: unknown storage ${this._o.storage}`);
    this._h = A.create({ nx: sc.nx, ny: sc.ny, nz: sc.nz, storage, device: +this._o.device || 0, block: this._o.block || 0 });
    this.info = Object.assign(sc.info(), { adapter: 'AMD Instinct (HIP, gfx950)', storage: this._o.storage || 'f32' });
    this.nzv = this.info.nzv;
    this._uploadColumns();
    this._applyMet(this._o.met || P.DEFAULT_MET);
    this._prog('Initialising the flow', 0.85);
    this.resetFlow();
  }

  _check() {
    if (this._destroyed) throw new Error('live solver: destroyed');
  }

  // device work in call order; a call made after destroy() rejects, work queued before it still runs (destroy waits
  // for it, as the reference's queued GPU work outlives buffer.destroy())
  _enqueue(fn) {
    if (this._destroyed) return Promise.reject(new Error('live solver: destroyed'));
    const p = this._q.then(() => fn());
    this._q = p.catch(() => {});
    return p;
  }

  _uploadColumns() {
    const c = this._sc.col;
    this._A.setColumns(this._h, c.kt, c.ktop, c.cd, c.q);
    this.info.nFluid = this._sc.nFluid;
  }

  _uploadParams() {
    this._A.setParams(this._h, this._sc.params, this._sc.profBlock());
  }

  _applyMet(m) {
    const met = this._sc.setMet(m);
    this._uploadParams();
    if (this._inflow.observe(met.dirFrom, met.u10)) {
      this._A.initFlow(this._h);        // theta kept, simTime keeps running
      this.flowReinits += 1;
    }
    return met;
  }

  /** { u10, dirFrom, wstar, blh, ustar }; applies from the next step (js/live/solver.js:245-285) */
  setMet(m) {
    this._check();
    return this._enqueue(() => this._applyMet(m || {}));
  }

  /** W per m^2 of operational data-center roof (js/live/solver.js:305-310) */
  setExistingFlux(wm2) {
    this._check();
    this._sc.setExistingFlux(wm2);
    this.info.heatExistingMW = this._sc.dcRoof * this._sc.flux / 1e6;
    return this._enqueue(() => this._uploadColumns());
  }

  /** [{ id, halls: [[x0, z0, x1, z1], ...] (scene metres), mw }]; replaces the previous list; cells uncovered by a
   *  removed campus restart at rest (js/live/solver.js:313-356) */
  setNewCampuses(list) {
    this._check();
    const g = this._sc.setCampuses(list);
    this.campusInfo = this._sc.ras.info;
    return this._enqueue(() => {
      this._uploadColumns();
      if (g.grow) this._A.reinitCells(this._h, g.oldKtop);
    });
  }

  /** theta -> 0, flow kept, simTime -> 0 */
  resetHeat() {
    this._check();
    this._n = 0; this._steps = 0;
    return this._enqueue(() => this._A.resetHeat(this._h));
  }

  /** equilibrium with the inflow profile at the local height above terrain, theta -> 0, simTime -> 0 */
  resetFlow() {
    this._check();
    this._n = 0; this._steps = 0;
    this._inflow.reset();
    return this._enqueue(() => { this._A.initFlow(this._h); this._A.resetHeat(this._h); });
  }

  /** n steps; resolves when the GPU has finished them */
  step(n, flags) {
    try { this._check(); } catch (e) { return Promise.reject(e); }
    n = Math.max(0, Math.floor(n === undefined ? 1 : n));
    this._n += n; this._steps += n;
    this._inflow.relax(n * this._sc.dt);
    return this._enqueue(() => (n > 0 ? this._A.step(this._h, n, flags || 0) : undefined));
  }

  /** snapshot of the near-ground state: { simTime, surf, u, v, vol, volMax, nzv, thetaMax, heatMW, steps } */
  read(opts) {
    const o = Object.assign({ vol: true, wind: true, volMax: 3.0 }, opts || {});
    try { this._check(); } catch (e) { return Promise.reject(e); }
    const simTime = this.simTime, steps = this._steps, heatMW = this._sc.heatMW;
    return this._enqueue(() => {
      const NXY = this._sc.nx * this._sc.ny, vm = o.volMax > 0 ? +o.volMax : 3.0;
      const surf = new Float32Array(NXY);
      const u = o.wind ? new Float32Array(NXY) : null, v = o.wind ? new Float32Array(NXY) : null;
      this._A.readSurface(this._h, this._sc.conv, surf, u, v);
      const vol = o.vol ? new Uint8Array(this.nzv * NXY) : null;
      const thetaMax = this._A.readVolume(this._h, this.nzv, vm, vol);
      return { simTime, surf, u, v, vol, volMax: vm, nzv: this.nzv, thetaMax, heatMW, steps };
    });
  }

  /** launch-shape autotuner of this grid (no counterpart in the reference): times every block x levels x tile-window
   *  candidate for `steps` steps, keeps the fastest and resolves to it; the state and simTime do not change */
  autotune(steps) {
    return this._enqueue(() => this._A.autotune(this._h, steps === undefined ? 4 : steps));
  }

  /** the launch shape in use, {block, levels, window} */
  get launch() { this._check(); return this._A.getLaunch(this._h); }

  /** steps the model and returns MLUPS over all nx*ny*nz cells, device time from HIP events */
  benchmark(nSteps) {
    const n = nSteps === undefined ? 300 : nSteps;
    return this.step(Math.min(20, n)).then(() => {
      this._n += n; this._steps += n;
      this._inflow.relax(n * this._sc.dt);
      return this._enqueue(() => this._A.benchmark(this._h, n)).then(ms => this.info.nCells * n / (ms * 1e3));
    });
  }

  /** later calls reject at once; the device buffers go when the work queued before this call has finished
   *  (js/live/solver.js:485-493; the caller js/live.js:44 destroys a solver whose step may still be in flight) */
  destroy() {
    if (this._destroyed) return;
    this._destroyed = true;
    const h = this._h, A = this._A;
    if (h) this._q = this._q.then(() => { A.destroy(h); this._h = null; });
  }

  /** validation hooks (js/live/solver.js:496-577) */
  get _dbg() {
    const self = this, A = this._A, sc = this._sc;
    const N = () => sc.nx * sc.ny * sc.nz;
    return {
      domain() {
        return { nx: sc.nx, ny: sc.ny, nz: sc.nz, dx: sc.dx, zbase: sc.dom.zbase, dt: sc.dt, conv: sc.conv,
          params: Object.assign({}, sc.params), inflow: Object.assign({}, sc.inflow), met: Object.assign({}, sc.met),
          uin: sc.prof.ul, vin: sc.prof.vl, kbg: sc.prof.kl, kt: sc.col.kt, ktop: sc.col.ktop, cd: sc.col.cd, q: sc.col.q,
          heatMW: sc.heatMW, steps: self._steps, campusInfo: self.campusInfo };
      },
      readState() {
        return self._enqueue(() => {
          const f = new Float32Array(19 * N()), g = new Float32Array(7 * N());
          A.readState(self._h, f, g);
          return { f, g };
        });
      },
      writeState(f, g) { return self._enqueue(() => A.writeState(self._h, f || null, g || null)); },
      stepOut(n) { return self.step(n, STEP_OUT); },
      readDbg() {
        return self._enqueue(() => { const o = new Float32Array(4 * N()); A.readOut(self._h, o); return o; });
      },
      stepAcc(n) { return self.step(n, STEP_ACC); },
      clearDbg() { return self._enqueue(() => A.meanReset(self._h)); },
      /** FP64 time means: surface (theta, u, v in m/s, north-up) and the heat budget in MW / MJ */
      mean() {
        return self._enqueue(() => {
          const NXY = sc.nx * sc.ny, s3 = new Float64Array(3 * NXY);
          const b = A.readMean(self._h, s3);
          for (let k = NXY; k < 3 * NXY; k++) s3[k] *= sc.conv;
          const dx = sc.dx;
          return {
            nAcc: b.n_acc, surf: s3.subarray(0, NXY), u: s3.subarray(NXY, 2 * NXY), v: s3.subarray(2 * NXY),
            heatInjectedMW: b.heat_in * P.RHO_CP * dx * dx * dx / sc.dt / 1e6,
            heatAdvectedOutMW: P.RHO_CP * b.heat_out * sc.conv * dx * dx / 1e6,
            heatContentMJ: b.content * P.RHO_CP * dx * dx * dx / 1e6,
            massImbalance: b.mass_imbalance,
          };
        });
      },
      stats() {
        return self._enqueue(() => {
          const d = A.stats(self._h), c = sc.conv, dx = sc.dx;
          return { simTime: self.simTime, steps: self._steps, nFluid: d.n_fluid, nonFinite: d.non_finite,
            thetaMax: d.theta_max, heatContentMJ: d.sum_theta * P.RHO_CP * dx * dx * dx / 1e6,
            keMean: d.sum_ke / Math.max(d.n_fluid, 1) * c * c, rhoDevMean: d.sum_drho / Math.max(d.n_fluid, 1),
            umax: d.umax * c };
        });
      },
    };
  }
}

// the reference's entry point (js/live/solver.js:44; its caller js/live.js:42 imports it by this name), also under a
// shorter alias
const createSolver = createLiveSolver;

module.exports = { createLiveSolver, createSolver, Solver, STEP_ACC, STEP_OUT, loadAddon };
