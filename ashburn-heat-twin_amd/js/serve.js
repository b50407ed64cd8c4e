'use strict';
// Serving loop: siting scenarios of different grid sizes and winds arrive as a Poisson process; each runs a spin-up
// and then accumulates the FP64 time mean of the near-ground warming until the mean window is complete.
//
// Continuous batching: every active scenario is a member of one aht batch (one persistent launch per step for all
// members, chunks captured as hipGraphs). Arrivals wait in a FIFO queue and are admitted at chunk boundaries while
// the summed device footprint stays under the HBM budget and a slot is free; finished scenarios are read, leave the
// batch and free their memory, which admits the next ones. Each member steps exactly its own step count and averages
// exactly its own window, so its result is bit-identical to running it alone (tests/test_gpu_serve.py).
//
// Iteration-level priority (policy 'short-first', the default): a launch covers only the members that step in it
// (a member sitting out costs no workgroups), so a member that has little work left (cells x remaining steps <=
// --short-work) steps in every launch while members with much more work step in the first 1/--long-share of the
// chunk's launches only, as long as any short member is active. Small scenarios that arrive behind a 600 x 600 x 64 one then finish in about their own compute time instead
// of one big step per launch, and the big one resumes its full rate as soon as they leave. Policy 'fifo' steps every
// member in every launch.
//
// Storage: one format for every request (--storage 0 f32, 4 f16s, ...), or `auto`: f16s (f16 deviations, stochastic
// rounding) for requests whose grid spacing is at most --f16s-max-dx metres (default 10) and f32 for coarser ones, each
// format in its own batch (a batch has one format), the two batches stepped chunk by chunk in turn. Measured on this
// grid mix against the same requests in f32: the 10 m requests differ by at most 5e-5 in mean warming, while the 30 m
// and 60 m ones, whose unsteady flow amplifies any perturbation (a 1e-6 relative change of the wind moves f32's result
// by up to 8e-4), differ by up to 1.5e-3 and 4.7e-3 in 16-bit storage; the coarse grids are a small share of the work.
// The baseline (`mode: 'sequential'`) runs the same requests one at a time with the plain per-solver path (f32 under
// `auto`).
// Usage: node serve.js --rate 5 --requests 24 [--mode batch|sequential] [--spinup 2200] [--mean 600] [--seed 1]
//        [--storage 0|1|2|4|auto] [--f16s-max-dx 10]
//        [--chunk 50] [--budget-gb 200] [--slots 64] [--policy short-first|fifo] [--short-work 3e10]
//        [--long-share 8] [--grid blocks (0 = online tuning; default one block per tile)]
//        [--block 64|128|256 (threads per block of the batched step; default tuned online)] [--out file.json]

const fs = require('fs');
const { Scenario } = require('./scenario');
const S = require('./synth');
const { loadAddon } = require('./solver');
const P = require('./physics');

const STEP_ACC = 1;
const STORAGE_NAMES = ['f32', 'f16', 'bf16', 'f64', 'f16s'];
// batch grid: one workgroup per tile (the aht batch clamps the grid to the tile count). Measured at 600 x 600 x 64
// plus four 200 x 200 x 48 domains, it beat the persistent grids (8 and 32 blocks per CU) for every mix; --grid 0
// lets the batch's online tuner choose.
const ONE_BLOCK_PER_TILE = 1 << 30;

/** seeded request list: Poisson arrivals, grid mix, wind, campus */
function makeRequests(o) {
  const rng = S.mulberry32(o.seed || 1);
  const grids = o.grids || S.SERVE_GRIDS;
  const out = [];
  let t = 0;
  for (let k = 0; k < o.requests; k++) {
    t += -Math.log(1 - rng()) / o.rate;                 // exponential inter-arrival times
    let u = rng(), g = grids[grids.length - 1][0];
    for (const [shape, p] of grids) { if (u < p) { g = shape; break; } u -= p; }
    out.push({
      id: k, arrival: t, grid: g, u10: 2 + 7 * rng(), dirFrom: 360 * rng(),
      campusMW: 60 + 120 * rng(), seed: 1000 * (o.seed || 1) + k,
    });
  }
  return out;
}

/** the scenario of one request: synthetic 6 x 6 km box at dx = 6000 / nx with a 300 x 200 m campus */
function buildScenario(req, o) {
  const [nx, ny, nz] = req.grid;
  const dx = (o.boxM || 6000) / nx;
  const sc = new Scenario({
    synthetic: { nx, ny, hills: 12, buildings: o.buildings === undefined ? 4000 : o.buildings,
      sources: o.sources === undefined ? 210 : o.sources, seed: req.seed,
      hillAmp: o.hillAmp, hillWidth: o.hillWidth, side: o.side, height: o.height },
    dx, nz, u10Max: req.u10, alpha: 0.2, existingFlux: 400,
    met: { u10: req.u10, dirFrom: req.dirFrom, wstar: 1.67, blh: 903 },
  });
  const halls = [S.placeCampus(sc.agg, o.campusW || 300, o.campusH || 200, req.seed)];
  sc.setCampuses([{ id: 'new', mw: req.campusMW, halls }]);
  const spin = Math.ceil(o.spinupS / sc.dt), mean = Math.ceil(o.meanS / sc.dt);
  return { sc, spin, mean, total: spin + mean };
}

function result(req, job, A, h, now) {
  const sc = job.sc, NXY = sc.nx * sc.ny;
  const s3 = new Float64Array(3 * NXY);
  const b = A.readMean(h, s3);
  let sum = 0, peak = -Infinity;
  for (let k = 0; k < NXY; k++) { sum += s3[k]; if (s3[k] > peak) peak = s3[k]; }
  const dx = sc.dx;
  return {
    id: req.id, grid: req.grid, steps: job.total, meanSteps: job.mean, dt: sc.dt,
    warmingMean: sum / NXY, warmingPeak: peak, nAcc: b.n_acc,
    heatInjectedMW: b.heat_in * P.RHO_CP * dx * dx * dx / sc.dt / 1e6,
    heatAdvectedOutMW: P.RHO_CP * b.heat_out * sc.conv * dx * dx / 1e6,
    massImbalance: b.mass_imbalance, arrival: req.arrival, done: now, latency: now - req.arrival,
    surfMean: job.keepSurf ? Array.from(s3.subarray(0, NXY)) : undefined,
  };
}

/** the storage format of a request: o.storage, or under 'auto' f16s up to o.f16sMaxDx metres of grid spacing, else f32 */
function storageOf(req, o) {
  if (o.storage !== 'auto') return o.storage || 0;
  return (o.boxM || 6000) / req.grid[0] <= (o.f16sMaxDx || 10) ? 4 : 0;
}

function createSolver(A, job, o, storage) {
  const sc = job.sc, c = sc.col;
  const h = A.create({ nx: sc.nx, ny: sc.ny, nz: sc.nz, storage, device: o.device || 0 });
  A.setColumns(h, c.kt, c.ktop, c.cd, c.q);
  A.setParams(h, sc.params, sc.profBlock());
  A.initFlow(h);
  A.resetHeat(h);
  A.meanReset(h);
  return h;
}

/** launches of the next chunk: members [{cells, total, spin, done}] -> n launches, per-member step counts ns and the
 *  first averaging launch af (relative to the chunk). Under 'short-first', while any member has at most shortWork
 *  cell-steps left (among these members, or anyShort: in another batch), members with more step only in the first
 *  ceil(chunk / longShare) launches. */
function planChunk(members, chunk, policy, shortWork, longShare, anyShort = false) {
  const m = members.length, ns = new Int32Array(m), af = new Int32Array(m);
  const work = members.map(a => a.cells * (a.total - a.done));
  const prio = policy === 'short-first' && (anyShort || work.some(w => w <= shortWork));
  let n = 0;
  for (let k = 0; k < m; k++) {
    const a = members[k];
    ns[k] = Math.min(chunk, a.total - a.done);
    if (prio && work[k] > shortWork) ns[k] = Math.min(ns[k], Math.ceil(chunk / longShare));
    af[k] = Math.max(0, a.spin - a.done);
    n = Math.max(n, ns[k]);
  }
  return { n, ns, af };
}

const sleep = ms => new Promise(r => setTimeout(r, ms));
// progress on stderr at most every 15 s (long runs must not look idle)
function progress(o, mode) {
  let last = Date.now();
  return (done, total, active, queued) => {
    if (!o.quiet && Date.now() - last > 15000) {
      last = Date.now();
      console.error(`[serve ${mode} rate ${o.rate}] ${done}/${total} done, ${active} active, ${queued} queued`);
    }
  };
}
const pct = (a, p) => { const s = a.slice().sort((x, y) => x - y); return s[Math.min(s.length - 1, Math.floor(p * (s.length - 1) + 0.5))]; };

function summary(mode, o, results, t0, t1, peakBytes, busyMs) {
  const lat = results.map(r => r.latency);
  const span = t1 - t0;
  return {
    mode, rate: o.rate, requests: results.length, makespan_s: +span.toFixed(3),
    scenarios_per_hour: +(results.length / span * 3600).toFixed(1),
    latency_p50_s: +pct(lat, 0.5).toFixed(3), latency_p99_s: +pct(lat, 0.99).toFixed(3),
    latency_mean_s: +(lat.reduce((a, b) => a + b, 0) / lat.length).toFixed(3),
    peak_hbm_gb: +(peakBytes / 1e9).toFixed(2), gpu_busy_frac: +(busyMs / 1e3 / span).toFixed(3),
  };
}

/** continuous batching over one aht batch per storage format in use */
async function serveBatched(reqs, o) {
  const A = loadAddon();
  const batches = new Map();   // storage -> batch
  const batchOf = st => {
    if (!batches.has(st)) {
      const B = A.batchCreate(o.device || 0, st);
      A.batchSetGrid(B, o.grid === undefined ? ONE_BLOCK_PER_TILE : o.grid, o.useGraph === false ? 0 : 1);
      A.batchSetBlock(B, o.block || 0);
      batches.set(st, B);
    }
    return batches.get(st);
  };
  const budget = (o.budgetGB || 200) * 1e9, slots = o.slots || 64, chunk = o.chunk || 50;
  const shortWork = o.shortWork || 3e10, longShare = o.longShare || 8;
  const future = reqs.slice().sort((a, b) => a.arrival - b.arrival);
  const pending = [], active = [], results = [];
  let used = 0, peak = 0, busy = 0;
  const t0 = Date.now() / 1e3;
  const now = () => Date.now() / 1e3 - t0;
  const prog = progress(o, 'batch');
  while (future.length || pending.length || active.length) {
    prog(results.length, reqs.length, active.length, pending.length);
    while (future.length && future[0].arrival <= now()) pending.push(future.shift());
    // admission: FIFO, by footprint and slots
    while (pending.length && active.length < slots) {
      const r = pending[0], st = storageOf(r, o);
      const fp = A.footprint({ nx: r.grid[0], ny: r.grid[1], nz: r.grid[2], storage: st });
      if (used + fp > budget && active.length > 0) break;
      pending.shift();
      const job = buildScenario(r, o);
      job.keepSurf = !!o.keepSurf;
      const h = createSolver(A, job, o, st);
      A.batchAdd(batchOf(st), h);
      active.push({ r, job, h, st, done: 0, fp, cells: r.grid[0] * r.grid[1] * r.grid[2] });
      used += fp;
      peak = Math.max(peak, used);
    }
    if (!active.length) {
      if (future.length) await sleep(Math.max(0, (future[0].arrival - now()) * 1e3));
      continue;
    }
    // one chunk per batch in use, members in add order; the short-first priority looks at the members of every batch
    const anyShort = active.some(a => a.cells * (a.job.total - a.done) <= shortWork);
    for (const [st, B] of batches) {
      const mem = active.filter(a => a.st === st);
      if (!mem.length) continue;
      const { n, ns, af } = planChunk(mem.map(a => ({ cells: a.cells, total: a.job.total, spin: a.job.spin, done: a.done })),
        chunk, o.policy || 'short-first', shortWork, longShare, anyShort);
      busy += await A.batchStep(B, n, ns, af);
      for (let k = 0; k < mem.length; k++) mem[k].done += ns[k];
    }
    for (let k = active.length - 1; k >= 0; k--) {
      const a = active[k];
      if (a.done >= a.job.total) {
        results.push(result(a.r, a.job, A, a.h, now()));
        A.batchRemove(batchOf(a.st), a.h);
        A.destroy(a.h);
        used -= a.fp;
        active.splice(k, 1);
      }
    }
  }
  const t1 = now();
  const blocks = {}, grids = {};
  for (const [st, B] of batches) {
    blocks[STORAGE_NAMES[st]] = A.batchBlock(B);
    grids[STORAGE_NAMES[st]] = A.batchGrid(B);
    A.batchDestroy(B);
  }
  results.sort((a, b) => a.id - b.id);
  return {
    summary: Object.assign(summary('batch', o, results, 0, t1, peak, busy),
      { storage: o.storage === 'auto' ? `auto (f16s up to ${o.f16sMaxDx || 10} m, else f32)` : STORAGE_NAMES[o.storage || 0],
        batch_grid: grids, batch_block: blocks, policy: o.policy || 'short-first' }),
    results,
  };
}

/** the baseline: one scenario at a time with the plain per-solver path */
async function serveSequential(reqs, o) {
  const A = loadAddon();
  const chunk = o.chunk || 50;
  const queue = reqs.slice().sort((a, b) => a.arrival - b.arrival);
  const results = [];
  let peak = 0, busy = 0;
  const t0 = Date.now() / 1e3;
  const now = () => Date.now() / 1e3 - t0;
  const prog = progress(o, 'sequential');
  for (const r of queue) {
    prog(results.length, reqs.length, 1, 0);
    if (r.arrival > now()) await sleep((r.arrival - now()) * 1e3);
    const job = buildScenario(r, o);
    job.keepSurf = !!o.keepSurf;
    const st = o.storage === 'auto' ? 0 : (o.storage || 0);   // the plain step: f32 under 'auto'
    const h = createSolver(A, job, o, st);
    peak = Math.max(peak, A.footprint({ nx: r.grid[0], ny: r.grid[1], nz: r.grid[2], storage: st }));
    for (let done = 0; done < job.total;) {
      const n = Math.min(chunk, job.total - done), spinLeft = job.spin - done;
      if (spinLeft >= n) busy += await A.benchmark(h, n);
      else {
        if (spinLeft > 0) busy += await A.benchmark(h, spinLeft);
        const t = Date.now();
        await A.step(h, n - Math.max(0, spinLeft), STEP_ACC);
        busy += Date.now() - t;
      }
      done += n;
      prog(results.length, reqs.length, 1, 0);
    }
    results.push(result(r, job, A, h, now()));
    A.destroy(h);
  }
  const t1 = now();
  results.sort((a, b) => a.id - b.id);
  return { summary: summary('sequential', o, results, 0, t1, peak, busy), results };
}

function parseArgs(argv) {
  const a = {};
  for (let i = 0; i < argv.length; i++) {
    if (argv[i].startsWith('--')) {
      const k = argv[i].slice(2), v = argv[i + 1];
      if (v === undefined || v.startsWith('--')) a[k] = true; else { a[k] = v; i++; }
    }
  }
  return a;
}

async function main() {
  const a = parseArgs(process.argv.slice(2));
  const o = {
    rate: +(a.rate || 5), requests: +(a.requests || 24), seed: +(a.seed || 1), chunk: +(a.chunk || 50),
    spinupS: +(a.spinup || 2200), meanS: +(a.mean || 600), budgetGB: +(a['budget-gb'] || 200), slots: +(a.slots || 64),
    storage: a.storage === 'auto' ? 'auto' : +(a.storage || 0), f16sMaxDx: +(a['f16s-max-dx'] || 10), grid: a.grid === undefined ? ONE_BLOCK_PER_TILE : +a.grid, block: +(a.block || 0),
    useGraph: a['no-graph'] ? false : true,
    policy: a.policy || 'short-first', shortWork: +(a['short-work'] || 3e10), longShare: +(a['long-share'] || 8),
  };
  if (a.grids) o.grids = JSON.parse(a.grids);
  const reqs = makeRequests(o);
  const modes = (a.mode || 'batch').split(',');
  const out = {};
  for (const mode of modes) {
    const r = mode === 'sequential' ? await serveSequential(reqs, o) : await serveBatched(reqs, o);
    out[mode] = r;
    console.log(JSON.stringify(r.summary));
  }
  if (a.out) fs.writeFileSync(a.out, JSON.stringify(out));
}

module.exports = { makeRequests, buildScenario, planChunk, storageOf, serveBatched, serveSequential };

if (require.main === module) main().catch(e => { console.error(e && e.stack || e); process.exit(1); });
