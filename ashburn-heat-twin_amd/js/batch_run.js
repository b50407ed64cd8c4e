'use strict';
// The batched uneven-domain workload: many independent siting scenarios (grids drawn from BATCH_GRIDS, random wind,
// one inserted campus each; dump.js batchMember) advanced together, one launch per step for all of them with hipGraph-
// captured chunks, against the same scenarios run one at a time with the plain per-solver step.
//
// Launch shapes: the one-at-a-time runs use the launch-shape autotuner's pick per grid shape (aht_autotune, run once per
// distinct grid on the first domain of that shape and reused for the others); the batched launch steps every member
// one level per thread, its block size tuned online by the batch (aht_batch_set_block 0).
//
// Usage: node batch_run.js [--domains 64] [--steps 3000] [--chunk 50] [--seed 4] [--mode batch,sequential]
//        [--tune-steps 4 (0: no per-grid tuning)] [--block 0|64|128|256] [--check all|N (members compared bit for
//        bit between the two modes)] [--out file.json]
// Prints one JSON summary per mode on stdout; progress on stderr.

const fs = require('fs');
const { batchMember } = require('./dump');
const { loadAddon } = require('./solver');
const S = require('./synth');

function log(...a) { console.error('[batch_run]', ...a); }

function createSolver(A, sc) {
  const c = sc.col;
  const h = A.create({ nx: sc.nx, ny: sc.ny, nz: sc.nz, storage: 0, device: 0 });
  A.setColumns(h, c.kt, c.ktop, c.cd, c.q);
  A.setParams(h, sc.params, sc.profBlock());
  A.initFlow(h);
  A.resetHeat(h);
  return h;
}

// FNV-1a over the state's bytes: equal digests <=> (with overwhelming probability) bit-identical states
function digest(A, h, sc) {
  const N = sc.nx * sc.ny * sc.nz, f = new Float32Array(19 * N), g = new Float32Array(7 * N);
  A.readState(h, f, g);
  let x = 0x811c9dc5;
  for (const a of [new Uint32Array(f.buffer), new Uint32Array(g.buffer)])
    for (let k = 0; k < a.length; k++) x = Math.imul(x ^ a[k], 0x01000193) >>> 0;
  return x.toString(16);
}

async function tuneShapes(A, doms, o) {
  const shapes = new Map();
  if (!o.tuneSteps) return shapes;
  for (const d of doms) {
    const key = d.grid.join('x');
    if (shapes.has(key)) continue;
    const h = createSolver(A, d.sc);
    shapes.set(key, A.autotune(h, o.tuneSteps));
    A.destroy(h);
    log('tuned', key, JSON.stringify(shapes.get(key)));
  }
  return shapes;
}

function applyShape(A, h, d, shapes) {
  const l = shapes.get(d.grid.join('x'));
  if (l) A.setLaunch(h, l);
}

function cellsOf(d) { return d.sc.nx * d.sc.ny * d.sc.nz; }

async function runBatched(A, doms, shapes, o) {
  const B = A.batchCreate(0, 0);
  A.batchSetGrid(B, 1 << 30, 1);             // one block per tile
  A.batchSetBlock(B, o.block || 0);
  const hs = doms.map(d => { const h = createSolver(A, d.sc); applyShape(A, h, d, shapes); A.batchAdd(B, h); return h; });
  let busy = 0, done = 0, last = Date.now();
  const t0 = Date.now();
  while (done < o.steps) {
    const n = Math.min(o.chunk, o.steps - done);
    busy += await A.batchStep(B, n, null, null);
    done += n;
    if (Date.now() - last > 15000) { last = Date.now(); log(`batch: ${done}/${o.steps} steps`); }
  }
  const wall = (Date.now() - t0) / 1e3;
  const block = A.batchBlock(B);
  const digests = o.check ? hs.slice(0, o.check).map((h, k) => digest(A, h, doms[k].sc)) : [];
  for (const h of hs) { A.batchRemove(B, h); A.destroy(h); }
  A.batchDestroy(B);
  return { wall, busy, block, digests };
}

async function runSequential(A, doms, shapes, o) {
  let busy = 0, last = Date.now();
  const digests = [];
  const t0 = Date.now();
  for (let k = 0; k < doms.length; k++) {
    const d = doms[k], h = createSolver(A, d.sc);
    applyShape(A, h, d, shapes);
    for (let done = 0; done < o.steps;) {
      const n = Math.min(o.chunk, o.steps - done);
      busy += await A.benchmark(h, n);
      done += n;
    }
    if (k < o.check) digests.push(digest(A, h, d.sc));
    A.destroy(h);
    if (Date.now() - last > 15000) { last = Date.now(); log(`sequential: ${k + 1}/${doms.length} domains`); }
  }
  return { wall: (Date.now() - t0) / 1e3, busy, digests };
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
  const W = S.BATCH_WORKLOAD;
  const o = {
    domains: +(a.domains || W.domains), steps: +(a.steps || W.steps), chunk: +(a.chunk || 50),
    seed: a.seed === undefined ? W.seed : +a.seed, tuneSteps: a['tune-steps'] === undefined ? 4 : +a['tune-steps'],
    block: +(a.block || 0),
  };
  o.check = a.check === 'all' ? o.domains : +(a.check || 0);
  const modes = (a.mode || 'batch,sequential').split(',');
  const A = loadAddon();
  const t0 = Date.now();
  const doms = [];
  for (let d = 0; d < o.domains; d++) doms.push(batchMember(d, o.seed));
  log(`${doms.length} domains built in ${((Date.now() - t0) / 1e3).toFixed(1)} s`);
  const shapes = await tuneShapes(A, doms, o);
  const cells = doms.reduce((s, d) => s + cellsOf(d), 0);
  const fluid = doms.reduce((s, d) => s + d.sc.nFluid, 0);
  const mix = {};
  for (const d of doms) mix[d.grid.join('x')] = (mix[d.grid.join('x')] || 0) + 1;
  const out = { workload: { domains: o.domains, steps: o.steps, chunk: o.chunk, seed: o.seed, mix, cells, fluid_cells: fluid,
    launch_shapes: Object.fromEntries(shapes) } };
  for (const mode of modes) {
    const r = mode === 'sequential' ? await runSequential(A, doms, shapes, o) : await runBatched(A, doms, shapes, o);
    const cellSteps = cells * o.steps;
    out[mode] = {
      wall_s: +r.wall.toFixed(3), device_ms: +r.busy.toFixed(2),
      mlups_device: +(cellSteps / (r.busy * 1e3)).toFixed(1), mlups_wall: +(cellSteps / (r.wall * 1e6)).toFixed(1),
      fluid_mlups_device: +(fluid * o.steps / (r.busy * 1e3)).toFixed(1),
    };
    if (r.block !== undefined) out[mode].batch_block = r.block;
    if (o.check) out[mode].digests = r.digests;
    console.log(JSON.stringify(Object.assign({ mode }, out[mode])));
  }
  if (out.batch && out.sequential && o.check)
    out.identical = out.batch.digests.map((x, k) => x === out.sequential.digests[k]);
  if (out.batch && out.sequential) out.speedup_device = +(out.sequential.device_ms / out.batch.device_ms).toFixed(3);
  console.log(JSON.stringify(out));
  if (a.out) fs.writeFileSync(a.out, JSON.stringify(out, null, 1));
}

module.exports = { digest };

if (require.main === module) main().catch(e => { console.error(e && e.stack || e); process.exit(1); });
