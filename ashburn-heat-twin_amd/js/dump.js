'use strict';
// Domain dumps: a scenario's columns, parameter block and profiles written as a directory of little-endian
// binaries plus domain.json. The GPU parity tests, the CPU oracle and bench.py all consume the same dump, so the
// JavaScript host stays the only geometry builder (the dump-and-replay pattern of tests/live/validate.html:51-59).

const fs = require('fs');
const path = require('path');
const { Scenario } = require('./scenario');
const S = require('./synth');

// the reference's validation scenarios (tests/live/validate.html:29-35): 25 Aug 2024 met and three 120 MW campuses
const MET_AUG25 = { u10: 3.09, dirFrom: 210, wstar: 1.67, blh: 903, ustar: 0.32 };
const CAMPUSES = {
  A: { id: 'A', mw: 120, halls: [[-205, 475, 175, 605], [-205, 635, 175, 765], [-205, 795, 175, 925]] },
  B: { id: 'B', mw: 120, halls: [[-2645, 1035, -2265, 1165], [-2645, 1195, -2265, 1325], [-2645, 1355, -2265, 1485]] },
  C: { id: 'C', mw: 120, halls: [[1555, -1495, 1935, -1365], [1555, -1335, 1935, -1205], [1555, -1175, 1935, -1045]] },
};
const campusList = s => (s || '').split('').filter(c => CAMPUSES[c]).map(c => CAMPUSES[c]);

function writeBin(dir, name, arr) {
  fs.writeFileSync(path.join(dir, name), Buffer.from(arr.buffer, arr.byteOffset, arr.byteLength));
}

/** write sc's current state to dir */
function dumpScenario(sc, dir, extra) {
  fs.mkdirSync(dir, { recursive: true });
  const c = sc.col;
  writeBin(dir, 'kt.u32', c.kt);
  writeBin(dir, 'ktop.u32', c.ktop);
  writeBin(dir, 'cd.f32', c.cd);
  writeBin(dir, 'q.f32', c.q);
  writeBin(dir, 'prof.f32', sc.profBlock());
  const meta = Object.assign({
    nx: sc.nx, ny: sc.ny, nz: sc.nz, dx: sc.dx, dt: sc.dt, conv: sc.conv, zbase: sc.dom.zbase,
    params: sc.params, inflow: sc.inflow, met: sc.met, heatMW: sc.heatMW, info: sc.info(),
    campusInfo: sc.ras.info, Kphys: Array.from(sc.prof.Kphys),
  }, extra || {});
  fs.writeFileSync(path.join(dir, 'domain.json'), JSON.stringify(meta, null, 1));
  return meta;
}

/** scenario options of a named workload (synthetic generator) */
function workloadOptions(name, over) {
  const w = S.WORKLOADS[name];
  if (!w) throw new Error('unknown workload ' + name);
  const o = {
    synthetic: { nx: w.nx, ny: w.ny, hills: w.hills, buildings: w.buildings, sources: w.sources, seed: w.seed },
    dx: w.dx, nz: w.nz, u10Max: w.met.u10, alpha: w.alpha, existingFlux: w.flux,
    met: { u10: w.met.u10, dirFrom: w.met.dirFrom, wstar: 1.67, blh: 903 },
  };
  return Object.assign(o, over || {});
}

/** member d of the batched uneven-domain workload (S.BATCH_WORKLOAD): a Scenario with its campus placed; the seeds
 *  derive from (workload seed, d), so every process builds the identical domain */
function batchMember(d, seed) {
  const W = S.BATCH_WORKLOAD, base = 1000 * (seed === undefined ? W.seed : seed) + d;
  const rng = S.mulberry32(base);
  const grid = W.grids[Math.floor(rng() * W.grids.length)];
  const u10 = W.u10[0] + (W.u10[1] - W.u10[0]) * rng(), dirFrom = 360 * rng();
  const [nx, ny, nz] = grid;
  const sc = new Scenario({
    synthetic: { nx, ny, hills: 12, buildings: W.buildings, sources: W.sources, seed: base },
    dx: 6000 / nx, nz, u10Max: u10, alpha: 0.2, existingFlux: W.flux,
    met: { u10, dirFrom, wstar: 1.67, blh: 903 },
  });
  const halls = [S.placeCampus(sc.agg, W.campus[0], W.campus[1], base + 7919)];
  sc.setCampuses([{ id: 'campus', mw: W.campusMW, halls }]);
  return { sc, grid, u10, dirFrom, seed: base };
}

function parseArgs(argv) {
  const a = {};
  for (let i = 0; i < argv.length; i++) {
    const t = argv[i];
    if (t.startsWith('--')) {
      const k = t.slice(2);
      const nxt = argv[i + 1];
      if (nxt === undefined || nxt.startsWith('--')) a[k] = true;
      else { a[k] = nxt; i++; }
    }
  }
  return a;
}

/** options from CLI flags:
 *  --workload c1|c2|c3|c2flat              synthetic benchmark workload (c2flat: c2 without terrain or buildings)
 *  --batch-member d[,seed]                  member d of the batched uneven-domain workload (handled in main)
 *  --serve-request k[,seed]                 request k of the serving loop's request list (handled in main)
 *  --tiny nx,ny,nz,seed[,buildings,sources] small synthetic grid (dx from --dx, default 30)
 *  (default)                                the Ashburn static10 grid at --dx
 *  --u10 --dir --wstar --blh --ustar --flux --nsw --nsp --oclip --campus ABC --u10max --alpha --side */
function optionsFromArgs(a) {
  let o;
  if (a.workload) o = workloadOptions(a.workload);
  else if (a.tiny) {
    const t = String(a.tiny).split(',').map(Number);
    o = {
      synthetic: { nx: t[0], ny: t[1], hills: 3, buildings: t[4] === undefined ? 6 : t[4], sources: t[5] === undefined ? 3 : t[5], seed: t[3] || 1,
        hillAmp: [0, 3 * (+a.dx || 30)], hillWidth: [4 * (+a.dx || 30), 10 * (+a.dx || 30)],
        side: [0.5 * (+a.dx || 30), 4 * (+a.dx || 30)], height: [0.3 * (+a.dx || 30), 2.5 * (+a.dx || 30)] },
      dx: +a.dx || 30, nz: t[2], u10Max: 8, met: Object.assign({}, MET_AUG25),
    };
  } else {
    o = { dx: +a.dx || 30, u10Max: 8, heightM: 720, met: Object.assign({}, MET_AUG25) };
    if (a.base) o.base = a.base;
  }
  const met = o.met || {};
  if (a.u10 !== undefined) { met.u10 = +a.u10; if (a.ustar === undefined) delete met.ustar; }
  if (a.dir !== undefined) met.dirFrom = +a.dir;
  if (a.wstar !== undefined) met.wstar = +a.wstar;
  if (a.blh !== undefined) met.blh = +a.blh;
  if (a.ustar !== undefined) met.ustar = +a.ustar;
  o.met = met;
  if (a.flux !== undefined) o.existingFlux = +a.flux;
  if (a.nsw !== undefined) o.nsw = +a.nsw;
  if (a.nsp !== undefined) o.nsp = +a.nsp;
  if (a.oclip !== undefined) o.outflowClip = +a.oclip !== 0;
  if (a.u10max !== undefined) o.u10Max = +a.u10max;
  if (a.alpha !== undefined) o.alpha = +a.alpha;
  if (a.side !== undefined) o.sideInflowDeg = +a.side;
  if (a.nz !== undefined && !a.tiny) o.nz = +a.nz;
  return o;
}

module.exports = { MET_AUG25, CAMPUSES, campusList, dumpScenario, workloadOptions, optionsFromArgs, parseArgs,
  batchMember };

if (require.main === module) {
  const a = parseArgs(process.argv.slice(2));
  if (!a.out) {
    console.error('usage: node dump.js --out DIR [--workload c2 | --tiny nx,ny,nz,seed | --batch-member d | --dx 30] [flags]');
    process.exit(2);
  }
  let sc, source;
  if (a['serve-request'] !== undefined) {
    // request k of the serving loop's seeded request list (js/serve.js makeRequests at 5 req/s), as it builds it
    const t = String(a['serve-request']).split(',').map(Number);
    const S5 = require('./serve');
    const reqs = S5.makeRequests({ rate: 5, requests: t[0] + 1, seed: t[1] === undefined ? 1 : t[1] });
    const job = S5.buildScenario(reqs[t[0]], { spinupS: 2200, meanS: 600 });
    sc = job.sc;
    source = `serving request ${t[0]} (seed ${t[1] === undefined ? 1 : t[1]}): ${reqs[t[0]].grid.join('x')}, spin-up ${job.spin} + mean ${job.mean} steps`;
  } else if (a['batch-member'] !== undefined) {
    const t = String(a['batch-member']).split(',').map(Number);
    const m = batchMember(t[0], t[1]);
    sc = m.sc;
    source = `batch member ${t[0]} (seed ${m.seed}): ${m.grid.join('x')}, ${m.u10.toFixed(3)} m/s from ${m.dirFrom.toFixed(1)} deg`;
  } else {
    sc = new Scenario(optionsFromArgs(a));
    if (a.campus) sc.setCampuses(campusList(a.campus));
    source = a.workload || (a.tiny ? 'tiny' : 'ashburn');
  }
  const meta = dumpScenario(sc, a.out, { source });
  console.log(JSON.stringify({ out: a.out, nx: meta.nx, ny: meta.ny, nz: meta.nz, nFluid: meta.info.nFluid, heatMW: meta.heatMW }));
}
