'use strict';
// The reference's stability matrix (tests/live/validate.html:162-198) through the Node host on the GPU: for dx 20 /
// 30 / 40 and six met / flux / campus combinations (one switching met, flux and campuses during the run), 10,000
// steps with read() + _dbg.stats() every 1,000 steps. Prints one JSON object with the same record per run and check
// as the reference's `stab` result. Usage: node tests/js/stab_test.js [--steps 10000] [--dxs 20,30,40]
const path = require('path');
const { createLiveSolver } = require(path.join(__dirname, '..', '..', 'ashburn-heat-twin_amd', 'js', 'solver.js'));
const { CAMPUSES } = require(path.join(__dirname, '..', '..', 'ashburn-heat-twin_amd', 'js', 'dump.js'));

const campusList = s => (s || '').split('').filter(c => CAMPUSES[c]).map(c => CAMPUSES[c]);
const arg = (k, d) => { const i = process.argv.indexOf('--' + k); return i > 0 ? process.argv[i + 1] : d; };

// the reference's combinations (tests/live/validate.html:165-173)
const COMBOS = [
  { name: 'calm-1ms-convective', met: { u10: 1, dirFrom: 300, wstar: 2.5, blh: 2000 }, flux: 800, campus: 'AB' },
  { name: 'strong-8ms-neutral', met: { u10: 8, dirFrom: 45, wstar: 0, blh: 500 }, flux: 800, campus: 'C' },
  { name: 'mid-4ms-south', met: { u10: 4, dirFrom: 180, wstar: 1.2, blh: 1200 }, flux: 0, campus: 'A' },
  { name: 'calm-1ms-neutral', met: { u10: 1, dirFrom: 210, wstar: 0, blh: 500 }, flux: 800, campus: 'AB' },
  { name: 'axis-270-2ms', met: { u10: 2, dirFrom: 270, wstar: 0.5, blh: 700 }, flux: 400, campus: 'BC' },
  { name: 'switching', met: { u10: 3.09, dirFrom: 210, wstar: 1.67, blh: 903 }, flux: 400, campus: '',
    changes: { 2500: { met: { u10: 8, dirFrom: 90 }, campus: 'AB' }, 5000: { met: { u10: 1, dirFrom: 315, wstar: 2.5, blh: 2000 }, flux: 800 },
      7500: { met: { u10: 5, dirFrom: 0, wstar: 0, blh: 500 }, campus: '' } } },
];

function surfStats(r) {
  let mx = -Infinity, sum = 0, nan = 0, sp = 0;
  for (let k = 0; k < r.surf.length; k++) {
    const t = r.surf[k];
    if (!Number.isFinite(t)) { nan++; continue; }
    if (t > mx) mx = t;
    sum += t;
    const a = r.u[k], b = r.v[k];
    if (!Number.isFinite(a) || !Number.isFinite(b)) nan++; else sp += Math.hypot(a, b);
  }
  return { surfMax: mx, windMean: sp / r.surf.length, nan };
}

async function main() {
  const nsteps = +arg('steps', 10000), dxs = arg('dxs', '20,30,40').split(',').map(Number);
  const res = [];
  for (const dx of dxs) {
    for (const c of COMBOS) {
      const t0 = Date.now();
      const s = await createLiveSolver({ dx, u10Max: 8, heightM: 720, met: Object.assign({}, c.met), existingFlux: c.flux });
      s.setNewCampuses(campusList(c.campus));
      const rec = { dx, name: c.name, checks: [] };
      const changeAt = Object.keys(c.changes || {}).map(Number);
      for (let done = 0; done < nsteps;) {
        const ch = c.changes && c.changes[done];
        if (ch) {
          if (ch.met) s.setMet(ch.met);
          if (ch.flux !== undefined) s.setExistingFlux(ch.flux);
          if (ch.campus !== undefined) s.setNewCampuses(campusList(ch.campus));
        }
        const nxt = Math.min(nsteps, done + 500, ...changeAt.filter(k => k > done));
        await s.step(nxt - done);
        done = nxt;
        if (done % 1000 === 0 || done === nsteps) {
          const r = await s.read({ vol: false });
          const ss = surfStats(r), st = await s._dbg.stats();
          let umax = 0;
          for (let k = 0; k < r.u.length; k++) umax = Math.max(umax, Math.hypot(r.u[k], r.v[k]));
          rec.checks.push({ step: done, t: r.simTime, surfMax: ss.surfMax, wind: ss.windMean, umax, thetaMax: st.thetaMax,
            heatMJ: st.heatContentMJ, umaxDomain: st.umax, nonFinite: st.nonFinite + ss.nan });
        }
      }
      rec.flowReinits = s.flowReinits;
      rec.dt = s.info.dt;
      rec.wall_s = (Date.now() - t0) / 1e3;
      s.destroy();
      console.error(`[stab] dx ${dx} ${c.name}: ${rec.wall_s.toFixed(1)} s`);
      res.push(rec);
    }
  }
  console.log(JSON.stringify({ results: res }));
}

main().catch(e => { console.error(e && e.stack || e); process.exit(1); });
