'use strict';
// Continuous batching keeps every result exact: serve a burst of requests through one batch, then run each request
// alone with the plain per-solver path and compare the time-mean warming fields bit for bit. Prints one JSON object.
// Default: six requests on small grids. --c5: eight requests of the serving grid mix (100 x 100 x 32, 200 x 200 x 48,
// 600 x 600 x 64 with the 6 km box's 4,000 buildings; seed 3 draws all three sizes) with a short spin-up and mean.
// --auto: the small-grid requests with storage chosen per request (f16s up to 200 m grid spacing, i.e. the 32- and
// 40-column grids of the 6 km box, f32 for the 24-column one), each compared with itself alone in its own storage.
const path = require('path');
const S = require(path.join(__dirname, '..', '..', 'ashburn-heat-twin_amd', 'js', 'serve.js'));

async function main() {
  const c5 = process.argv.includes('--c5'), auto = process.argv.includes('--auto');
  const o = c5
    ? { rate: 5, requests: 8, seed: 3, chunk: 50, spinupS: 120, meanS: 60, budgetGB: 200, slots: 64, keepSurf: true }
    : { rate: 50, requests: 6, seed: 3, chunk: 16, spinupS: 200, meanS: 100, budgetGB: 50, slots: 4,
      grids: [[[24, 20, 12], 0.4], [[32, 28, 14], 0.3], [[40, 24, 16], 0.3]], buildings: 40, sources: 4,
      keepSurf: true };
  if (auto) Object.assign(o, { storage: 'auto', f16sMaxDx: 200 });
  const reqs = S.makeRequests(o);
  const batched = await S.serveBatched(reqs, o);
  const out = { summary: batched.summary, identical: [], nAcc: [], grids: reqs.map(r => r.grid),
    storages: reqs.map(r => S.storageOf(r, o)) };
  for (const r of reqs) {
    const alone = await S.serveSequential([Object.assign({}, r, { arrival: 0 })],
      Object.assign({}, o, { storage: S.storageOf(r, o) }));
    const a = alone.results[0], b = batched.results.find(x => x.id === r.id);
    let same = a.surfMean.length === b.surfMean.length;
    for (let k = 0; same && k < a.surfMean.length; k++) if (a.surfMean[k] !== b.surfMean[k]) same = false;
    out.identical.push(same && a.warmingMean === b.warmingMean && a.heatAdvectedOutMW === b.heatAdvectedOutMW);
    out.nAcc.push([a.nAcc, b.nAcc, a.meanSteps]);
  }
  console.log(JSON.stringify(out));
}
main().catch(e => { console.error(e && e.stack || e); process.exit(1); });
