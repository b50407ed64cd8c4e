'use strict';
// Drives the N-API addon built against the CPU stand-in of the C-ABI (tests/js/aht_standin.cpp): jobs issued without
// awaiting, destroy() of a solver and of a batch with work in flight, then the end of the process, either naturally
// or through process.exit(). The stand-in aborts on any threading or lifetime violation, including device state
// left on a thread that outlives the runtime at exit. Usage: node exit_test.js <addon.node> natural|exit
const A = require(process.argv[2]);
const mode = process.argv[3];

async function main() {
  const mk = () => A.create({ nx: 8, ny: 8, nz: 8, storage: 0, device: 0 });
  const out = {};
  const h1 = mk(), h2 = mk();
  await A.step(h1, 5, 0);
  out.bench = await A.benchmark(h2, 3);
  // destroy with a step in flight (js/live.js:44, :76): the step settles, later calls throw
  const inflight = A.step(h2, 20, 0);
  A.destroy(h2);
  out.stepBeforeDestroy = await inflight.then(() => 'resolved', e => 'rejected: ' + e.message);
  try { A.step(h2, 1, 0); out.stepAfterDestroy = 'no error'; } catch (e) { out.stepAfterDestroy = 'threw'; }
  // synchronous calls on a solver (or batch) with a job in flight throw; after it settles they run
  const h5 = mk();
  const s5 = A.step(h5, 20, 0);
  const tryCall = f => { try { f(); return 'ran'; } catch (e) { return 'threw'; } };
  out.statsDuringStep = tryCall(() => A.stats(h5));
  out.setParamsDuringStep = tryCall(() => A.resetHeat(h5));
  await s5;
  out.statsAfterStep = tryCall(() => A.stats(h5));
  const B5 = A.batchCreate(0, 0);
  A.batchAdd(B5, h5);
  const b5 = A.batchStep(B5, 5, null, null);
  out.memberStatsDuringBatchStep = tryCall(() => A.stats(h5));
  out.setBlockDuringBatchStep = tryCall(() => A.batchSetBlock(B5, 128));
  await b5;
  out.setBlockAfterBatchStep = tryCall(() => A.batchSetBlock(B5, 128));
  A.batchRemove(B5, h5);
  A.batchDestroy(B5);
  A.destroy(h5);
  // a batch and one of its members destroyed under a batch step
  const B = A.batchCreate(0, 0), h3 = mk();
  A.batchAdd(B, h1);
  A.batchAdd(B, h3);
  const bj = A.batchStep(B, 10, null, null);
  A.destroy(h3);
  A.batchDestroy(B);
  out.batchStep = await bj.then(ms => ms, e => 'rejected: ' + e.message);
  // jobs of one solver issued without awaiting run one at a time, in issue order
  const h4 = mk(), order = [];
  const jobs = [];
  for (let k = 0; k < 6; k++) jobs.push(A.step(h4, 2, 0).then(() => order.push(k)));
  await Promise.all(jobs);
  out.order = order;
  // a step left in flight when the script ends: natural exit waits for it
  out.lastStepQueued = true;
  A.step(h1, 5, 0).then(() => console.error('last step settled'));
  console.log(JSON.stringify(out));
  if (mode === 'exit') { await new Promise(r => setTimeout(r, 50)); process.exit(0); }
}

main().catch(e => { console.error(e && e.stack || e); process.exit(1); });
