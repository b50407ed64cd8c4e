'use strict';
// Interface behaviour of the Node host, after the reference's `api` test mode (tests/live/validate.html:281-328):
// rejection messages, reads without volume or wind, 8 concurrent reads while stepping, campus add/remove, flux,
// resets, flow re-initialisation on a large turn, use-after-destroy. Prints one JSON object.
// Usage: node tests/js/api_test.js [--cpu]   (--cpu: only the checks that need no GPU)
const path = require('path');
const { createLiveSolver: createSolver, loadAddon } = require(path.join(__dirname, '..', '..', 'ashburn-heat-twin_amd', 'js', 'solver.js'));
const { CAMPUSES, MET_AUG25 } = require(path.join(__dirname, '..', '..', 'ashburn-heat-twin_amd', 'js', 'dump.js'));

const expectReject = async p => { try { await p; return 'resolved (unexpected)'; } catch (e) { return 'rejected: ' + e.message; } };

async function main() {
  const cpuOnly = process.argv.includes('--cpu');
  const out = {};
  out.badDx = await expectReject(createSolver({ dx: 25 }));
  out.badU10Max = await expectReject(createSolver({ u10Max: 0 }));
  out.badHeight = await expectReject(createSolver({ heightM: 100 }));
  if (cpuOnly) {
    out.noGpu = await expectReject(createSolver({ dx: 40 }));
    console.log(JSON.stringify(out));
    return;
  }
  const s = await createSolver({ dx: 20, met: MET_AUG25, existingFlux: 400 });
  out.info = s.info;
  await s.step(300);
  const r0 = await s.read({ vol: false, wind: false });
  out.readNoVolNoWind = { surf: r0.surf.length, u: r0.u, v: r0.v, vol: r0.vol, thetaMax: r0.thetaMax, simTime: r0.simTime };
  const ps = [];
  for (let k = 0; k < 8; k++) { ps.push(s.read()); s.step(5); }
  const rs = await Promise.all(ps);
  out.concurrentReads = rs.map(r => +r.simTime.toFixed(2));
  out.concurrentSteps = rs.map(r => r.steps);
  await s.setNewCampuses([CAMPUSES.A, CAMPUSES.B]);
  const nf1 = s.info.nFluid, rA = await s.read();
  await s.step(1000);
  const stA = await s._dbg.stats();
  await s.setNewCampuses([]);
  const nf2 = s.info.nFluid;
  await s.step(1000);
  const stB = await s._dbg.stats(), rB = await s.read();
  out.campuses = { nFluidWith: nf1, nFluidWithout: nf2, heatMWwith: rA.heatMW, heatMWwithout: rB.heatMW,
    campusInfo: s.campusInfo, nonFinite: stA.nonFinite + stB.nonFinite, umaxDomain: Math.max(stA.umax, stB.umax) };
  await s.setExistingFlux(800);
  out.heatExistingMW800 = s.info.heatExistingMW;
  await s.resetHeat();
  const rH = await s.read();
  const windMean = r => { let sp = 0; for (let k = 0; k < r.u.length; k++) sp += Math.hypot(r.u[k], r.v[k]); return sp / r.u.length; };
  out.afterResetHeat = { simTime: rH.simTime, thetaMax: rH.thetaMax, windMean: windMean(rH) };
  await s.step(200);
  await s.resetFlow();
  const rF = await s.read();
  out.afterResetFlow = { simTime: rF.simTime, thetaMax: rF.thetaMax, windMean: windMean(rF) };
  await s.setMet({ u10: 4, dirFrom: 230 });
  const n1 = s.flowReinits;
  await s.setMet({ u10: 4, dirFrom: 30 });
  const n2 = s.flowReinits;
  out.flowReinits = { afterSmallTurn: n1, afterBigTurn: n2 };
  out.mlups = await s.benchmark(200);
  const tBefore = s.simTime, rBefore = await s.read();
  out.autotune = await s.autotune(2);
  const rAfter = await s.read();
  let same = rBefore.surf.length === rAfter.surf.length;
  for (let k = 0; same && k < rBefore.surf.length; k++) {
    same = rBefore.surf[k] === rAfter.surf[k] && rBefore.u[k] === rAfter.u[k];
  }
  out.autotuneKeepsState = same && s.simTime === tBefore;
  out.launch = s.launch;
  s.destroy();
  out.afterDestroy = await expectReject(Promise.resolve().then(() => s.step(1)));
  out.readAfterDestroy = await expectReject(Promise.resolve().then(() => s.read()));
  // destroy with a step in flight, as the reference's caller does (js/live.js:44, :76): the step issued before
  // destroy() settles, later calls reject
  const s2 = await createSolver({ dx: 40, met: MET_AUG25 });
  const inflight = s2.step(200);
  s2.destroy();
  out.stepBeforeDestroy = await inflight.then(() => 'resolved', e => 'rejected: ' + e.message);
  out.stepAfterDestroy = await expectReject(s2.step(1));
  // the same on the addon directly: destroy() while the step job runs on a worker is deferred to its completion
  const A = loadAddon(), sc = s2._sc, c = sc.col;
  const h = A.create({ nx: sc.nx, ny: sc.ny, nz: sc.nz, storage: 0, device: 0 });
  A.setColumns(h, c.kt, c.ktop, c.cd, c.q);
  A.setParams(h, sc.params, sc.profBlock());
  A.initFlow(h);
  const job = A.step(h, 300, 0);
  A.destroy(h);
  out.addonStepBeforeDestroy = await job.then(() => 'resolved', e => 'rejected: ' + e.message);
  try { A.step(h, 1, 0); out.addonStepAfterDestroy = 'no error (unexpected)'; } catch (e) { out.addonStepAfterDestroy = 'threw: ' + e.message; }
  console.log(JSON.stringify(out));
}

main().catch(e => { console.error(e && e.stack || e); process.exit(1); });
