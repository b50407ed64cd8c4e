'use strict';
// Seeded synthetic siting domains for the benchmark workloads: Gaussian-hill terrain, axis-aligned box buildings
// rasterised by exact area overlap, a random subset of heat-source roofs, per-column roughness. The output has the
// shape of an aggregated grid (dem, frac, hmean, dcarea, cd; north-up), so the reference's build_domain rules
// (js/live/domain.js:166-190, restated in domain.js) turn it into solver columns exactly as for the Ashburn inputs.

const { dragCoefficient, overlapCells } = require('./domain');

/** mulberry32: 32-bit state, returns uniform [0, 1) doubles */
function mulberry32(seed) {
  let a = seed >>> 0;
  return function next() {
    a = (a + 0x6D2B79F5) >>> 0;
    let t = a;
    t = Math.imul(t ^ (t >>> 15), t | 1);
    t ^= t + Math.imul(t ^ (t >>> 7), t | 61);
    return ((t ^ (t >>> 14)) >>> 0) / 4294967296;
  };
}

const uniform = (rng, a, b) => a + (b - a) * rng();

const Z0_CHOICES = [0.03, 0.1, 0.5, 1.0];

/** The named workloads. Grids share one 6 x 6 km box unless stated (dx = 6000 / nx). */
const WORKLOADS = {
  // plain method on the reference CPU path: 200 x 200 x 32 at 30 m
  c1: { nx: 200, ny: 200, nz: 32, dx: 30, hills: 12, buildings: 1500, sources: 120, flux: 400,
    met: { u10: 5, dirFrom: 225 }, alpha: 0.2, steps: 2000, meanSeconds: 600, seed: 1 },
  // plain drop-in on one GPU: 600 x 600 x 64 at 10 m (23 M cells)
  c2: { nx: 600, ny: 600, nz: 64, dx: 10, hills: 12, buildings: 4000, sources: 210, flux: 400,
    met: { u10: 4, dirFrom: 315 }, alpha: 0.2, steps: 20000, meanSeconds: 600, seed: 2 },
  // timing reference: the same grid and met without terrain, buildings or source roofs (every cell on the fast path)
  c2flat: { nx: 600, ny: 600, nz: 64, dx: 10, hills: 0, buildings: 0, sources: 0, flux: 400,
    met: { u10: 4, dirFrom: 315 }, alpha: 0.2, steps: 20000, meanSeconds: 600, seed: 2 },
  // reduced-precision storage: 1200 x 1200 x 96 at 5 m (138 M cells)
  c3: { nx: 1200, ny: 1200, nz: 96, dx: 5, hills: 12, buildings: 16000, sources: 850, flux: 400,
    met: { u10: 4, dirFrom: 315 }, alpha: 0.2, steps: 5000, meanSeconds: 600, seed: 3 },
};

/** grids of the batched and served workloads (6 km box, dx = 6000 / nx) */
const BATCH_GRIDS = [[100, 100, 32], [150, 150, 32], [200, 200, 48], [300, 300, 48]];
/** the batched uneven-domain workload: 64 domains with grids drawn uniformly from BATCH_GRIDS, wind 2-9 m/s from any
 *  direction, the 6 km box's seeded buildings and source roofs, one inserted 120 MW campus of 300 x 200 m each,
 *  3,000 steps */
const BATCH_WORKLOAD = { domains: 64, grids: BATCH_GRIDS, u10: [2, 9], campusMW: 120, campus: [300, 200],
  buildings: 4000, sources: 210, flux: 400, steps: 3000, seed: 4 };
const SERVE_GRIDS = [[[100, 100, 32], 0.50], [[200, 200, 48], 0.35], [[600, 600, 64], 0.15]];

/**
 * spec: { nx, ny, dx, hills, buildings, sources, seed, baseM = 80, hillAmp = [0, 40], hillWidth = [300, 1500],
 *         side = [20, 150], height = [8, 30] }
 * Returns an aggregated-grid object (north-up) plus the building list and the origin of scene coordinates
 * (x east, z south, metres; the box spans [x0, x0 + nx dx) x [z0, z0 + ny dx)).
 */
function generate(spec) {
  const nx = spec.nx, ny = spec.ny, dx = spec.dx, n = nx * ny, cellA = dx * dx;
  const rng = mulberry32(spec.seed === undefined ? 1 : spec.seed);
  const Lx = nx * dx, Lz = ny * dx;
  const baseM = spec.baseM === undefined ? 80 : spec.baseM;
  const amp = spec.hillAmp || [0, 40], wid = spec.hillWidth || [300, 1500];
  const side = spec.side || [20, 150], height = spec.height || [8, 30];
  // terrain: base + sum of Gaussian hills (sigma = width / 2), sampled at the cell centres
  const hills = [];
  for (let h = 0; h < (spec.hills === undefined ? 12 : spec.hills); h++) {
    hills.push({ a: uniform(rng, amp[0], amp[1]), s: uniform(rng, wid[0], wid[1]) / 2, x: uniform(rng, 0, Lx), z: uniform(rng, 0, Lz) });
  }
  const dem = new Float64Array(n);
  for (let r = 0; r < ny; r++) {
    const z = (r + 0.5) * dx;
    for (let c = 0; c < nx; c++) {
      const x = (c + 0.5) * dx;
      let e = baseM;
      for (const h of hills) {
        const ddx = x - h.x, ddz = z - h.z;
        e += h.a * Math.exp(-(ddx * ddx + ddz * ddz) / (2 * h.s * h.s));
      }
      dem[r * nx + c] = e;
    }
  }
  // buildings: axis-aligned boxes, overlaps allowed
  const buildings = [];
  for (let b = 0; b < (spec.buildings || 0); b++) {
    const wx = uniform(rng, side[0], side[1]), wz = uniform(rng, side[0], side[1]);
    const cx = uniform(rng, 0, Lx), cz = uniform(rng, 0, Lz), hh = uniform(rng, height[0], height[1]);
    buildings.push({ xa: cx - wx / 2, xb: cx + wx / 2, za: cz - wz / 2, zb: cz + wz / 2, h: hh, source: false });
  }
  // heat-source roofs: a seeded random subset (partial Fisher-Yates)
  const ns = Math.min(spec.sources || 0, buildings.length);
  const idx = buildings.map((_, i) => i);
  for (let s = 0; s < ns; s++) {
    const t = s + Math.floor(rng() * (idx.length - s));
    const tmp = idx[s]; idx[s] = idx[t]; idx[t] = tmp;
    buildings[idx[s]].source = true;
  }
  const cover = new Float64Array(n), hsum = new Float64Array(n), dcarea = new Float64Array(n);
  for (const b of buildings) {
    overlapCells(b.xa, b.xb, b.za, b.zb, nx, ny, dx, (k, a) => {
      cover[k] += a; hsum[k] += a * b.h;
      if (b.source) dcarea[k] += a;
    });
  }
  const frac = new Float64Array(n), hmean = new Float64Array(n), cd = new Float64Array(n);
  for (let k = 0; k < n; k++) {
    frac[k] = Math.min(1, cover[k] / cellA);
    hmean[k] = cover[k] > 0 ? hsum[k] / cover[k] : 0;
    if (dcarea[k] > cellA) dcarea[k] = cellA;
    cd[k] = dragCoefficient(Z0_CHOICES[Math.floor(rng() * Z0_CHOICES.length)], dx);
  }
  let roof = 0;
  for (let k = 0; k < n; k++) roof += dcarea[k];
  return {
    nx, ny, dx, r: 1, dem, frac, hmean, dcarea, cd, buildings, hills,
    origin: { x0: -Lx / 2, z0: -Lz / 2 }, sourceRoofArea_m2: roof,
  };
}

/** a location for a w x h campus roof (scene metres): among 64 seeded candidates inside the central 80 % of the box,
 *  the one overlapping the least building footprint area (the first fully free one if any) */
function placeCampus(agg, w, h, seed) {
  const rng = mulberry32(seed);
  const Lx = agg.nx * agg.dx, Lz = agg.ny * agg.dx;
  let best = null, bestA = Infinity;
  for (let tries = 0; tries < 64 && bestA > 0; tries++) {
    const xa = uniform(rng, 0.1 * Lx, Math.max(0.1 * Lx, 0.9 * Lx - w)), za = uniform(rng, 0.1 * Lz, Math.max(0.1 * Lz, 0.9 * Lz - h));
    let a = 0;
    for (const b of agg.buildings || []) {
      const ox = Math.min(b.xb, xa + w) - Math.max(b.xa, xa), oz = Math.min(b.zb, za + h) - Math.max(b.za, za);
      if (ox > 0 && oz > 0) a += ox * oz;
    }
    if (a < bestA) { bestA = a; best = [xa, za]; }
  }
  return [agg.origin.x0 + best[0], agg.origin.z0 + best[1], agg.origin.x0 + best[0] + w, agg.origin.z0 + best[1] + h];
}

module.exports = { mulberry32, generate, placeCampus, overlapCells, WORKLOADS, BATCH_GRIDS, BATCH_WORKLOAD, SERVE_GRIDS,
  Z0_CHOICES };
