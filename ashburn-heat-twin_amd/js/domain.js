'use strict';
// Solver geometry: the packed 10 m Ashburn inputs (static10), their block aggregation to dx, the build_domain
// rules (kt, nb, ktop, nz, zbase), campus rasterisation and the north-up <-> solver row flip.
// Semantics follow js/live/domain.js:32-190 of the reference; this is a Node (fs + zlib) implementation.
//
// Conventions: 2-D arrays returned here are north-up, [row*nx + col], row 0 = north edge (scene z = -2995),
// col 0 = west edge (scene x = -2995). The solver flips rows (j = 0 at the south edge) when it uploads them.

const fs = require('fs');
const path = require('path');
const zlib = require('zlib');
const { KAPPA } = require('./physics');

const SCENE_X0 = -2995, SCENE_Z0 = -2995, AOI_SIZE = 6000, BASE_ELEV = 83.53166;
// WorldCover class -> roughness length z0 (m); the reference's table (js/live/domain.js:9-10)
const Z0_BY_CLASS = { 10: 1.0, 20: 0.3, 30: 0.03, 40: 0.05, 50: 0.5, 60: 0.01, 80: 0.0002, 90: 0.05 };
const Z0_DEFAULT = 0.1;

const staticCache = new Map();

/** load and decode static10.{json,bin.gz} from a directory; integer-exact planes, checksums verified
 *  (format: data/live/static10.json; decoder semantics js/live/domain.js:32-69) */
function loadStatic10(dir) {
  const key = path.resolve(dir);
  if (staticCache.has(key)) return staticCache.get(key);
  const meta = JSON.parse(fs.readFileSync(path.join(key, 'static10.json'), 'utf8'));
  let raw = fs.readFileSync(path.join(key, meta.file || 'static10.bin.gz'));
  if (raw[0] === 0x1f && raw[1] === 0x8b) raw = zlib.gunzipSync(raw);
  if (raw.length !== meta.raw_bytes) throw new Error(`static10 has ${raw.length} bytes, expected ${meta.raw_bytes}`);
  const ny = meta.grid.nrows, nx = meta.grid.ncols, n = nx * ny, P = meta.planes;
  const bytes = new Uint8Array(raw.buffer, raw.byteOffset, raw.length);
  const demCm = new Int32Array(n);
  const lo = P.dem.offset, hi = P.dem.offset + n;
  for (let r = 0; r < ny; r++) {
    let acc = 0;
    for (let c = 0; c < nx; c++) {
      const k = r * nx + c;
      let d = bytes[lo + k] | (bytes[hi + k] << 8);
      if (d >= 32768) d -= 65536;
      acc += d;
      demCm[k] = acc;
    }
  }
  const hm = new Uint16Array(n);
  for (let k = 0; k < n; k++) hm[k] = bytes[P.hmean.offset + 2 * k] | (bytes[P.hmean.offset + 2 * k + 1] << 8);
  const st = {
    nrows: ny, ncols: nx, meta, demCm,
    lc: bytes.slice(P.lc.offset, P.lc.offset + n),
    frac25: bytes.slice(P.frac25.offset, P.frac25.offset + n),
    hmeanCm: hm,
    dca4: bytes.slice(P.dca4.offset, P.dca4.offset + n),
  };
  const sum = a => { let s = 0; for (let i = 0; i < a.length; i++) s += a[i]; return s; };
  const ck = meta.checksums || {};
  const got = { dem_sum_cm: sum(demCm), lc_sum: sum(st.lc), frac25_sum: sum(st.frac25), hmean_sum_cm: sum(st.hmeanCm), dca4_sum: sum(st.dca4) };
  for (const k of Object.keys(got)) {
    if (ck[k] !== undefined && got[k] !== ck[k]) throw new Error(`static10 checksum mismatch (${k})`);
  }
  staticCache.set(key, st);
  return st;
}

/** numpy rounding (half to even) (js/live/domain.js:73-76) */
function roundHalfEven(x) {
  const r = Math.round(x);
  return Math.abs(x - Math.trunc(x)) === 0.5 ? 2 * Math.round(x / 2) : r;
}

/** north-up [row][col] <-> solver [j][i] (j = 0 south); its own inverse (js/live/domain.js:79-83) */
function flipRows(a, nx, ny) {
  const o = new a.constructor(a.length);
  for (let r = 0; r < ny; r++) o.set(a.subarray(r * nx, (r + 1) * nx), (ny - 1 - r) * nx);
  return o;
}

/** drag coefficient of a roughness length at the first-cell centre 0.5 dx */
function dragCoefficient(z0, dx) { return Math.pow(KAPPA / Math.log((0.5 * dx) / z0), 2); }

/** 10 m -> dx block aggregation (js/live/domain.js:89-118): dem block mean, frac block mean, area-weighted hmean,
 *  dcarea block sum, stress-conserving mean drag coefficient */
function aggregate(st, dx) {
  const r = Math.round(dx / 10);
  if (Math.abs(r * 10 - dx) > 1e-9 || st.ncols % r || st.nrows % r) throw new Error(`dx must be a multiple of 10 m dividing 6000 m (got ${dx})`);
  const nx = st.ncols / r, ny = st.nrows / r, n = nx * ny, NC = st.ncols;
  const cdTab = new Float64Array(256);
  for (let c = 0; c < 256; c++) {
    const z0 = Z0_BY_CLASS[c] !== undefined ? Z0_BY_CLASS[c] : Z0_DEFAULT;
    cdTab[c] = dragCoefficient(z0, dx);
  }
  const dem = new Float64Array(n), cnt = new Int32Array(n), frac = new Float64Array(n), hmean = new Float64Array(n),
    dcarea = new Float64Array(n), cd = new Float64Array(n);
  for (let R = 0; R < ny; R++) {
    for (let C = 0; C < nx; C++) {
      let sd = 0, sc = 0, sh = 0, sa = 0, scd = 0;
      for (let a = 0; a < r; a++) {
        const base = (R * r + a) * NC + C * r;
        for (let b = 0; b < r; b++) {
          const k = base + b, v = st.frac25[k];
          sd += st.demCm[k]; sc += v; sh += v * st.hmeanCm[k]; sa += st.dca4[k]; scd += cdTab[st.lc[k]];
        }
      }
      const o = R * nx + C;
      dem[o] = sd / (100 * r * r);
      cnt[o] = sc;
      frac[o] = sc / (25 * r * r);
      hmean[o] = sc > 0 ? sh / sc / 100 : 0;
      dcarea[o] = 4 * sa;
      cd[o] = scd / (r * r);
    }
  }
  return { nx, ny, dx, r, dem, cnt, frac, hmean, dcarea, cd };
}

/** exact rectangle-cell overlap of campus halls [x0, z0, x1, z1] (scene metres) with the grid; each campus spreads
 *  its MW over its total hall area (js/live/domain.js:124-160). Returns north-up newArea (m2) and heatW (W). */
function rasterizeCampuses(list, dx, nx, ny, origin) {
  const X0 = origin ? origin.x0 : SCENE_X0, Z0 = origin ? origin.z0 : SCENE_Z0;
  const n = nx * ny, newArea = new Float64Array(n), heatW = new Float64Array(n), info = [];
  const X1 = X0 + nx * dx, Z1 = Z0 + ny * dx, cellA = dx * dx;
  for (const cp of list || []) {
    const per = new Map();
    let total = 0;
    for (const h of cp.halls || []) {
      const xa = Math.min(h[0], h[2]), xb = Math.max(h[0], h[2]), za = Math.min(h[1], h[3]), zb = Math.max(h[1], h[3]);
      if (!(xb > xa && zb > za)) continue;
      total += (xb - xa) * (zb - za);
      const x0 = Math.max(xa, X0), x1 = Math.min(xb, X1), z0 = Math.max(za, Z0), z1 = Math.min(zb, Z1);
      if (x1 <= x0 || z1 <= z0) continue;
      const c0 = Math.max(0, Math.floor((x0 - X0) / dx)), c1 = Math.min(nx - 1, Math.floor((x1 - X0) / dx));
      const r0 = Math.max(0, Math.floor((z0 - Z0) / dx)), r1 = Math.min(ny - 1, Math.floor((z1 - Z0) / dx));
      for (let r = r0; r <= r1; r++) {
        const oz = Math.min(z1, Z0 + (r + 1) * dx) - Math.max(z0, Z0 + r * dx);
        if (oz <= 0) continue;
        for (let c = c0; c <= c1; c++) {
          const ox = Math.min(x1, X0 + (c + 1) * dx) - Math.max(x0, X0 + c * dx);
          if (ox <= 0) continue;
          const k = r * nx + c;
          per.set(k, (per.get(k) || 0) + ox * oz);
        }
      }
    }
    const mw = Math.max(0, +cp.mw || 0);
    const flux = total > 0 ? mw * 1e6 / total : 0;
    let inDom = 0;
    for (const [k, a0] of per) {
      const a = Math.min(a0, cellA);
      newArea[k] += a; heatW[k] += a * flux; inDom += a;
    }
    info.push({ id: cp.id, mw, roofArea_m2: total, inDomain_m2: inDom, flux_wm2: flux, heatMW: inDom * flux / 1e6, cells: per.size });
  }
  for (let k = 0; k < n; k++) if (newArea[k] > cellA) newArea[k] = cellA;
  return { newArea, heatW, info };
}

/** build_domain rules on an aggregated grid (js/live/domain.js:166-190). opts: { heightM = 720, newArea = null,
 *  extraHeight = 14, nzMin = 0, nz = null (fixed level count; raised like heightM if ktop leaves < 9 levels) } */
function buildDomain(agg, opts) {
  opts = opts || {};
  const heightM = opts.heightM === undefined ? 720 : opts.heightM;
  const newArea = opts.newArea || null;
  const extraHeight = opts.extraHeight === undefined ? 14.0 : opts.extraHeight;
  const nzMin = opts.nzMin || 0;
  const { nx, ny, dx } = agg, n = nx * ny;
  let dmin = Infinity;
  for (let k = 0; k < n; k++) if (agg.dem[k] < dmin) dmin = agg.dem[k];
  const zbase = Math.floor(dmin) - dx;
  const kt = new Int32Array(n), nb = new Int32Array(n), ktop = new Int32Array(n), cd = new Float32Array(n);
  let ktopMax = 0, nbCells = 0, nbColumns = 0;
  for (let k = 0; k < n; k++) {
    let frac = agg.frac[k], hmean = agg.hmean[k];
    if (newArea && newArea[k] > 0) {
      frac = Math.max(frac, newArea[k] / (dx * dx));
      hmean = Math.max(hmean, extraHeight);
    }
    kt[k] = Math.max(1, Math.ceil((agg.dem[k] - zbase) / dx - 0.5));
    nb[k] = frac >= 0.5 ? roundHalfEven(hmean / dx) : 0;
    ktop[k] = kt[k] + nb[k];
    if (ktop[k] > ktopMax) ktopMax = ktop[k];
    nbCells += nb[k];
    if (nb[k] > 0) nbColumns++;
    cd[k] = agg.cd[k];
  }
  let nz = opts.nz ? Math.max(opts.nz, nzMin) : Math.max(Math.round(heightM / dx), nzMin);
  if (ktopMax >= nz - 8) nz = ktopMax + 9;
  return { nx, ny, nz, dx, zbase, kt, nb, ktop, cd, ktopMax, nbCells, nbColumns, dcarea: agg.dcarea, dem: agg.dem };
}

module.exports = {
  SCENE_X0, SCENE_Z0, AOI_SIZE, BASE_ELEV, Z0_BY_CLASS, Z0_DEFAULT,
  loadStatic10, roundHalfEven, flipRows, dragCoefficient, aggregate, rasterizeCampuses, buildDomain,
};
