'use strict';
// Geometry of the solver grid, from the Ashburn 10 m input rasters (static10) or from an aggregated synthetic grid.
//
// Three stages, each a pass over whole planes:
//   1. decodePlanes: the packed static10 file is a list of planes described by its JSON header (offset, dtype,
//      encoding); each is decoded by the routine its encoding names, into integer planes (centimetres, classes,
//      25ths, 4 m2 units), and checked against the header's integer checksums.
//   2. aggregate (blockSums): every plane is reduced to the coarse grid by streaming its rows once: each 10 m row adds its values
//      into the accumulators of the coarse cells it crosses. Within a coarse cell the values arrive row by row, left
//      to right, so floating sums come out the same whatever the pass structure.
//   3. buildDomain (columnLevels): per column, the number of terrain-solid levels (kt), building levels (nb) and the first fluid
//      level (ktop = kt + nb), from the coarse terrain height, building cover and height, and the roof area of new
//      campuses (rasterised by exact rectangle overlap, overlapCells).
// The rules of stage 3 (zbase, the half-level terrain rounding, the 50 % cover rule with half-to-even height
// rounding, the 14 m minimum campus height, nz from the lid height with 8 free levels above the tallest column), the
// roughness table and the static10 format are the reference's (js/live/domain.js:9-10, :32-69, :166-190;
// data/live/static10.json): the recorded domain facts of its validation runs pin them.
//
// Planes are stored north-up: index row * nx + col, row 0 on the north edge of the box, col 0 on the west edge; the
// box's north-west corner is at scene (x, z) = (-2995, -2995) m, z pointing south. The solver numbers rows from the
// south (flipRows converts, in either direction).

const fs = require('fs');
const path = require('path');
const zlib = require('zlib');
const { KAPPA } = require('./physics');

const SCENE_X0 = -2995, SCENE_Z0 = -2995, AOI_SIZE = 6000, BASE_ELEV = 83.53166;
// roughness length z0 (m) per ESA WorldCover class; classes not listed take Z0_DEFAULT
const Z0_BY_CLASS = { 10: 1.0, 20: 0.3, 30: 0.03, 40: 0.05, 50: 0.5, 60: 0.01, 80: 0.0002, 90: 0.05 };
const Z0_DEFAULT = 0.1;

// ---- stage 1: static10 ---------------------------------------------------------------------------------------------

// little-endian integer readers by dtype
const READERS = {
  uint8: { size: 1, Out: Uint8Array, at: (b, i) => b[i] },
  uint16: { size: 2, Out: Uint16Array, at: (b, i) => b[2 * i] | (b[2 * i + 1] << 8) },
  int16: { size: 2, Out: Int16Array, at: (b, i) => ((b[2 * i] | (b[2 * i + 1] << 8)) << 16) >> 16 },
};

// plane decoders by encoding name: (bytes of the plane, dtype, rows, cols) -> integer plane
const DECODERS = {
  raw(bytes, dtype, rows, cols) {
    const R = READERS[dtype], n = rows * cols, out = new R.Out(n);
    for (let i = 0; i < n; i++) out[i] = R.at(bytes, i);
    return out;
  },
  // signed 16-bit differences along each row, stored as a plane of low bytes followed by a plane of high bytes;
  // the values are the running sums of each row (exceed 16 bits, hence Int32)
  'rowdelta-byteplanes'(bytes, dtype, rows, cols) {
    if (dtype !== 'int16') throw new Error(`static10: rowdelta planes are int16 (got ${dtype})`);
    const n = rows * cols, lo = bytes.subarray(0, n), hi = bytes.subarray(n, 2 * n), out = new Int32Array(n);
    for (let r = 0; r < rows; r++) {
      let run = 0;
      for (let i = r * cols, end = i + cols; i < end; i++) {
        run += ((lo[i] | (hi[i] << 8)) << 16) >> 16;
        out[i] = run;
      }
    }
    return out;
  },
};

// the header's checksum of each plane (sum of its integer values)
const CHECKSUM_KEY = { dem: 'dem_sum_cm', lc: 'lc_sum', frac25: 'frac25_sum', hmean: 'hmean_sum_cm', dca4: 'dca4_sum' };

function decodePlanes(meta, bytes) {
  const rows = meta.grid.nrows, cols = meta.grid.ncols, planes = {};
  for (const [name, P] of Object.entries(meta.planes)) {
    const dec = DECODERS[P.encoding];
    if (!dec) throw new Error(`static10: plane ${name} has an unknown encoding ${P.encoding}`);
    planes[name] = dec(bytes.subarray(P.offset, P.offset + P.bytes), P.dtype, rows, cols);
    const key = CHECKSUM_KEY[name], want = meta.checksums && key ? meta.checksums[key] : undefined;
    if (want !== undefined) {
      let sum = 0;
      for (const v of planes[name]) sum += v;
      if (sum !== want) throw new Error(`static10 checksum mismatch (${key}: ${sum}, header ${want})`);
    }
  }
  return planes;
}

const loaded = new Map();

/** the static10 rasters of a directory (static10.json + its packed file, gzip or not), decoded once per directory:
 *  { rows, cols, meta, planes: { dem (cm), lc (class), frac25 (cover in 25ths), hmean (cm), dca4 (4 m2 units) } } */
function loadStatic10(dir) {
  const key = path.resolve(dir);
  if (!loaded.has(key)) {
    const meta = JSON.parse(fs.readFileSync(path.join(key, 'static10.json'), 'utf8'));
    let buf = fs.readFileSync(path.join(key, meta.file || 'static10.bin.gz'));
    if (buf.length >= 2 && buf[0] === 0x1f && buf[1] === 0x8b) buf = zlib.gunzipSync(buf);
    if (buf.length !== meta.raw_bytes) throw new Error(`static10 has ${buf.length} bytes, expected ${meta.raw_bytes}`);
    const bytes = new Uint8Array(buf.buffer, buf.byteOffset, buf.length);
    loaded.set(key, { rows: meta.grid.nrows, cols: meta.grid.ncols, meta, planes: decodePlanes(meta, bytes) });
  }
  return loaded.get(key);
}

// ---- helpers -------------------------------------------------------------------------------------------------------

/** nearest integer, ties to the even one (numpy's rounding) */
function roundHalfEven(x) {
  const f = Math.floor(x), d = x - f;
  if (d !== 0.5) return d < 0.5 ? f : f + 1;
  return f % 2 === 0 ? f : f + 1;
}

/** north-up rows <-> solver rows (j = 0 on the south edge); applying it twice gives the input back */
function flipRows(a, nx, ny) {
  const out = new a.constructor(a.length);
  for (let j = 0; j < ny; j++) out.set(a.subarray((ny - 1 - j) * nx, (ny - j) * nx), j * nx);
  return out;
}

/** log-law drag coefficient (kappa / ln(z / z0))^2 at the first cell's centre z = dx / 2 */
function dragCoefficient(z0, dx) { return Math.pow(KAPPA / Math.log((0.5 * dx) / z0), 2); }

// ---- stage 2: coarsening -------------------------------------------------------------------------------------------

/** sums of value(i) over r x r blocks of a rows x cols plane, row by row (see the header) */
function blockSums(rows, cols, r, value) {
  const nx = cols / r, ny = rows / r, acc = new Float64Array(nx * ny);
  for (let y = 0; y < rows; y++) {
    const rowAcc = ((y / r) | 0) * nx;
    for (let x = 0, i = y * cols; x < cols; x++, i++) acc[rowAcc + ((x / r) | 0)] += value(i);
  }
  return acc;
}

/** the 10 m rasters on the dx grid: mean terrain height (m), building cover fraction, cover-weighted mean building
 *  height (m), data-center roof area (m2), and the block mean of the drag coefficient (the stress-conserving mean)
 *  of every 10 m cell's roughness class, evaluated at dx / 2 */
function aggregate(st, dx) {
  const r = Math.round(dx / 10);
  if (Math.abs(r * 10 - dx) > 1e-9 || st.cols % r || st.rows % r)
    throw new Error(`dx must be a multiple of 10 m dividing 6000 m (got ${dx})`);
  const P = st.planes, rows = st.rows, cols = st.cols, nx = cols / r, ny = rows / r, area = r * r;
  const cdOfClass = new Float64Array(256);
  for (let c = 0; c < 256; c++) cdOfClass[c] = dragCoefficient(Z0_BY_CLASS[c] !== undefined ? Z0_BY_CLASS[c] : Z0_DEFAULT, dx);
  const demSum = blockSums(rows, cols, r, i => P.dem[i]);
  const cover = blockSums(rows, cols, r, i => P.frac25[i]);
  const heightCover = blockSums(rows, cols, r, i => P.frac25[i] * P.hmean[i]);
  const roof = blockSums(rows, cols, r, i => P.dca4[i]);
  const cdSum = blockSums(rows, cols, r, i => cdOfClass[P.lc[i]]);
  const n = nx * ny;
  const out = { nx, ny, dx, r, dem: new Float64Array(n), cnt: new Int32Array(n), frac: new Float64Array(n),
    hmean: new Float64Array(n), dcarea: new Float64Array(n), cd: new Float64Array(n) };
  for (let k = 0; k < n; k++) {
    out.dem[k] = demSum[k] / (100 * area);
    out.cnt[k] = cover[k];
    out.frac[k] = cover[k] / (25 * area);
    out.hmean[k] = cover[k] > 0 ? heightCover[k] / cover[k] / 100 : 0;
    out.dcarea[k] = 4 * roof[k];
    out.cd[k] = cdSum[k] / area;
  }
  return out;
}

// ---- campuses ------------------------------------------------------------------------------------------------------

/** exact overlap of the rectangle [xa,xb) x [za,zb) (metres from the north-west corner, z southward) with the grid;
 *  calls fn(cell, area) per covered north-up cell */
function overlapCells(xa, xb, za, zb, nx, ny, dx, fn) {
  const X1 = nx * dx, Z1 = ny * dx;
  const x0 = Math.max(xa, 0), x1 = Math.min(xb, X1), z0 = Math.max(za, 0), z1 = Math.min(zb, Z1);
  if (x1 <= x0 || z1 <= z0) return;
  const c0 = Math.floor(x0 / dx), c1 = Math.min(nx - 1, Math.floor(x1 / dx));
  const r0 = Math.floor(z0 / dx), r1 = Math.min(ny - 1, Math.floor(z1 / dx));
  for (let r = r0; r <= r1; r++) {
    const oz = Math.min(z1, (r + 1) * dx) - Math.max(z0, r * dx);
    if (oz <= 0) continue;
    for (let c = c0; c <= c1; c++) {
      const ox = Math.min(x1, (c + 1) * dx) - Math.max(x0, c * dx);
      if (ox <= 0) continue;
      fn(r * nx + c, ox * oz);
    }
  }
}


/** new data-center campuses on the grid: list = [{ id, mw, halls: [[x0, z0, x1, z1], ...] }] in scene metres, origin
 *  = the box's north-west corner (default the Ashburn box). Each campus spreads its power evenly over the total area
 *  of its halls (parts outside the box are dropped, with their share of the power); overlap areas are exact and
 *  clipped to one cell's area per campus and in total. Returns north-up roof area (m2) and heat (W) per cell, and per
 *  campus { id, mw, roofArea_m2, inDomain_m2, flux_wm2, heatMW, cells }. */
function rasterizeCampuses(list, dx, nx, ny, origin) {
  const X0 = origin ? origin.x0 : SCENE_X0, Z0 = origin ? origin.z0 : SCENE_Z0;
  const n = nx * ny, cellA = dx * dx, newArea = new Float64Array(n), heatW = new Float64Array(n), info = [];
  for (const cp of list || []) {
    const covered = new Map();     // cell -> overlap area of this campus, cells in order of first touch
    let hallArea = 0;
    for (const h of cp.halls || []) {
      const xa = Math.min(h[0], h[2]), xb = Math.max(h[0], h[2]), za = Math.min(h[1], h[3]), zb = Math.max(h[1], h[3]);
      if (!(xb > xa && zb > za)) continue;
      hallArea += (xb - xa) * (zb - za);
      overlapCells(xa - X0, xb - X0, za - Z0, zb - Z0, nx, ny, dx, (k, a) => covered.set(k, (covered.get(k) || 0) + a));
    }
    const mw = Math.max(0, +cp.mw || 0), flux = hallArea > 0 ? mw * 1e6 / hallArea : 0;
    let inside = 0;
    for (const [k, a] of covered) {
      const ac = Math.min(a, cellA);
      newArea[k] += ac;
      heatW[k] += ac * flux;
      inside += ac;
    }
    info.push({ id: cp.id, mw, roofArea_m2: hallArea, inDomain_m2: inside, flux_wm2: flux, heatMW: inside * flux / 1e6,
      cells: covered.size });
  }
  for (let k = 0; k < n; k++) newArea[k] = Math.min(newArea[k], cellA);
  return { newArea, heatW, info };
}

// ---- stage 3: columns ----------------------------------------------------------------------------------------------

/** solid levels of one column: terrain (kt >= 1) and buildings (nb, when at least half the cell is covered) */
function columnLevels(dem, frac, hmean, zbase, dx) {
  const kt = Math.max(1, Math.ceil((dem - zbase) / dx - 0.5));
  const nb = frac >= 0.5 ? roundHalfEven(hmean / dx) : 0;
  return { kt, nb };
}

/** solver columns of an aggregated grid. opts: { heightM = 720 (lid height above zbase), newArea = null (north-up
 *  m2 of new campus roof per cell: cover at least newArea / dx^2, height at least extraHeight), extraHeight = 14,
 *  nzMin = 0, nz = null (fixed level count) }; nz grows to leave 8 free levels above the tallest column. Returns
 *  north-up kt, nb, ktop = kt + nb, cd (f32), nz, zbase (height of level 0's bottom), and counts. */
function buildDomain(agg, opts) {
  const o = opts || {};
  const heightM = o.heightM === undefined ? 720 : o.heightM, extraHeight = o.extraHeight === undefined ? 14.0 : o.extraHeight;
  const newArea = o.newArea || null, nzMin = o.nzMin || 0;
  const { nx, ny, dx } = agg, n = nx * ny, cellA = dx * dx;
  let lowest = Infinity;
  for (const e of agg.dem) lowest = Math.min(lowest, e);
  const zbase = Math.floor(lowest) - dx;
  const kt = new Int32Array(n), nb = new Int32Array(n), ktop = new Int32Array(n), cd = Float32Array.from(agg.cd);
  for (let k = 0; k < n; k++) {
    const campus = newArea !== null && newArea[k] > 0;
    const frac = campus ? Math.max(agg.frac[k], newArea[k] / cellA) : agg.frac[k];
    const hmean = campus ? Math.max(agg.hmean[k], extraHeight) : agg.hmean[k];
    const c = columnLevels(agg.dem[k], frac, hmean, zbase, dx);
    kt[k] = c.kt;
    nb[k] = c.nb;
    ktop[k] = c.kt + c.nb;
  }
  let ktopMax = 0, nbCells = 0, nbColumns = 0;
  for (let k = 0; k < n; k++) {
    ktopMax = Math.max(ktopMax, ktop[k]);
    nbCells += nb[k];
    nbColumns += nb[k] > 0 ? 1 : 0;
  }
  let nz = Math.max(o.nz ? o.nz : Math.round(heightM / dx), nzMin);
  if (nz - 8 <= ktopMax) nz = ktopMax + 9;
  return { nx, ny, nz, dx, zbase, kt, nb, ktop, cd, ktopMax, nbCells, nbColumns, dcarea: agg.dcarea, dem: agg.dem };
}

module.exports = {
  SCENE_X0, SCENE_Z0, AOI_SIZE, BASE_ELEV, Z0_BY_CLASS, Z0_DEFAULT,
  loadStatic10, roundHalfEven, flipRows, dragCoefficient, aggregate, overlapCells, rasterizeCampuses, buildDomain,
};
