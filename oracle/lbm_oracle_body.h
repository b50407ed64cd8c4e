/*
 * Body of the CPU oracle (TEST INFRASTRUCTURE ONLY), included once per type pair by lbm_oracle.c with
 * REAL (arithmetic type), STORE (population storage type) and SFX (name suffix) defined.
 *
 * Every expression keeps the operation order of the WGSL text that js/live/kernel.js generates: sums run left
 * to right in lattice order q = 0..18 (or 0..6), products associate left to right, literals are the f32-rounded
 * values the generator prints (fmt(), js/live/kernel.js:31-33) converted to REAL.
 */
#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* lattice weights: the generator prints fmt(w) = the f32-rounded weight (js/live/kernel.js:31-33, :357) */
#define LIT(x) ((REAL)(float)(x))
/* other literals (0.15, 1.0/3.0): exact constants, converted once to the arithmetic type */
#define CST(x) ((REAL)(x))

/* sqrt in the arithmetic type: sqrtf for float (correctly rounded, as WGSL on IEEE hardware), sqrt for double */
#define RSQRT(x) (sizeof(REAL) == 4 ? (REAL)sqrtf((float)(x)) : (REAL)sqrt((double)(x)))

/* solidity of cell (kk, column (si,sj)): flag != 0 <=> kk < ktop[column] (js/live/kernel.js:7-8, :95-98) */
static inline int FN(solid_)(const oracle_dom *d, int kk, int si, int sj) {
  return kk < (int)d->ktop[sj * d->nx + si];
}

/* js/live/kernel.js:122-148 (secOutflowPrelude) + :159-243 (secStreamF, secStreamG) for one cell:
 * pulls the 19 + 7 post-stream populations of cell (i,j,k) into f[] and g[]. */
static void FN(stream_cell_)(const oracle_dom *d, const STORE *fA, const STORE *gA, int i, int j, int k, REAL uI,
                             REAL vI, REAL f[19], REAL g[7]) {
  const int NX = d->nx, NY = d->ny, NZ = d->nz;
  const int64_t N = (int64_t)NX * NY * NZ;
  const int64_t nu = ((int64_t)k * NY + j) * NX + i;
  /* outflow prelude (js/live/kernel.js:122-148) */
  REAL rn = 1.0, uxn = 0.0, uyn = 0.0, uzn = 0.0;
  const int onb = (i == 0 && d->inflow_w == 0) || (i == NX - 1 && d->inflow_e == 0) ||
                  (j == 0 && d->inflow_s == 0) || (j == NY - 1 && d->inflow_n == 0);
  if (onb) {
    REAL s = (REAL)fA[nu];
    for (int q = 1; q < 19; q++) s = s + (REAL)fA[q * N + nu];
    rn = s;
    for (int comp = 0; comp < 3; comp++) {
      REAL m = 0.0;
      for (int q = 0; q < 19; q++) {
        if (C19[q][comp] == 1) m = m + (REAL)fA[q * N + nu];
        else if (C19[q][comp] == -1) m = m - (REAL)fA[q * N + nu];
      }
      m = m / rn;
      if (comp == 0) uxn = m; else if (comp == 1) uyn = m; else uzn = m;
    }
  }
  const REAL usqn = uxn * uxn + uyn * uyn + uzn * uzn;
  REAL uxc = uxn, uyc = uyn;
  if (d->oclip != 0) {
    if (i == 0 && d->inflow_w == 0 && uxc > 0.0) uxc = 0.0;
    if (i == NX - 1 && d->inflow_e == 0 && uxc < 0.0) uxc = 0.0;
    if (j == 0 && d->inflow_s == 0 && uyc > 0.0) uyc = 0.0;
    if (j == NY - 1 && d->inflow_n == 0 && uyc < 0.0) uyc = 0.0;
  }
  const REAL usqc = uxc * uxc + uyc * uyc + uzn * uzn;

  f[0] = (REAL)fA[nu];
  for (int q = 1; q < 19; q++) {
    const int cx = C19[q][0], cy = C19[q][1], cz = C19[q][2];
    const int si = i - cx, sj = j - cy, sk = k - cz;
    const int outx = (si < 0) || (si >= NX), outy = (sj < 0) || (sj >= NY);
    REAL v;
    if (sk >= NZ) {
      /* free-slip lid: mirrored population of the lateral neighbour at the same level (:168-181) */
      const int m = MIRZ19[q];
      if (si < 0 || si >= NX || sj < 0 || sj >= NY) v = (REAL)fA[m * N + nu];
      else if (FN(solid_)(d, k, si, sj)) v = (REAL)fA[m * N + nu];
      else v = (REAL)fA[m * N + ((int64_t)k * NY + sj) * NX + si];
    } else if (outx || outy) {
      const int inflow = (cx == 1 && si < 0 && d->inflow_w != 0) || (cx == -1 && si >= NX && d->inflow_e != 0) ||
                         (cy == 1 && sj < 0 && d->inflow_s != 0) || (cy == -1 && sj >= NY && d->inflow_n != 0);
      const REAL w = LIT(W19[q]);
      if (inflow) {
        /* inflow face: equilibrium of rho = 1, U = (uI, vI, 0) (:187-189) */
        REAL cu = 0.0;
        if (cx) cu = cu + (REAL)cx * uI;
        if (cy) cu = cu + (REAL)cy * vI;
        v = w * (((REAL)1.0 + (REAL)3.0 * cu + (REAL)4.5 * cu * cu) - (REAL)1.5 * (uI * uI + vI * vI));
      } else {
        /* pressure outlet: non-equilibrium extrapolation towards rho = 1 (:190-205) */
        REAL cu = 0.0;
        if (cx) cu = cu + (REAL)cx * uxn;
        if (cy) cu = cu + (REAL)cy * uyn;
        if (cz) cu = cu + (REAL)cz * uzn;
        const REAL own = (REAL)fA[q * N + nu];
        if (d->oclip != 0) {
          REAL cuc = 0.0;
          if (cx) cuc = cuc + (REAL)cx * uxc;
          if (cy) cuc = cuc + (REAL)cy * uyc;
          if (cz) cuc = cuc + (REAL)cz * uzn;
          const REAL ec = ((REAL)1.0 + (REAL)3.0 * cuc + (REAL)4.5 * cuc * cuc) - (REAL)1.5 * usqc;
          const REAL en = ((REAL)1.0 + (REAL)3.0 * cu + (REAL)4.5 * cu * cu) - (REAL)1.5 * usqn;
          v = own + w * (ec - rn * en);
        } else {
          const REAL en = ((REAL)1.0 + (REAL)3.0 * cu + (REAL)4.5 * cu * cu) - (REAL)1.5 * usqn;
          v = own + w * ((REAL)1.0 - rn) * en;
        }
      }
    } else {
      /* interior: half-way bounce-back from a solid source (:207-213) */
      if (FN(solid_)(d, sk, si, sj)) v = (REAL)fA[OPP19[q] * N + nu];
      else v = (REAL)fA[q * N + ((int64_t)sk * NY + sj) * NX + si];
    }
    f[q] = v;
  }

  /* D3Q7 (:219-243) */
  g[0] = (REAL)gA[nu];
  for (int q = 1; q < 7; q++) {
    const int cx = C7[q][0], cy = C7[q][1], cz = C7[q][2];
    const int si = i - cx, sj = j - cy, sk = k - cz;
    REAL v;
    if (sk >= NZ) {
      v = (REAL)gA[q * N + nu];                       /* zero-gradient top */
    } else if (si < 0 || si >= NX || sj < 0 || sj >= NY) {
      const int inflow = (cx == 1 && si < 0 && d->inflow_w != 0) || (cx == -1 && si >= NX && d->inflow_e != 0) ||
                         (cy == 1 && sj < 0 && d->inflow_s != 0) || (cy == -1 && sj >= NY && d->inflow_n != 0);
      v = inflow ? (REAL)0.0 : (REAL)gA[q * N + nu];
    } else {
      if (FN(solid_)(d, sk, si, sj)) v = (REAL)gA[OPP7[q] * N + nu];
      else v = (REAL)gA[q * N + ((int64_t)sk * NY + sj) * NX + si];
    }
    g[q] = v;
  }
}

/* first moments and theta (:264-269) */
static void FN(moments1_)(const REAL f[19], const REAL g[7], REAL *rho, REAL *jx, REAL *jy, REAL *jz, REAL *th) {
  REAL r = f[0];
  for (int q = 1; q < 19; q++) r = r + f[q];
  *rho = r;
  REAL m[3];
  for (int c = 0; c < 3; c++) {
    REAL s = 0.0;
    for (int q = 0; q < 19; q++) {
      if (C19[q][c] == 1) s = s + f[q];
      else if (C19[q][c] == -1) s = s - f[q];
    }
    m[c] = s;
  }
  *jx = m[0]; *jy = m[1]; *jz = m[2];
  REAL t = g[0];
  for (int q = 1; q < 7; q++) t = t + g[q];
  *th = t;
}

/* second moment Pi_pr (:254-262, :271-276) */
static REAL FN(mom2_)(const REAL f[19], int p, int r) {
  REAL s = 0.0;
  for (int q = 0; q < 19; q++) {
    const int v = C19[q][p] * C19[q][r];
    if (v == 1) s = s + f[q];
    else if (v == -1) s = s - f[q];
  }
  return s;
}

typedef struct {
  REAL Fx, Fy, Fz, sg, ux, uy, uz;
} FN(forces_t);

/* secForces (:279-306) */
static void FN(forces_)(const oracle_dom *d, int k, int kg, int ktc, float cdc, REAL uI, REAL vI, REAL rho, REAL jx,
                        REAL jy, REAL jz, REAL th, FN(forces_t) *o) {
  const int NZ = d->nz;
  REAL Fx = 0.0, Fy = 0.0;
  REAL Fz = rho * (REAL)d->gbeta * th;
  if (k == kg && kg == ktc) {
    const REAL ux0 = jx / rho, uy0 = jy / rho;
    const REAL sp = RSQRT(ux0 * ux0 + uy0 * uy0);
    const REAL cd = (REAL)cdc;
    Fx = -rho * cd * sp * ux0;
    Fy = -rho * cd * sp * uy0;
    const REAL lim = (REAL)0.5 * rho;
    if (cd * sp > (REAL)0.5) {
      Fx = -lim * ux0;
      Fy = -lim * uy0;
    }
  }
  REAL sg = 0.0;
  if (k >= NZ - d->nsp) {
    const REAL aa = ((REAL)(float)(k - (NZ - d->nsp)) + (REAL)1.0) / (REAL)(float)d->nsp;
    sg = (REAL)d->sigmax * aa * aa;
    Fx += rho * sg * (uI - jx / rho);
    Fy += rho * sg * (vI - jy / rho);
    Fz += -rho * sg * (jz / rho);
  }
  o->Fx = Fx; o->Fy = Fy; o->Fz = Fz; o->sg = sg;
  o->ux = (jx + (REAL)0.5 * Fx) / rho;
  o->uy = (jy + (REAL)0.5 * Fy) / rho;
  o->uz = (jz + (REAL)0.5 * Fz) / rho;
}

static int FN(zagl_)(int k, int ktc) { int z = k - ktc; return z < 0 ? 0 : z; }   /* secProfile (:112-120) */

void FN(oracle_step)(const oracle_dom *d, const STORE *fA, STORE *fB, const STORE *gA, STORE *gB, STORE *out4) {
  const int NX = d->nx, NY = d->ny, NZ = d->nz;
  const int64_t N = (int64_t)NX * NY * NZ;
  const REAL third = CST(1.0 / 3.0);
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
  for (int k = 0; k < NZ; k++) {
    for (int j = 0; j < NY; j++) {
      for (int i = 0; i < NX; i++) {
        const int col = j * NX + i;
        const int ktc = (int)d->kt[col], kg = (int)d->ktop[col];
        if (k < kg) continue;                                           /* solid (:413-415) */
        const int64_t nu = ((int64_t)k * NY + j) * NX + i;
        const int zagl = FN(zagl_)(k, ktc);
        const REAL uI = (REAL)d->prof[4 * zagl + 0], vI = (REAL)d->prof[4 * zagl + 1];
        REAL f[19], g[7];
        FN(stream_cell_)(d, fA, gA, i, j, k, uI, vI, f, g);
        REAL rho, jx, jy, jz, th;
        FN(moments1_)(f, g, &rho, &jx, &jy, &jz, &th);
        const REAL Pxx = FN(mom2_)(f, 0, 0), Pyy = FN(mom2_)(f, 1, 1), Pzz = FN(mom2_)(f, 2, 2);
        const REAL Pxy = FN(mom2_)(f, 0, 1), Pxz = FN(mom2_)(f, 0, 2), Pyz = FN(mom2_)(f, 1, 2);
        FN(forces_t) F;
        FN(forces_)(d, k, kg, ktc, d->cd[col], uI, vI, rho, jx, jy, jz, th, &F);
        const REAL Fx = F.Fx, Fy = F.Fy, Fz = F.Fz, ux = F.ux, uy = F.uy, uz = F.uz, sg = F.sg;
        /* secRelax (:308-338) */
        const REAL usq = ux * ux + uy * uy + uz * uz;
        const REAL r3 = rho * third;
        const REAL Nxx = Pxx - r3 - rho * ux * ux;
        const REAL Nyy = Pyy - r3 - rho * uy * uy;
        const REAL Nzz = Pzz - r3 - rho * uz * uz;
        const REAL Nxy = Pxy - rho * ux * uy;
        const REAL Nxz = Pxz - rho * ux * uz;
        const REAL Nyz = Pyz - rho * uy * uz;
        const REAL Pn = RSQRT(Nxx * Nxx + Nyy * Nyy + Nzz * Nzz + (REAL)2.0 * (Nxy * Nxy + Nxz * Nxz + Nyz * Nyz));
        const REAL kb = (REAL)d->prof[4 * zagl + 2];
        REAL tb = (REAL)d->tau0 + (REAL)3.0 * (REAL)d->momfac * kb;
        int dmin = 1000;
        if (d->inflow_w == 0 && i < dmin) dmin = i;
        if (d->inflow_e == 0 && NX - 1 - i < dmin) dmin = NX - 1 - i;
        if (d->inflow_s == 0 && j < dmin) dmin = j;
        if (d->inflow_n == 0 && NY - 1 - j < dmin) dmin = NY - 1 - j;
        if (dmin < d->nsw) {
          const REAL rr = ((REAL)(float)d->nsw - (REAL)(float)dmin) / (REAL)(float)d->nsw;
          tb += CST(0.15) * rr * rr;
        }
        const REAL tau = (REAL)0.5 * (tb + RSQRT(tb * tb + (REAL)d->cs2smag * Pn / rho));
        const REAL nusgs = (tau - tb) * third;
        const REAL kap = nusgs / (REAL)d->prt + kb + (REAL)d->kfloor;
        const REAL taug = (REAL)0.5 + (REAL)4.0 * kap;
        const REAL om = (REAL)1.0 - (REAL)1.0 / tau;
        const REAL fpre = (REAL)1.0 - (REAL)0.5 / tau;
        const REAL trN = (Nxx + Nyy + Nzz) * third;
        const REAL Nd[3] = {Nxx, Nyy, Nzz};
        /* secCollision (:340-367) */
        for (int q = 0; q < 19; q++) {
          const int cx = C19[q][0], cy = C19[q][1], cz = C19[q][2];
          REAL cu = 0.0;
          if (cx) cu = cu + (REAL)cx * ux;
          if (cy) cu = cu + (REAL)cy * uy;
          if (cz) cu = cu + (REAL)cz * uz;
          const REAL w = LIT(W19[q]);
          const REAL feq = w * rho * (((REAL)1.0 + (REAL)3.0 * cu + (REAL)4.5 * cu * cu) - (REAL)1.5 * usq);
          /* (c c):N, terms in the generator's order: diagonal xx,yy,zz then cross xy,xz,yz */
          REAL qn = 0.0;
          int first = 1;
          for (int p = 0; p < 3; p++) {
            if (C19[q][p] != 0) { qn = first ? Nd[p] : qn + Nd[p]; first = 0; }
          }
          if (cx && cy) { const REAL t = (REAL)(2.0 * (cx * cy)) * Nxy; qn = first ? t : qn + t; first = 0; }
          if (cx && cz) { const REAL t = (REAL)(2.0 * (cx * cz)) * Nxz; qn = first ? t : qn + t; first = 0; }
          if (cy && cz) { const REAL t = (REAL)(2.0 * (cy * cz)) * Nyz; qn = first ? t : qn + t; first = 0; }
          const REAL fr = LIT(4.5 * W19[q]) * (qn - trN);
          REAL cF = 0.0;
          if (cx) cF = cF + (REAL)cx * Fx;
          if (cy) cF = cF + (REAL)cy * Fy;
          if (cz) cF = cF + (REAL)cz * Fz;
          const REAL Fq = fpre * w * ((REAL)3.0 * (cF - (ux * Fx + uy * Fy + uz * Fz)) + (REAL)9.0 * cu * cF);
          fB[q * N + nu] = (STORE)(feq + om * fr + Fq);
        }
        /* secScalarCollision (:369-386) */
        REAL S = 0.0;
        if (k == kg) S = (REAL)d->q[col] * (REAL)d->srcscale;
        const REAL omg = (REAL)1.0 / taug;
        for (int q = 0; q < 7; q++) {
          const int cx = C7[q][0], cy = C7[q][1], cz = C7[q][2];
          REAL cu = 0.0;
          if (cx) cu = cu + (REAL)cx * ux;
          if (cy) cu = cu + (REAL)cy * uy;
          if (cz) cu = cu + (REAL)cz * uz;
          const REAL w = LIT(W7[q]);
          const REAL geq = w * th * ((REAL)1.0 + (REAL)4.0 * cu);
          gB[q * N + nu] = (STORE)((g[q] + omg * (geq - g[q]) + w * S) * ((REAL)1.0 - sg));
        }
        if (out4) {   /* lbm_step do_out / stepOut (:427) */
          out4[0 * N + nu] = (STORE)th;
          out4[1 * N + nu] = (STORE)ux;
          out4[2 * N + nu] = (STORE)uy;
          out4[3 * N + nu] = (STORE)uz;
        }
      }
    }
  }
}

/* genInitWGSL (:524-553): equilibrium with the inflow profile at the local height above terrain, all cells */
void FN(oracle_init)(const oracle_dom *d, STORE *f) {
  const int NX = d->nx, NY = d->ny, NZ = d->nz;
  const int64_t N = (int64_t)NX * NY * NZ;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < NZ; k++)
    for (int j = 0; j < NY; j++)
      for (int i = 0; i < NX; i++) {
        const int ktc = (int)d->kt[j * NX + i];
        int zagl = k - ktc;
        if (zagl < 0) zagl = 0;
        if (zagl > NZ - 1) zagl = NZ - 1;
        const REAL U = (REAL)d->prof[4 * zagl], V = (REAL)d->prof[4 * zagl + 1];
        const REAL usq = U * U + V * V;
        const int64_t nu = ((int64_t)k * NY + j) * NX + i;
        for (int q = 0; q < 19; q++) {
          const REAL cu = (REAL)C19[q][0] * U + (REAL)C19[q][1] * V;
          f[q * N + nu] = (STORE)(LIT(W19[q]) * (((REAL)1.0 + (REAL)3.0 * cu + (REAL)4.5 * cu * cu) - (REAL)1.5 * usq));
        }
      }
}

/* genSurfWGSL (:435-471): pre-collision theta and force-corrected u, v (times conv) at k = ktop, north-up */
void FN(oracle_surface)(const oracle_dom *d, const STORE *fA, const STORE *gA, float conv, float *th, float *u,
                        float *v) {
  const int NX = d->nx, NY = d->ny, NZ = d->nz;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < NX * NY; c++) {
    const int i = c % NX, j = c / NX;
    const int ktc = (int)d->kt[c], kg = (int)d->ktop[c], k = kg;
    const int o = (NY - 1 - j) * NX + i;
    if (k >= NZ) { th[o] = 0.0f; u[o] = 0.0f; v[o] = 0.0f; continue; }
    const int zagl = FN(zagl_)(k, ktc);
    const REAL uI = (REAL)d->prof[4 * zagl + 0], vI = (REAL)d->prof[4 * zagl + 1];
    REAL f[19], g[7];
    FN(stream_cell_)(d, fA, gA, i, j, k, uI, vI, f, g);
    REAL rho, jx, jy, jz, t;
    FN(moments1_)(f, g, &rho, &jx, &jy, &jz, &t);
    FN(forces_t) F;
    FN(forces_)(d, k, kg, ktc, d->cd[c], uI, vI, rho, jx, jy, jz, t, &F);
    th[o] = (float)t;
    u[o] = (float)(F.ux * (REAL)conv);
    v[o] = (float)(F.uy * (REAL)conv);
  }
}

/* genStatsWGSL (:580-633) + the host reduction of _dbg.stats() (js/live/solver.js:559-576), sums in f64 */
void FN(oracle_stats)(const oracle_dom *d, const STORE *fA, const STORE *gA, oracle_stats *st) {
  const int NX = d->nx, NY = d->ny, NZ = d->nz;
  const int64_t N = (int64_t)NX * NY * NZ;
  double sth = 0, ske = 0, sdr = 0, snf = 0, tmax = 0, umax = 0;
  int64_t bad = 0;
  for (int64_t n = 0; n < N; n++) {
    const int64_t c = n % ((int64_t)NX * NY);
    const int k = (int)(n / ((int64_t)NX * NY));
    if (k < (int)d->ktop[c]) continue;
    REAL rho = (REAL)fA[n];
    for (int q = 1; q < 19; q++) rho = rho + (REAL)fA[q * N + n];
    REAL j3[3];
    for (int cc = 0; cc < 3; cc++) {
      REAL s = 0.0;
      for (int q = 0; q < 19; q++) {
        if (C19[q][cc] == 1) s = s + (REAL)fA[q * N + n];
        else if (C19[q][cc] == -1) s = s - (REAL)fA[q * N + n];
      }
      j3[cc] = s;
    }
    REAL th = (REAL)gA[n];
    for (int q = 1; q < 7; q++) th = th + (REAL)gA[q * N + n];
    if (!isfinite((double)rho) || !isfinite((double)th)) { bad++; continue; }
    const REAL jj = j3[0] * j3[0] + j3[1] * j3[1] + j3[2] * j3[2];
    sth += (double)th;
    ske += (double)((REAL)0.5 * jj / rho);
    sdr += (double)(rho - (REAL)1.0);
    snf += 1.0;
    if ((double)th > tmax) tmax = (double)th;
    const double um = (double)(RSQRT(jj) / rho);
    if (um > umax) umax = um;
  }
  st->sum_theta = sth; st->sum_ke = ske; st->sum_drho = sdr; st->n_fluid = snf;
  st->theta_max = tmax; st->umax = umax; st->non_finite = bad;
}

/* genReinitWGSL (:557-575): cells that turned from solid to fluid restart at rest (f = w_q, g = 0) */
void FN(oracle_reinit)(const oracle_dom *d, const uint32_t *old_ktop, STORE *f, STORE *g) {
  const int NX = d->nx, NY = d->ny, NZ = d->nz;
  const int64_t N = (int64_t)NX * NY * NZ, NXY = (int64_t)NX * NY;
  for (int64_t c = 0; c < NXY; c++) {
    const int k1 = (int)old_ktop[c] < NZ ? (int)old_ktop[c] : NZ;
    for (int k = (int)d->ktop[c]; k < k1; k++) {
      const int64_t nu = k * NXY + c;
      for (int q = 0; q < 19; q++) f[q * N + nu] = (STORE)LIT(W19[q]);
      for (int q = 0; q < 7; q++) g[q * N + nu] = (STORE)0.0;
    }
  }
}

#undef CAT2
#undef CAT
#undef FN
#undef LIT
#undef RSQRT
#undef CST
