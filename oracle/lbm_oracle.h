/*
 * CPU oracle of the coupled D3Q19 + D3Q7 lattice-Boltzmann step — TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library, and only as the
 * reference the GPU results are compared with (or as the timed CPU baseline). The product library (libaht.so) never links it and has no CPU path.
 *
 * It restates the reference kernel generator js/live/kernel.js (step :395-431 with its sections :102-386,
 * near-surface read :435-471, init :524-553, geometry reinit :557-575, diagnostics :580-633) as plain C, one
 * function per generated kernel, evaluating every expression in the generated WGSL's operation order.
 * Three builds of the same source differ only in types:
 *   _f32      : f32 arithmetic, f32 storage   (the WGSL kernel)
 *   _f64s32   : f64 arithmetic, f32 storage   (the numba kernel: f64-promoted arithmetic, f32 arrays;
 *                                              js/live/kernel.js:12-13, tools/live/validate_live.py:97-101)
 *   _f64      : f64 arithmetic, f64 storage   (the FP64 reference for reduced-precision error statements)
 *
 * Layout (as the reference, js/live/kernel.js:17): populations [q][cell], cell = (k*ny + j)*nx + i, j = 0 at
 * the south edge; per-column arrays [j*nx + i].
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H
#include <stdint.h>

typedef struct {
  int32_t nx, ny, nz, pad0;
  const uint32_t *kt;    /* terrain-solid cells per column (kt >= 1)            */
  const uint32_t *ktop;  /* first fluid level per column (kt <= ktop <= nz)     */
  const float *cd;       /* log-law drag coefficient per column                 */
  const float *q;        /* source in K per step per column (first fluid cell)  */
  const float *prof;     /* [nz][4]: inflow u, v (lattice units), K-profile kb, 0 */
  /* the reference's uniform Params block (js/live/kernel.js:44-50), same order */
  float tau0, cs2smag, gbeta, prt, kfloor, momfac, srcscale, sigmax;
  int32_t inflow_w, inflow_e, inflow_s, inflow_n, nsp, nsw, oclip, pad1;
} oracle_dom;

/* per-fluid-cell sums over the stored state (js/live/kernel.js:580-633), accumulated in f64 here */
typedef struct {
  double sum_theta, sum_ke, sum_drho, n_fluid;
  double theta_max, umax;
  int64_t non_finite;
} oracle_stats;

#define ORACLE_DECL(SFX, STORE)                                                                              \
  void oracle_init_##SFX(const oracle_dom *d, STORE *f);                                                      \
  void oracle_step_##SFX(const oracle_dom *d, const STORE *fA, STORE *fB, const STORE *gA, STORE *gB,         \
                         STORE *out4);                                                                        \
  void oracle_surface_##SFX(const oracle_dom *d, const STORE *fA, const STORE *gA, float conv, float *th,     \
                            float *u, float *v);                                                              \
  void oracle_stats_##SFX(const oracle_dom *d, const STORE *fA, const STORE *gA, oracle_stats *st);           \
  void oracle_reinit_##SFX(const oracle_dom *d, const uint32_t *old_ktop, STORE *f, STORE *g);

ORACLE_DECL(f32, float)
ORACLE_DECL(f64s32, float)
ORACLE_DECL(f64, double)

int oracle_num_threads(void);

#endif
