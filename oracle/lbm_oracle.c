/*
 * CPU oracle of the coupled D3Q19 + D3Q7 lattice-Boltzmann step — TEST INFRASTRUCTURE ONLY (see lbm_oracle.h).
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp: no FMA contraction, so f32 results follow the
 * WGSL expression order exactly).
 */
#include "lbm_oracle.h"
#include <math.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* lattice tables, js/live/kernel.js:19-28 (C19/W19/OPP19/MIRZ19, C7/W7/OPP7) */
static const int C19[19][3] = {{0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
                               {1, 1, 0},  {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 0, 1},  {-1, 0, -1}, {1, 0, -1},
                               {-1, 0, 1}, {0, 1, 1},   {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};
static const double W19[19] = {1.0 / 3,  1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18,
                               1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36,
                               1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
static const int OPP19[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
static const int MIRZ19[19] = {0, 1, 2, 3, 4, 6, 5, 7, 8, 9, 10, 13, 14, 11, 12, 17, 18, 15, 16};
static const int C7[7][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
static const double W7[7] = {0.25, 0.125, 0.125, 0.125, 0.125, 0.125, 0.125};
static const int OPP7[7] = {0, 2, 1, 4, 3, 6, 5};

#define REAL float
#define STORE float
#define SFX f32
#include "lbm_oracle_body.h"
#undef REAL
#undef STORE
#undef SFX

#define REAL double
#define STORE float
#define SFX f64s32
#include "lbm_oracle_body.h"
#undef REAL
#undef STORE
#undef SFX

#define REAL double
#define STORE double
#define SFX f64
#include "lbm_oracle_body.h"
#undef REAL
#undef STORE
#undef SFX

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* table export for the tests that re-derive OPP/MIRZ from C19 */
const int *oracle_c19(void) { return &C19[0][0]; }
const int *oracle_opp19(void) { return OPP19; }
const int *oracle_mirz19(void) { return MIRZ19; }
