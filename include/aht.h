/*
 * aht.h — C-ABI of the MI355X-native coupled lattice-Boltzmann step (D3Q19 regularized-BGK flow + D3Q7
 * potential-temperature transport) of the Ashburn heat twin. Implemented in HIP for gfx950 (libaht.so).
 *
 * The library owns every device buffer. Callers pass plain host pointers and sizes; there is no torch, no CPU path
 * and no fallback: every entry point fails with a negative AHT_E* code when the GPU or the kernels are unavailable.
 *
 * Host arrays (as the reference, js/live/kernel.js:17): populations [q][cell], cell = (k*ny + j)*nx + i, j = 0 at the
 * south edge; per-column arrays [j*nx + i] in that solver orientation; surface outputs north-up [row*nx + col], row 0 =
 * north. (On the device, lattice pairs below 4 GiB are blocked, f32 by 2048 cells, f16/bf16 by 256; aht_read_state /
 * aht_write_state convert.)
 *
 * Threading: one host thread per solver at a time. Work is enqueued on the solver's own HIP stream; aht_step returns
 * once enqueued, every read returns after the data is on the host.
 */
#ifndef AHT_H
#define AHT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AHT_VERSION 1

/* return codes */
#define AHT_OK 0
#define AHT_EINVAL (-1)   /* bad argument (null pointer, size, geometry rule)          */
#define AHT_ENOMEM (-2)   /* device allocation failed                                  */
#define AHT_EDEVICE (-3)  /* HIP runtime error or no GPU                               */
#define AHT_ESTATE (-4)   /* call out of order (no columns / params yet, destroyed)    */

/* population storage formats */
#define AHT_STORE_F32 0   /* f32 populations (208 B per fluid-cell update)                            */
#define AHT_STORE_F16 1   /* f16 deviations (f_q - w_q)*2^10; theta in f32 + f16 (g_q - w_q theta)*2^10;
                             112 B per update, f32 math                                                */
#define AHT_STORE_BF16 2  /* bf16 deviations, same layout as F16, stochastic rounding in the step      */
#define AHT_STORE_F64 3   /* f64 populations and f64 arithmetic: the FP64 twin (416 B per update)     */
#define AHT_STORE_F16S 4  /* f16 (f_q - w_q)*2^10 and g_q*2^10 (no theta plane), stochastic rounding in the
                             step; 104 B per update, f32 math                                          */

/* aht_step flags */
#define AHT_STEP_ACC 1u   /* accumulate the FP64 time means (surface theta/u/v, face fluxes, heat content)  */
#define AHT_STEP_OUT 2u   /* the last step of the call writes (theta, ux, uy, uz) of every fluid cell        */

typedef struct aht_solver aht_solver;

typedef struct {
  int32_t nx, ny, nz;
  int32_t storage;       /* AHT_STORE_*                                          */
  int32_t device;        /* HIP device ordinal                                   */
  int32_t block;         /* threads per block of the step kernel (0 = default)   */
  int32_t layout;        /* AHT_LAYOUT_*: population layout on the device        */
  int32_t reserved;
} aht_dims;

/* population layout on the device (the state exchange is [q][cell] either way): AUTO blocks f32 / f16 / bf16 lattice
 * pairs below 4 GiB (one array [cell >> B][19 D3Q19 + 7 D3Q7][cell % 2^B] per pair, B = 11 for f32, 8 for the 16-bit
 * formats with their f32 theta plane after it; faster), QCELL keeps the reference's [q][cell] for every format */
#define AHT_LAYOUT_AUTO 0
#define AHT_LAYOUT_QCELL 1

/* the scalar block of the reference's uniform Params (js/live/kernel.js:44-50; values js/live/solver.js:225-238) */
typedef struct {
  float tau0, cs2smag, gbeta, prt, kfloor, momfac, srcscale, sigmax;
  int32_t inflow_w, inflow_e, inflow_s, inflow_n, nsp, nsw, oclip, pad;
} aht_params;

/* diagnostics of the stored state over fluid cells (js/live/kernel.js:580-633 + js/live/solver.js:559-576) */
typedef struct {
  double sum_theta;   /* sum of theta (K * cells)                      */
  double sum_ke;      /* sum of 0.5 |j|^2 / rho (lattice units)         */
  double sum_drho;    /* sum of rho - 1                                 */
  double n_fluid;     /* finite fluid cells                             */
  double theta_max;   /* max(theta, 0)                                  */
  double umax;        /* max |j| / rho (lattice units)                  */
  int64_t non_finite; /* cells with non-finite rho or theta             */
} aht_diag;

/* time-mean budgets, lattice units (the host scales by rho*cp, dx, dt, conv); semantics of heat_diag,
 * tools/live/validate_live.py:226-240, evaluated from the FP64 time means of the face cells */
typedef struct {
  int64_t n_acc;           /* accumulated steps                                                   */
  double heat_in;          /* sum over columns of q (K per step per cell, i.e. cells*K/step)       */
  double heat_out;         /* sum_E th*u - sum_W th*u + sum_N th*v - sum_S th*v of the mean fields */
  double content;          /* sum over fluid cells of the mean theta (K * cells)                  */
  double q_w, q_e, q_s, q_n; /* mean volume flux through each face (sum of mean u or v)           */
  double mass_imbalance;   /* |q_w - q_e + q_s - q_n| / (|q_w| + |q_e| + |q_s| + |q_n|)            */
} aht_budget;

/* lifetime (replaces createLiveSolver's device and buffer set-up, js/live/solver.js:44-101, :137-158) */
int aht_create(const aht_dims *dims, aht_solver **out);
void aht_destroy(aht_solver *s);                                 /* destroy(), js/live/solver.js:485-493 */
const char *aht_last_error(const aht_solver *s);                 /* message of the last failure          */
const char *aht_error_string(int code);

/* geometry and sources per column: kt >= 1, kt <= ktop <= nz (replaces writeBuffer of ktop and the packed colv,
 * js/live/solver.js:338, :359-376; numba ABI kt/ktop/cd/q, tools/live/validate_live.py:110-112) */
int aht_set_columns(aht_solver *s, const uint32_t *kt, const uint32_t *ktop, const float *cd, const float *q);
/* parameter block and the per-level profile prof[nz][4] = (u_in, v_in, kb, 0), lattice units
 * (replaces _writeParams, js/live/solver.js:225-242) */
int aht_set_params(aht_solver *s, const aht_params *p, const float *prof);

/* state (replaces resetFlow/_reinitFlow's init kernel, js/live/solver.js:288-293, :389-398; genInitWGSL) */
int aht_init_flow(aht_solver *s);                 /* f := equilibrium of the inflow profile, both lattices   */
int aht_reset_heat(aht_solver *s);                /* g := 0, both lattices (resetHeat, solver.js:380-386)     */
int aht_reinit_cells(aht_solver *s, const uint32_t *old_ktop); /* solid->fluid cells at rest (genReinitWGSL) */

/* stepping (replaces step(n) / _encodeSteps, js/live/solver.js:401-427; numba lbm_step loop,
 * tools/live/validate_live.py:105-117) */
int aht_step(aht_solver *s, int32_t n, uint32_t flags);
int aht_sync(aht_solver *s);
int64_t aht_steps_done(const aht_solver *s);
/* n steps between two HIP events recorded on the solver's stream (replaces benchmark(), solver.js:474-483);
 * *ms = event-measured device time of the n steps */
int aht_benchmark(aht_solver *s, int32_t n, double *ms);

/* read-back (replaces read()'s surf kernel, js/live/solver.js:439-471, genSurfWGSL kernel.js:435-471):
 * pre-collision theta and force-corrected u*conv, v*conv at k = ktop, north-up; u or v may be NULL */
int aht_read_surface(aht_solver *s, float conv, float *theta, float *u, float *v);
/* the u8 theta volume of the lowest nzv levels, [k][row][col] north-up, round(theta*255/volmax) clamped,
 * plus the max theta over those fluid cells (genVolWGSL, kernel.js:475-520) */
int aht_read_volume(aht_solver *s, int32_t nzv, float volmax, uint8_t *vol, float *theta_max);
/* diagnostics (replaces _dbg.stats, js/live/solver.js:559-576) */
int aht_stats(aht_solver *s, aht_diag *out);
/* aht_stats without device allocation (no counterpart in the reference): for callers that cannot call hipMalloc
 * (under HIP graph capture, in tight serving loops, harnesses that forbid allocation). Same kernels and the same
 * results bit for bit, but the partials live in a fixed library-owned scratch of 114,752 bytes per GPU (7 x 2,048
 * partials + 8 result slots; the diagnostics kernel strides over any solver size within 2,048 blocks, and the call
 * fails with AHT_EINVAL if it ever needed more). Every call on one GPU shares that GPU's scratch: calls that overlap
 * on one GPU (solvers stepped from different threads) race and must be serialised by the caller. */
int aht_stats_noalloc(aht_solver *s, aht_diag *out);
/* the (theta, ux, uy, uz) [4][N] written by the last AHT_STEP_OUT step (stepOut/readDbg, solver.js:548-556) */
int aht_read_out(aht_solver *s, float *out4);

/* FP64 time means accumulated by AHT_STEP_ACC steps (replaces stepAcc + the offline means of
 * tools/live/validate_live.py:258-290): surf3 = [3][nx*ny] mean (theta, ux, uy), north-up, lattice units
 * (may be NULL); budget may be NULL */
int aht_mean_reset(aht_solver *s);
int aht_read_mean(aht_solver *s, double *surf3, aht_budget *budget);
/* face means [2][nz][ny or nx] (theta, normal velocity) of face 0..3 = W, E, S, N, lattice units */
int aht_read_face_mean(aht_solver *s, int32_t face, double *out);

/* raw state at the current parity, f32 [19][N] and [7][N] (replaces _dbg.readState/writeState,
 * js/live/solver.js:540-547); the other storages convert on the device; f or g may be NULL */
int aht_read_state(aht_solver *s, float *f, float *g);
int aht_write_state(aht_solver *s, const float *f, const float *g);
/* the same in f64 host format (exact for AHT_STORE_F64; no counterpart in the reference, whose storage is f32 only) */
int aht_read_state_f64(aht_solver *s, double *f, double *g);
int aht_write_state_f64(aht_solver *s, const double *f, const double *g);

/* ---- batches: many uneven domains advanced by one persistent launch per step (no counterpart in the reference, whose
 * host runs one domain per device). Members keep their own state, accumulators and read-back entry points. */
typedef struct aht_batch aht_batch;
int aht_batch_create(int32_t device, int32_t storage, aht_batch **out);
void aht_batch_destroy(aht_batch *b);
const char *aht_batch_last_error(const aht_batch *b);
/* members: same device and storage and one population layout (lattice pairs below 4 GiB are blocked on the device
 * in every format but f64 — f32 by 2048 cells, f16 / bf16 by 256 with their theta plane — larger pairs and f64 keep
 * [q][cell]; a batch cannot mix the two). Batched launches step every member one level per thread, with the batch's
 * block size; a member's own launch shape (aht_set_launch / aht_autotune) applies to its steps outside the batch. */
int aht_batch_add(aht_batch *b, aht_solver *s);
int aht_batch_remove(aht_batch *b, aht_solver *s);
int aht_batch_size(const aht_batch *b);
/* one chunk of n launches: member d (in add order) steps nsteps[d] <= n times (NULL: n) and accumulates the FP64
 * means from its launch acc_from[d] on (NULL or >= nsteps[d]: none). Launch `it` covers only the members with
 * nsteps > it: a member that sits out a launch costs no workgroups. The n launches are captured once as a hipGraph
 * and replayed while the block size, the grid and every launch's member and tile counts stay the same (member tables
 * are uploaded before each chunk, so a change of members with the same counts replays too). *ms (may be NULL) = event
 * time of the chunk. */
int aht_batch_step(aht_batch *b, int32_t n, const int32_t *nsteps, const int32_t *acc_from, double *ms);
int aht_batch_sync(aht_batch *b);
/* grid size of the batched step: blocks > 0 fixes it (clamped to the tile count: one block per tile), 0 = online
 * tuning between one block per tile and 32 blocks per CU (persistent, looping over tiles); use_graph 0 launches
 * without graph capture */
int aht_batch_set_grid(aht_batch *b, int32_t blocks, int32_t use_graph);
/* threads per block of the batched step: 64, 128 or 256 (tiles of that many columns), 0 = online tuning over the
 * three. The tuner times one (block, grid, cap) candidate per chunk as device time per cell-step, where a grid of one
 * block per tile is also tried with its wavefronts capped at 5 and 4 per SIMD (aht_launch::waves), and tunes again
 * whenever a chunk's cell-steps move by more than 2x from where the last tuning started; every candidate computes the
 * same bits (no counterpart in the reference: its kernel has one workgroup size, js/live/kernel.js:395) */
int aht_batch_set_block(aht_batch *b, int32_t block);
int aht_batch_grid(const aht_batch *b);              /* blocks launched for all members by the last chunk (0: none yet) */
int aht_batch_block(const aht_batch *b);             /* block size of the last chunk (0: none yet)                      */

/* launch shape of one solver's step kernel (no counterpart in the reference: its WGSL kernel has one fixed shape) */
typedef struct {
  int32_t block;         /* threads per block: 64, 128, 192 or 256                                        */
  int32_t levels;        /* cells per thread along z: 1, 2 or 4                                           */
  int32_t window;        /* tile order: 1 linear, W = 2, 4, ..., 4096 = runs of W consecutive tiles per XCD,
                            0 = one contiguous slab of tiles per XCD, -1 = plane slabs (XCD x steps the
                            x-th eighth of the columns of every level, all XCDs level by level together)  */
  int32_t sweep;         /* 0: one step per launch. Otherwise two steps per launch (f32 storage in the blocked
                            pair layout; other formats ignore it): a z sweep over level chunks of |sweep|
                            levels, regions of 32 x 8 columns (sweep > 0) or 16 x 16 (sweep < 0), with the
                            first step's state kept on chip (LDS-staged halo planes); steps with FP64 means or
                            debug outputs, and an odd last step, take the one-step kernel (block, levels).
                            Bit-identical results. Opt-in, the autotuner does not try it: at 600 x 600 x 64 it
                            matches the one-step kernel at best (sweep 12, window 4), on smaller grids it is
                            slower                                                                          */
  int32_t waves;         /* 0: as many wavefronts per SIMD as the kernel's registers allow. 1-8: at most that many
                            of the one-step kernel's wavefronts share a SIMD (the launch reserves LDS the kernel
                            does not use, so that fewer blocks fit on a CU). Bit-identical results. At
                            600 x 600 x 64 the f32 step runs 3 % faster at 4 (0.799 against 0.823 ms at its
                            register-bound 6); the 16-bit steps run slower with a cap                       */
} aht_launch;

/* launch-shape autotuner of one solver: times every block x levels x window candidate over `steps` steps each,
 * re-times the six fastest in three rounds over 10 x `steps`, keeps the fastest (written to *best, may be NULL), and
 * restores the state and counters */
int aht_autotune(aht_solver *s, int32_t steps, aht_launch *best);
int aht_set_launch(aht_solver *s, const aht_launch *l);
int aht_get_launch(const aht_solver *s, aht_launch *l);

/* device bytes held by the solver (admission by footprint in the serving loop) */
int64_t aht_footprint_bytes(const aht_dims *dims);

#ifdef __cplusplus
}
#endif
#endif
