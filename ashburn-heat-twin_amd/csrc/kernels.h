// Internal interface between the C-ABI (api.hip) and the kernels (step.hip). Not installed.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lattice.h"

namespace aht {

// one solver's device view: the step kernel reads it from a device copy through a pointer (as an argument it would
// be pinned in scalar registers and spilled); the other kernels take it by value, batch members hold a copy each
struct Grid {
  int nx, ny, nz;
  int kmin;            // lowest fluid level over all columns (first plane the step launches)
  uint32_t NXY;
  uint32_t N;          // nx * ny * nz (< 2^32; per-q offsets use 64-bit arithmetic)
  float inv_nx;        // 1 / nx, for the column -> (i, j) split
  uint32_t win;        // tile order: 1 = linear, W > 1 = windows of W consecutive tiles per XCD, 0 = one slab per XCD
  uint32_t th_off;     // theta-referenced D3Q7 lattices: byte offset of the f32 theta plane (0 otherwise)
  int32_t lb;          // population layout: 0 = [q][cell]; LB > 0 = blocked [cell >> LB][q][cell & (2^LB - 1)]
  const uint32_t *colk; // per column: kt | ktop << 10 | kfast << 20 | near << 30 (read by every cell), solver orientation
  const uint32_t *colnb; // per column: ktop of the 8 lateral neighbours, 4 bits each (see pack_neighbours in api.hip),
                         // read by near-solid cells only
  const float2 *colcq;  // per column: (cd, q), read only at the first fluid cell
  const uint32_t *ktop;
  const float4 *prof;  // [nz] (u_in, v_in, kb, 0)
  Params P;
  ParamRecip rp;       // 1/prt, 1/nsp, 1/nsw
};

// FP64 time-mean accumulators (AHT_STEP_ACC)
struct Acc {
  double *surf;     // [3][NXY]  theta, ux, uy at k = ktop
  double *face;     // W [2][nz*ny], E [2][nz*ny], S [2][nz*nx], N [2][nz*nx] (theta, normal velocity)
  double *content;  // [nz][ceil(NXY / 64)]: sum of theta over the fluid cells of one level of 64 columns
};

// Stream bases of the single-domain fast path, host-built once per lattice parity (the step that reads lattice pair A
// and writes pair B): rd = the base of every read stream with its pull offset folded in (population plane q shifted
// back by c_q . (1, nx, nx ny) cells; in a blocked layout by the uniform part of the pull, pull_shift), wr = every
// written plane, lo = the blocked pull's remainder within a block (pull_rem; the lane's compare operand). A wavefront
// loads these through the scalar cache instead of forming 60 64-bit addresses with scalar multiplies and adds.
struct Streams {
  uint64_t rd[33];   // f_q (0..18), g_q (19..25), the theta plane pulled along D3Q7 velocity q (26..32, theta_ref)
  uint64_t wr[27];   // f_q planes (0..18), g_q planes (19..25), the theta plane (26, theta_ref)
  uint32_t lo[19];   // blocked layouts: pull remainder of velocity q (D3Q7 q < 7 shares it)
};

constexpr int STEP_BLOCK = 256;
// Grid::win value of the plane-slab tile order (aht_launch window -1): XCD x steps the x-th eighth of the column
// blocks of every level chunk, level chunk after level chunk
constexpr uint32_t WIN_PLANE_SLABS = 0xffffffffu;
constexpr uint32_t NUM_XCD = 8;   // MI355X: workgroups are dispatched round-robin over 8 XCDs, each with its own L2

// one member of a batched launch (device-resident table, uploaded by the host before every chunk)
struct Member {
  Grid g;
  void *f[2];
  void *gq[2];
  Acc acc;
  uint32_t tiles;      // (block of BS columns, one level) tiles of one step
  uint32_t tile_base;  // first tile in the concatenated list
  uint32_t bx;         // column blocks per level chunk: ceil(NXY / BS)
  float inv_bx;        // 1 / bx
  int parity0;         // lattice holding the state at launch index 0 of the chunk
  int nsteps;          // launches of the chunk in which this member steps (the table is ordered by it, descending)
  int acc_from;        // first launch index that accumulates the FP64 means (>= nsteps: none)
  uint32_t step0;      // the member's step index at launch index 0 of the chunk
  Streams s[2];        // stream bases of the member's lattice parities (make_streams), read as the single step does
};

size_t face_offset(int face, int nx, int ny, int nz);   // offset of face 0..3 in Acc::face
size_t face_len(int face, int nx, int ny, int nz);      // doubles per face (2 * nz * ny or 2 * nz * nx)
uint32_t content_slots(uint32_t nxy, int nz);           // heat-content slots: one per level and 64-column group

int storage_bytes(int storage);                         // bytes per population element of an AHT_STORE_* format
bool storage_theta_ref(int storage);                    // D3Q7 lattice = f32 theta + deviations (f16, bf16)
size_t theta_offset(int storage, int lb, size_t N);     // byte offset of the theta plane from the D3Q7 base (or 0)
int population_layout(int storage, size_t N);           // Grid::lb for a lattice of N cells in this format
size_t lattice_cells(int lb, size_t N);                 // cells per population, rounded up to whole blocks
// bytes kept free before the first and after the last blocked lattice pair of an allocation: the fast path's pulls
// of face cells at levels 1 and nz - 2 reach up to nx + 1 cells outside the lattice (their values are replaced)
size_t lattice_guard_bytes(int storage, int lb, int nx);
size_t f_lattice_bytes(int storage, int lb, size_t N);  // bytes of one D3Q19 lattice (a blocked pair: the whole pair)
size_t g_lattice_bytes(int storage, int lb, size_t N);  // bytes of one D3Q7 lattice (a blocked pair: 0, inside the D3Q19)
size_t g_lattice_offset(int storage, int lb);           // a blocked pair's D3Q7 base from its D3Q19 base (else 0)
hipError_t launch_zero_g_blocked(int storage, const Grid &g, void *gq, hipStream_t st);   // blocked pair: D3Q7 = 0

// kernels are templated on the storage format; the launchers dispatch on the AHT_STORE_* id
// levels = cells per thread along z (1, 2 or 4): the tile is (block of columns, chunk of `levels` levels);
// g.win sets the order in which blocks take tiles
// gd = device copy of g: the kernel reads the grid through a pointer, so the compiler loads fields where they are used
// (as a by-value argument the grid was pinned in SGPRs, which spilled into VGPR lanes: ~120 v_readlane/v_writelane
// per cell of the f32 fast path); sd = device copy of make_streams(storage, g, fA, fB, gA, gB)
// step = the solver's step index of this launch (random bits of the stochastic-rounding formats)
// dynamic LDS (bytes) that caps a kernel of `block` threads at `waves` wavefronts per SIMD on a CU with `lds_cu`
// bytes of LDS (0: no cap, or the cap cannot be expressed within `lds_block` bytes per block)
uint32_t occupancy_lds(int block, int waves, uint32_t lds_cu, uint32_t lds_block);
hipError_t launch_step(int storage, const Grid &g, const Grid *gd, const Streams *sd, const void *fA, void *fB,
                       const void *gA, void *gB, const Acc *acc, float *out4, int block, int levels, uint32_t step,
                       hipStream_t st, uint32_t lds = 0);
Streams make_streams(int storage, const Grid &g, const void *fA, void *fB, const void *gA, void *gB);
// two steps per launch (the z sweep with LDS-staged halo planes; f32 pair layout only, no means or debug outputs):
// regions of tx x (256 / tx) columns (tx = 16 or 32), chunks of kc levels, tiles in the XCD-aware order win (plane
// slabs fall back to linear); step = the step index of the first of the two steps; reads pair A, writes pair B
bool sweep2_supported(int storage, int lb);
hipError_t launch_sweep2(int storage, const Grid &g, const Grid *gd, const void *fA, void *fB, const void *gA,
                         void *gB, int tx, int kc, uint32_t win, uint32_t step, hipStream_t st);

// one step of the first nmem batch members (launch index `it` of a chunk; their tiles are the first total_tiles) on
// a grid of min(grid, total_tiles) blocks of `block` (64, 128 or 256) threads, one level per thread
// lds: occupancy_lds's reservation for a grid of one block per tile (a persistent grid launches without one)
hipError_t launch_batch_step(int storage, int lb, int block, const Member *mem, int nmem, uint32_t total_tiles, int it,
                             int grid, hipStream_t st, uint32_t lds = 0);
hipError_t launch_init(int storage, const Grid &g, void *f0, void *f1, hipStream_t st);
hipError_t launch_reinit(int storage, const Grid &g, const uint32_t *old_ktop, void *f, void *gq, hipStream_t st);
hipError_t launch_surface(int storage, const Grid &g, const void *fA, const void *gA, float conv, float *out3,
                          hipStream_t st);
hipError_t launch_volume(int storage, const Grid &g, const void *gA, int nzv, float volscale, uint32_t *vout,
                         uint32_t *tmax_bits, hipStream_t st);
// diagnostics, two-level deterministic reduction: a grid of stats_blocks(g) <= STATS_MAX_BLOCKS blocks writes
// part[7 * stats_blocks(g)], then one block per statistic writes out8[0..6]
constexpr uint32_t STATS_MAX_BLOCKS = 2048;
hipError_t launch_stats(int storage, const Grid &g, const void *fA, const void *gA, double *part, double *out8,
                        hipStream_t st);
uint32_t stats_blocks(const Grid &g);
// fixed-order sum of n doubles into *out (one block)
hipError_t launch_sum(const double *x, uint32_t n, double *out, hipStream_t st);
// population q of a lattice of n cells (layout lb) between storage and an f32 (or f64) host-format plane; w = rest
// weight for f populations; g populations of theta-referenced formats decode / encode against the lattice's theta
// plane `th` (w = the population's D3Q7 weight); otherwise th is unused
hipError_t launch_decode(int storage, int lb, const void *lat, int q, void *dst, int dst_f64, uint32_t n, float w,
                         int is_f, const float *th, hipStream_t st);
hipError_t launch_encode(int storage, int lb, const void *src, int src_f64, void *lat, int q, uint32_t n, float w,
                         int is_f, const float *th, hipStream_t st);

}  // namespace aht
