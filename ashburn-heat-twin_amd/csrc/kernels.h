// Internal interface between the C-ABI (api.hip) and the kernels (step.hip). Not installed.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lattice.h"

namespace aht {

// one solver's device view, passed by value as the kernel argument
struct Grid {
  int nx, ny, nz;
  int kmin;            // lowest fluid level over all columns (first plane the step launches)
  uint32_t NXY;
  uint32_t N;          // nx * ny * nz (< 2^32; per-q offsets use 64-bit arithmetic)
  float inv_nx;        // 1 / nx, for the column -> (i, j) split
  int pad;
  const uint4 *colv;   // per column (kt, ktop, cd bits, q bits), solver orientation
  const uint32_t *ktop;
  const float4 *prof;  // [nz] (u_in, v_in, kb, 0)
  Params P;
};

// FP64 time-mean accumulators (AHT_STEP_ACC)
struct Acc {
  double *surf;     // [3][NXY]  theta, ux, uy at k = ktop
  double *face;     // W [2][nz*ny], E [2][nz*ny], S [2][nz*nx], N [2][nz*nx] (theta, normal velocity)
  double *content;  // one slot per wavefront of the step grid: sum of theta over its fluid cells
};

constexpr int STEP_BLOCK = 256;

size_t face_offset(int face, int nx, int ny, int nz);   // offset of face 0..3 in Acc::face
size_t face_len(int face, int nx, int ny, int nz);      // doubles per face (2 * nz * ny or 2 * nz * nx)
uint32_t content_slots(const Grid &g, int block);       // wavefront slots of the step grid

hipError_t launch_step(const Grid &g, const float *fA, float *fB, const float *gA, float *gB, const Acc *acc,
                       float *out4, int block, hipStream_t st);
hipError_t launch_init(const Grid &g, float *f0, float *f1, hipStream_t st);
hipError_t launch_reinit(const Grid &g, const uint32_t *old_ktop, float *f, float *gq, hipStream_t st);
hipError_t launch_surface(const Grid &g, const float *fA, const float *gA, float conv, float *out3, hipStream_t st);
hipError_t launch_volume(const Grid &g, const float *gA, int nzv, float volscale, uint32_t *vout, uint32_t *tmax_bits,
                         hipStream_t st);
// two-pass deterministic reduction: per-block partials into part (>= stats_partials(g) entries), then one block
hipError_t launch_stats(const Grid &g, const float *fA, const float *gA, double *part, double *out8, hipStream_t st);
uint32_t stats_partials(const Grid &g);
// sums content slots, surface and face accumulators into out (deterministic order)
hipError_t launch_sum(const double *x, uint32_t n, double *out, hipStream_t st);

}  // namespace aht
