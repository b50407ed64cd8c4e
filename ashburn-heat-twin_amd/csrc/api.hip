// C-ABI of the coupled lattice-Boltzmann step (include/aht.h). Owns the device buffers of one solver: two f32
// population lattices per distribution (A/B ping-pong, js/live/solver.js:149-150, :184), the packed columns, the
// profile block and the FP64 accumulators. Every failure returns a negative code and leaves a message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "aht.h"
#include "kernels.h"

struct aht_batch;

namespace {
// Neighbour word of a column for the near-solid cells of the fast path: for each of the 8 lateral neighbours
// (dx, dy) in {-1, 0, 1}^2 minus (0, 0), nibble n = (dy + 1) * 3 + (dx + 1) (minus one past the centre) holds
// e = ktop(neighbour) - ktop(column) + 1, so that the source (i + dx, j + dy, k - c_z) of a cell at height
// r = k - ktop above its column top is solid iff r - c_z + 1 < e (e <= 0 never is: stored as 0). Returns false when
// some e exceeds 15 (a neighbour more than 14 levels taller): such a column's near-solid cells take the boundary
// path. Neighbours are clamped at the faces as the boundary path clamps them.
bool pack_neighbours(const uint32_t *ktop, int nx, int ny, int i, int j, uint32_t &word) {
  const int kc = (int)ktop[(size_t)j * nx + i];
  word = 0;
  bool ok = true;
  int slot = 0;
  for (int dy = -1; dy <= 1; dy++)
    for (int dx = -1; dx <= 1; dx++) {
      if (dx == 0 && dy == 0) continue;
      const int a = std::min(std::max(i + dx, 0), nx - 1), b = std::min(std::max(j + dy, 0), ny - 1);
      const int e = std::max((int)ktop[(size_t)b * nx + a] - kc + 1, 0);
      if (e > 15) ok = false;
      word |= (uint32_t)std::min(e, 15) << (4 * slot);
      slot++;
    }
  return ok;
}
}  // namespace

// what the step kernel reads through pointers: the grid and the stream tables of both lattice parities
struct DevGrid {
  aht::Grid g;
  aht::Streams s[2];   // s[p]: the step that reads lattice pair p and writes pair 1 - p
};

struct aht_solver {
  aht_dims dims{};
  int device = 0;
  hipStream_t st = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  aht::Grid g{};
  int block = aht::STEP_BLOCK;
  int levels = 1;             // cells per thread along z
  int sweep = 0;              // aht_launch::sweep: 0 = one step per launch, else the two-step sweep's chunk / region
  int waves = 0;              // aht_launch::waves: wavefronts per SIMD cap of the one-step kernel (0 = none)
  uint32_t lds_cu = 0, lds_block = 0;   // the device's LDS per CU and per block (bytes)
  int storage = AHT_STORE_F32;
  int eb = 4;                 // bytes per stored population element
  size_t fbytes = 0;          // bytes of one D3Q19 lattice (a blocked f32 pair: the whole array, D3Q7 included)
  size_t gbytes = 0;          // bytes of one D3Q7 lattice (theta-referenced: + the f32 theta plane; blocked pair: 0)
  void *f[2] = {nullptr, nullptr}, *gq[2] = {nullptr, nullptr};   // parts of `lattices`
  void *lattices = nullptr;
  void *stage = nullptr;      // one population plane in f64, for state conversion
  DevGrid *gdev = nullptr;    // device copy of g and of its stream tables, read by the step kernel
  DevGrid *gpin = nullptr;    // its pinned staging copy
  bool gdirty = true;         // g changed since the last upload
  uint32_t *colk = nullptr, *colnb = nullptr;
  float2 *colcq = nullptr;
  uint32_t *ktop = nullptr, *oldk = nullptr;
  float4 *prof = nullptr;
  double *acc_surf = nullptr, *acc_face = nullptr, *acc_content = nullptr, *red_out = nullptr;
  uint32_t content_cap = 0;
  double *stat_part = nullptr;
  float *out4 = nullptr, *surf3 = nullptr;
  uint32_t *vol = nullptr, *tmax = nullptr;
  int parity = 0;
  bool has_cols = false, has_params = false, has_out = false;
  // ordering against a batch's stream: the batch records `after_batch` after each chunk, the solver records `own`
  // before the batch steps it
  hipEvent_t after_batch = nullptr, own = nullptr;
  bool batch_pending = false;
  aht_batch *batch = nullptr;
  int64_t steps = 0, n_acc = 0;
  std::vector<float> h_q;
  std::string err;
};

namespace {

int fail(aht_solver *s, int code, const char *fmt, ...) {
  if (s) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    s->err = buf;
  }
  return code;
}

#define HIP_TRY(s, call)                                                                                    \
  do {                                                                                                      \
    hipError_t e_ = (call);                                                                                 \
    if (e_ != hipSuccess) return fail((s), e_ == hipErrorOutOfMemory ? AHT_ENOMEM : AHT_EDEVICE, "%s: %s", \
                                      #call, hipGetErrorString(e_));                                        \
  } while (0)

size_t face_total(int nx, int ny, int nz) { return aht::face_offset(4, nx, ny, nz); }

// heat-content slots (one per level and group of 64 columns, whatever the launch shape)
uint32_t max_content_slots(const aht_dims &d) { return aht::content_slots((uint32_t)d.nx * (uint32_t)d.ny, d.nz); }

int check_ready(aht_solver *s) {
  if (!s) return AHT_EINVAL;
  if (!s->has_cols || !s->has_params) return fail(s, AHT_ESTATE, "columns and parameters must be set first");
  return AHT_OK;
}

int set_device(aht_solver *s) {
  HIP_TRY(s, hipSetDevice(s->device));
  if (s->batch_pending) {                 // work the solver enqueues next comes after the batch chunk that stepped it
    HIP_TRY(s, hipStreamWaitEvent(s->st, s->after_batch, 0));
    s->batch_pending = false;
  }
  return AHT_OK;
}

}  // namespace

extern "C" {

const char *aht_error_string(int code) {
  switch (code) {
    case AHT_OK: return "ok";
    case AHT_EINVAL: return "invalid argument";
    case AHT_ENOMEM: return "device out of memory";
    case AHT_EDEVICE: return "HIP runtime error";
    case AHT_ESTATE: return "call out of order";
    default: return "unknown error";
  }
}

const char *aht_last_error(const aht_solver *s) { return s ? s->err.c_str() : "null solver"; }

int64_t aht_footprint_bytes(const aht_dims *d) {
  if (!d || d->nx <= 0 || d->ny <= 0 || d->nz <= 0) return -1;
  const int64_t N = (int64_t)d->nx * d->ny * d->nz, NXY = (int64_t)d->nx * d->ny;
  const int eb = d->storage == AHT_STORE_F32 ? 4 : (d->storage == AHT_STORE_F64 ? 8 : 2);
  const int lb = d->layout == AHT_LAYOUT_QCELL ? 0 : aht::population_layout(d->storage, (size_t)N);
  // both lattice pairs in one allocation of 2 MiB-aligned parts, as aht_create allocates them
  const int64_t A = (int64_t)2 << 20;
  const int64_t fb = ((int64_t)aht::f_lattice_bytes(d->storage, lb, (size_t)N) + A - 1) / A * A;
  const int64_t gb = ((int64_t)aht::g_lattice_bytes(d->storage, lb, (size_t)N) + A - 1) / A * A;
  const int64_t guard = ((int64_t)aht::lattice_guard_bytes(d->storage, lb, d->nx) + A - 1) / A * A;
  const int64_t pop = 2 * (fb + gb) + 2 * guard;
  (void)eb;
  const int64_t cols = NXY * (4 + 8 + 4 + 4) + (int64_t)d->nz * 16;
  const int64_t acc = 3 * NXY * 8 + (int64_t)face_total(d->nx, d->ny, d->nz) * 8 + (int64_t)max_content_slots(*d) * 8;
  const int64_t misc = 7 * (int64_t)aht::STATS_MAX_BLOCKS * 8 + 3 * NXY * 4 + (N + 3) / 4 * 4 + 4096;
  return pop + cols + acc + misc;
}

int aht_create(const aht_dims *dims, aht_solver **out) {
  if (!dims || !out) return AHT_EINVAL;
  *out = nullptr;
  const aht_dims d = *dims;
  if (d.nx < 2 || d.ny < 2 || d.nz < 2 || d.nz > 1022) return AHT_EINVAL;
  const uint64_t N = (uint64_t)d.nx * d.ny * d.nz;
  if (d.storage < AHT_STORE_F32 || d.storage > AHT_STORE_F16S) return AHT_EINVAL;
  if (d.layout != AHT_LAYOUT_AUTO && d.layout != AHT_LAYOUT_QCELL) return AHT_EINVAL;
  // byte offsets within one population plane and within a D3Q7 lattice (theta plane) stay 32-bit
  if (N * (uint64_t)aht::storage_bytes(d.storage) >= (1ull << 32)) return AHT_EINVAL;
  if (aht::storage_theta_ref(d.storage) && (uint64_t)aht::g_lattice_bytes(d.storage, 0, (size_t)N) >= (1ull << 32))
    return AHT_EINVAL;
  const int block = d.block ? d.block : aht::STEP_BLOCK;
  if (block < 64 || block > 256 || block % 64) return AHT_EINVAL;
  aht_solver *s = new aht_solver();
  s->dims = d;
  s->storage = d.storage;
  s->eb = aht::storage_bytes(d.storage);
  s->g.lb = d.layout == AHT_LAYOUT_QCELL ? 0 : aht::population_layout(d.storage, (size_t)N);
  s->fbytes = aht::f_lattice_bytes(d.storage, s->g.lb, (size_t)N);
  s->gbytes = aht::g_lattice_bytes(d.storage, s->g.lb, (size_t)N);
  s->device = d.device;
  s->block = block;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= d.device || d.device < 0) {
    // keep the solver so the caller can read the message, then let it be destroyed
    *out = s;
    return fail(s, AHT_EDEVICE, "no HIP device %d (found %d)", d.device, ndev);
  }
  *out = s;
  int rc = set_device(s);
  if (rc) return rc;
  {
    int cu = 0, blk = 0;
    HIP_TRY(s, hipDeviceGetAttribute(&cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, s->device));
    HIP_TRY(s, hipDeviceGetAttribute(&blk, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device));
    s->lds_cu = (uint32_t)std::max(cu, 0);
    s->lds_block = (uint32_t)std::max(blk, 0);
  }
  HIP_TRY(s, hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
  HIP_TRY(s, hipEventCreate(&s->ev0));
  HIP_TRY(s, hipEventCreate(&s->ev1));
  HIP_TRY(s, hipEventCreateWithFlags(&s->after_batch, hipEventDisableTiming));
  HIP_TRY(s, hipEventCreateWithFlags(&s->own, hipEventDisableTiming));
  const uint32_t NXY = (uint32_t)d.nx * d.ny;
  aht::Grid &g = s->g;
  g.nx = d.nx; g.ny = d.ny; g.nz = d.nz; g.kmin = d.nz; g.NXY = NXY; g.N = (uint32_t)N;
  g.inv_nx = 1.0f / (float)d.nx;
  g.win = 1;
  g.th_off = (uint32_t)aht::theta_offset(d.storage, s->g.lb, (size_t)N);
  {
    // both lattice pairs in one allocation (2 MiB-aligned parts): one large mapping keeps the 52+ streams of a step
    // inside few translation entries (the f16 step's L1 TLB misses went from 4 M per launch to none at 600x600x64)
    const size_t A = (size_t)2 << 20;
    const size_t fb = (s->fbytes + A - 1) / A * A, gb = (s->gbytes + A - 1) / A * A;
    // blocked pairs: guard parts before the first and after the second pair (pull_guard_bytes; read, never used)
    const size_t guard = (aht::lattice_guard_bytes(d.storage, s->g.lb, d.nx) + A - 1) / A * A;
    char *alloc = nullptr;
    HIP_TRY(s, hipMalloc((void **)&alloc, 2 * (fb + gb) + 2 * guard));
    s->lattices = alloc;
    char *base = alloc + guard;
    // a blocked pair is one array (gb = 0) whose D3Q7 populations start 19 populations into the first block
    const size_t go = aht::g_lattice_offset(d.storage, s->g.lb);
    s->f[0] = base; s->f[1] = base + fb + gb;
    s->gq[0] = gb ? base + fb : base + go;
    s->gq[1] = gb ? base + 2 * fb + gb : base + fb + go;
    HIP_TRY(s, hipMemsetAsync(alloc, 0, 2 * (fb + gb) + 2 * guard, s->st));
  }
  HIP_TRY(s, hipMalloc(&s->colk, NXY * sizeof(uint32_t)));
  HIP_TRY(s, hipMalloc(&s->colnb, NXY * sizeof(uint32_t)));
  HIP_TRY(s, hipMalloc(&s->colcq, NXY * sizeof(float2)));
  HIP_TRY(s, hipMalloc(&s->ktop, NXY * sizeof(uint32_t)));
  HIP_TRY(s, hipMalloc(&s->oldk, NXY * sizeof(uint32_t)));
  HIP_TRY(s, hipMalloc(&s->prof, d.nz * sizeof(float4)));
  HIP_TRY(s, hipMalloc(&s->acc_surf, 3 * (size_t)NXY * sizeof(double)));
  HIP_TRY(s, hipMalloc(&s->acc_face, face_total(d.nx, d.ny, d.nz) * sizeof(double)));
  s->content_cap = max_content_slots(d);
  HIP_TRY(s, hipMalloc(&s->acc_content, s->content_cap * sizeof(double)));
  HIP_TRY(s, hipMalloc(&s->red_out, 8 * sizeof(double)));
  HIP_TRY(s, hipMalloc(&s->stat_part, 7 * (size_t)aht::STATS_MAX_BLOCKS * sizeof(double)));
  HIP_TRY(s, hipMalloc(&s->surf3, 3 * (size_t)NXY * sizeof(float)));
  HIP_TRY(s, hipMalloc(&s->vol, ((N + 3) / 4) * sizeof(uint32_t)));
  HIP_TRY(s, hipMalloc(&s->tmax, sizeof(uint32_t)));
  s->g.colk = s->colk; s->g.colnb = s->colnb; s->g.colcq = s->colcq; s->g.ktop = s->ktop; s->g.prof = s->prof;
  HIP_TRY(s, hipMalloc(&s->gdev, sizeof(DevGrid)));
  HIP_TRY(s, hipHostMalloc((void **)&s->gpin, sizeof(DevGrid), hipHostMallocDefault));
  rc = aht_mean_reset(s);
  if (rc) return rc;
  HIP_TRY(s, hipStreamSynchronize(s->st));
  return AHT_OK;
}

void aht_destroy(aht_solver *s) {
  if (!s) return;
  if (s->batch) aht_batch_remove(s->batch, s);     // waits for the batch's last chunk
  if (s->st) {
    (void)hipSetDevice(s->device);
    (void)hipStreamSynchronize(s->st);
  }
  void *bufs[] = {s->lattices, s->stage, s->gdev, s->colk, s->colnb, s->colcq, s->ktop, s->oldk, s->prof, s->acc_surf, s->acc_face,
                  s->acc_content, s->red_out, s->stat_part, s->out4, s->surf3, s->vol, s->tmax};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (s->gpin) (void)hipHostFree(s->gpin);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->after_batch) (void)hipEventDestroy(s->after_batch);
  if (s->own) (void)hipEventDestroy(s->own);
  if (s->st) (void)hipStreamDestroy(s->st);
  delete s;
}

int aht_set_columns(aht_solver *s, const uint32_t *kt, const uint32_t *ktop, const float *cd, const float *q) {
  if (!s || !kt || !ktop || !cd || !q) return s ? fail(s, AHT_EINVAL, "null column array") : AHT_EINVAL;
  const uint32_t NXY = s->g.NXY;
  const int nz = s->g.nz;
  const int nx = s->g.nx, ny = s->g.ny;
  std::vector<uint32_t> ckv(NXY), nbv(NXY);
  std::vector<float2> cqv(NXY);
  int kmin = nz;
  for (uint32_t c = 0; c < NXY; c++) {
    if (kt[c] < 1 || kt[c] > ktop[c] || (int)ktop[c] > nz)
      return fail(s, AHT_EINVAL, "column %u: need 1 <= kt (%u) <= ktop (%u) <= nz (%d)", c, kt[c], ktop[c], nz);
    if (!std::isfinite(cd[c]) || !std::isfinite(q[c])) return fail(s, AHT_EINVAL, "column %u: non-finite cd or q", c);
    kmin = std::min(kmin, (int)ktop[c]);
  }
  for (int j = 0; j < ny; j++) {
    for (int i = 0; i < nx; i++) {
      const uint32_t c = (uint32_t)j * nx + i;
      // kfast: first level whose 3 x 3 x 3 source neighbourhood is all fluid (clamped at the faces)
      uint32_t m = 0;
      for (int b = std::max(j - 1, 0); b <= std::min(j + 1, ny - 1); b++)
        for (int a = std::max(i - 1, 0); a <= std::min(i + 1, nx - 1); a++) m = std::max(m, ktop[(size_t)b * nx + a]);
      const uint32_t kfast = std::min<uint32_t>(m + 1, 1023u);
      const bool near = pack_neighbours(ktop, nx, ny, i, j, nbv[c]);
      ckv[c] = kt[c] | (ktop[c] << 10) | (kfast << 20) | ((uint32_t)near << 30);
      cqv[c] = make_float2(cd[c], q[c]);
    }
  }
  int rc = set_device(s);
  if (rc) return rc;
  HIP_TRY(s, hipMemcpyAsync(s->colk, ckv.data(), NXY * sizeof(uint32_t), hipMemcpyHostToDevice, s->st));
  HIP_TRY(s, hipMemcpyAsync(s->colnb, nbv.data(), NXY * sizeof(uint32_t), hipMemcpyHostToDevice, s->st));
  HIP_TRY(s, hipMemcpyAsync(s->colcq, cqv.data(), NXY * sizeof(float2), hipMemcpyHostToDevice, s->st));
  HIP_TRY(s, hipMemcpyAsync(s->ktop, ktop, NXY * sizeof(uint32_t), hipMemcpyHostToDevice, s->st));
  HIP_TRY(s, hipStreamSynchronize(s->st));
  s->g.kmin = kmin;
  s->gdirty = true;

  s->h_q.assign(q, q + NXY);
  s->has_cols = true;
  return AHT_OK;
}

int aht_set_params(aht_solver *s, const aht_params *p, const float *prof) {
  if (!s || !p || !prof) return s ? fail(s, AHT_EINVAL, "null params") : AHT_EINVAL;
  if (p->nsp < 1 || p->nsw < 1) return fail(s, AHT_EINVAL, "nsp and nsw must be >= 1");
  if (!(p->prt > 0.0f)) return fail(s, AHT_EINVAL, "prt must be > 0");
  const float *pf = &p->tau0;
  for (int k = 0; k < 8; k++)
    if (!std::isfinite(pf[k])) return fail(s, AHT_EINVAL, "non-finite parameter %d", k);
  static_assert(sizeof(aht::Params) == sizeof(aht_params), "parameter block layout");
  memcpy(&s->g.P, p, sizeof(aht_params));
  s->gdirty = true;
  s->g.rp = aht::ParamRecip{(float)(1.0 / (double)p->prt), (float)(1.0 / (double)p->nsp), (float)(1.0 / (double)p->nsw)};
  int rc = set_device(s);
  if (rc) return rc;
  HIP_TRY(s, hipMemcpyAsync(s->prof, prof, s->g.nz * sizeof(float4), hipMemcpyHostToDevice, s->st));
  HIP_TRY(s, hipStreamSynchronize(s->st));
  s->has_params = true;
  return AHT_OK;
}

int aht_init_flow(aht_solver *s) {
  int rc = check_ready(s);
  if (rc) return rc;
  if ((rc = set_device(s))) return rc;
  HIP_TRY(s, aht::launch_init(s->storage, s->g, s->f[0], s->f[1], s->st));
  return AHT_OK;
}

int aht_reset_heat(aht_solver *s) {
  if (!s) return AHT_EINVAL;
  int rc = set_device(s);
  if (rc) return rc;
  if (s->g.lb) {   // the D3Q7 populations of a blocked pair lie between D3Q19 ones; the theta plane is one range
    for (int p = 0; p < 2; p++) {
      HIP_TRY(s, aht::launch_zero_g_blocked(s->storage, s->g, s->gq[p], s->st));
      if (aht::storage_theta_ref(s->storage))
        HIP_TRY(s, hipMemsetAsync((char *)s->gq[p] + s->g.th_off, 0, (size_t)s->g.N * 4, s->st));
    }
    return AHT_OK;
  }
  const size_t b = s->gbytes;                        // zero bits are 0.0 in every storage format (theta too)
  HIP_TRY(s, hipMemsetAsync(s->gq[0], 0, b, s->st));
  HIP_TRY(s, hipMemsetAsync(s->gq[1], 0, b, s->st));
  return AHT_OK;
}

int aht_reinit_cells(aht_solver *s, const uint32_t *old_ktop) {
  int rc = check_ready(s);
  if (rc) return rc;
  if (!old_ktop) return fail(s, AHT_EINVAL, "null old_ktop");
  if ((rc = set_device(s))) return rc;
  HIP_TRY(s, hipMemcpyAsync(s->oldk, old_ktop, s->g.NXY * sizeof(uint32_t), hipMemcpyHostToDevice, s->st));
  HIP_TRY(s, aht::launch_reinit(s->storage, s->g, s->oldk, s->f[s->parity], s->gq[s->parity], s->st));
  HIP_TRY(s, hipStreamSynchronize(s->st));
  return AHT_OK;
}

// the step kernel's device copy of the grid follows s->g; uploaded (rarely: columns, parameters, launch shape) after
// the stream has drained, so no launch still reads the old copy and the pinned staging copy is free to overwrite
static int upload_grid(aht_solver *s) {
  if (!s->gdirty) return AHT_OK;
  HIP_TRY(s, hipStreamSynchronize(s->st));
  s->gpin->g = s->g;
  for (int p = 0; p < 2; p++)
    s->gpin->s[p] = aht::make_streams(s->storage, s->g, s->f[p], s->f[1 - p], s->gq[p], s->gq[1 - p]);
  HIP_TRY(s, hipMemcpyAsync(s->gdev, s->gpin, sizeof(DevGrid), hipMemcpyHostToDevice, s->st));
  s->gdirty = false;
  return AHT_OK;
}

int aht_step(aht_solver *s, int32_t n, uint32_t flags) {
  int rc = check_ready(s);
  if (rc) return rc;
  if (n < 0) return fail(s, AHT_EINVAL, "negative step count");
  if ((rc = set_device(s))) return rc;
  if ((rc = upload_grid(s))) return rc;
  const bool acc = flags & AHT_STEP_ACC, out = flags & AHT_STEP_OUT;
  if (out && !s->out4) HIP_TRY(s, hipMalloc(&s->out4, 4 * (size_t)s->g.N * sizeof(float)));
  aht::Acc a{s->acc_surf, s->acc_face, s->acc_content};
  const bool sweep = s->sweep != 0 && !acc && aht::sweep2_supported(s->storage, s->g.lb);
  for (int32_t it = 0; it < n; it++) {
    const int p = s->parity;
    if (sweep && it + 1 < n && !(out && it + 1 == n - 1)) {   // steps it and it + 1 in one launch
      HIP_TRY(s, aht::launch_sweep2(s->storage, s->g, &s->gdev->g, s->f[p], s->f[1 - p], s->gq[p], s->gq[1 - p],
                                    s->sweep > 0 ? 32 : 16, s->sweep > 0 ? s->sweep : -s->sweep, s->g.win,
                                    (uint32_t)s->steps, s->st));
      s->parity ^= 1;
      s->steps += 2;
      it++;
      continue;
    }
    float *o4 = (out && it == n - 1) ? s->out4 : nullptr;
    if (o4) HIP_TRY(s, hipMemsetAsync(o4, 0, 4 * (size_t)s->g.N * sizeof(float), s->st));
    HIP_TRY(s, aht::launch_step(s->storage, s->g, &s->gdev->g, &s->gdev->s[p], s->f[p], s->f[1 - p], s->gq[p],
                                s->gq[1 - p], acc ? &a : nullptr, o4, s->block, s->levels, (uint32_t)s->steps, s->st,
                                aht::occupancy_lds(s->block, s->waves, s->lds_cu, s->lds_block)));
    s->parity ^= 1;
    s->steps++;
    if (acc) s->n_acc++;
    if (o4) s->has_out = true;
  }
  return AHT_OK;
}

int aht_sync(aht_solver *s) {
  if (!s) return AHT_EINVAL;
  int rc = set_device(s);
  if (rc) return rc;
  HIP_TRY(s, hipStreamSynchronize(s->st));
  return AHT_OK;
}

int64_t aht_steps_done(const aht_solver *s) { return s ? s->steps : -1; }

int aht_benchmark(aht_solver *s, int32_t n, double *ms) {
  int rc = check_ready(s);
  if (rc) return rc;
  if (!ms || n < 1) return fail(s, AHT_EINVAL, "benchmark needs n >= 1 and an output");
  if ((rc = set_device(s))) return rc;
  HIP_TRY(s, hipEventRecord(s->ev0, s->st));
  if ((rc = aht_step(s, n, 0))) return rc;
  HIP_TRY(s, hipEventRecord(s->ev1, s->st));
  HIP_TRY(s, hipEventSynchronize(s->ev1));
  float t = 0.0f;
  HIP_TRY(s, hipEventElapsedTime(&t, s->ev0, s->ev1));
  *ms = t;
  return AHT_OK;
}

int aht_read_surface(aht_solver *s, float conv, float *theta, float *u, float *v) {
  int rc = check_ready(s);
  if (rc) return rc;
  if (!theta) return fail(s, AHT_EINVAL, "null theta");
  if ((rc = set_device(s))) return rc;
  const int p = s->parity;
  const size_t NXY = s->g.NXY;
  HIP_TRY(s, aht::launch_surface(s->storage, s->g, s->f[p], s->gq[p], conv, s->surf3, s->st));
  HIP_TRY(s, hipMemcpyAsync(theta, s->surf3, NXY * sizeof(float), hipMemcpyDeviceToHost, s->st));
  if (u) HIP_TRY(s, hipMemcpyAsync(u, s->surf3 + NXY, NXY * sizeof(float), hipMemcpyDeviceToHost, s->st));
  if (v) HIP_TRY(s, hipMemcpyAsync(v, s->surf3 + 2 * NXY, NXY * sizeof(float), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(s, hipStreamSynchronize(s->st));
  return AHT_OK;
}

int aht_read_volume(aht_solver *s, int32_t nzv, float volmax, uint8_t *vol, float *theta_max) {
  int rc = check_ready(s);
  if (rc) return rc;
  if (nzv < 1 || nzv > s->g.nz) return fail(s, AHT_EINVAL, "nzv must be within 1..nz");
  if ((rc = set_device(s))) return rc;
  const float vm = volmax > 0.0f ? volmax : 3.0f;
  HIP_TRY(s, hipMemsetAsync(s->tmax, 0, sizeof(uint32_t), s->st));
  HIP_TRY(s, aht::launch_volume(s->storage, s->g, s->gq[s->parity], nzv, 255.0f / vm, s->vol, s->tmax, s->st));
  if (vol) HIP_TRY(s, hipMemcpyAsync(vol, s->vol, (size_t)nzv * s->g.NXY, hipMemcpyDeviceToHost, s->st));
  uint32_t tb = 0;
  HIP_TRY(s, hipMemcpyAsync(&tb, s->tmax, sizeof(uint32_t), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(s, hipStreamSynchronize(s->st));
  if (theta_max) memcpy(theta_max, &tb, 4);
  return AHT_OK;
}

int aht_stats(aht_solver *s, aht_diag *out) {
  int rc = check_ready(s);
  if (rc) return rc;
  if (!out) return fail(s, AHT_EINVAL, "null stats");
  if ((rc = set_device(s))) return rc;
  const int p = s->parity;
  HIP_TRY(s, aht::launch_stats(s->storage, s->g, s->f[p], s->gq[p], s->stat_part, s->red_out, s->st));
  double r[8];
  HIP_TRY(s, hipMemcpyAsync(r, s->red_out, 7 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(s, hipStreamSynchronize(s->st));
  out->sum_theta = r[0]; out->sum_ke = r[1]; out->sum_drho = r[2]; out->n_fluid = r[3];
  out->theta_max = r[4]; out->umax = r[5]; out->non_finite = (int64_t)r[6];
  return AHT_OK;
}

}  // extern "C"

// Scratch of aht_stats_noalloc: the partials of the diagnostics (7 per stats block, at most STATS_MAX_BLOCKS blocks)
// and the 7 results after them (8 slots), 114,752 bytes held by every GPU the library runs on until the process ends. In host
// code `scratch` names a host-side stand-in, not device memory: the device address comes from hipGetSymbolAddress,
// once per device with that device current, and only that address is passed to kernels.
constexpr size_t SCRATCH_DOUBLES = 7 * (size_t)aht::STATS_MAX_BLOCKS + 8;
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

namespace {
constexpr int SCRATCH_DEVICES = 64;
double *scratch_dev[SCRATCH_DEVICES];   // device address of `scratch` per device ordinal (null: not looked up yet)
std::mutex scratch_mu;

// the device address of this device's scratch (s's device is current)
int device_scratch(aht_solver *s, double **out) {
  if (s->device < 0 || s->device >= SCRATCH_DEVICES)
    return fail(s, AHT_EINVAL, "allocation-free diagnostics: device %d beyond the %d devices served", s->device,
                SCRATCH_DEVICES);
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[s->device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(s, AHT_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", s->device, hipGetErrorString(e));
    scratch_dev[s->device] = (double *)p;
  }
  *out = scratch_dev[s->device];
  return AHT_OK;
}
}  // namespace

extern "C" {

int aht_stats_noalloc(aht_solver *s, aht_diag *out) {
  int rc = check_ready(s);
  if (rc) return rc;
  if (!out) return fail(s, AHT_EINVAL, "null stats");
  if ((rc = set_device(s))) return rc;
  const uint32_t nb = aht::stats_blocks(s->g);
  if (7 * (size_t)nb + 8 > SCRATCH_DOUBLES)
    return fail(s, AHT_EINVAL, "allocation-free diagnostics: %u partial blocks exceed the scratch's %u", nb,
                aht::STATS_MAX_BLOCKS);
  double *ptr = nullptr;
  if ((rc = device_scratch(s, &ptr))) return rc;
  double *res = ptr + 7 * (size_t)aht::STATS_MAX_BLOCKS;   // every partial and result is written before it is read
  const int p = s->parity;
  HIP_TRY(s, aht::launch_stats(s->storage, s->g, s->f[p], s->gq[p], ptr, res, s->st));
  double r[8];
  HIP_TRY(s, hipMemcpyAsync(r, res, 7 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(s, hipStreamSynchronize(s->st));
  out->sum_theta = r[0]; out->sum_ke = r[1]; out->sum_drho = r[2]; out->n_fluid = r[3];
  out->theta_max = r[4]; out->umax = r[5]; out->non_finite = (int64_t)r[6];
  return AHT_OK;
}

int aht_read_out(aht_solver *s, float *out4) {
  if (!s || !out4) return AHT_EINVAL;
  if (!s->has_out) return fail(s, AHT_ESTATE, "no AHT_STEP_OUT step has run");
  int rc = set_device(s);
  if (rc) return rc;
  HIP_TRY(s, hipMemcpyAsync(out4, s->out4, 4 * (size_t)s->g.N * sizeof(float), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(s, hipStreamSynchronize(s->st));
  return AHT_OK;
}

int aht_mean_reset(aht_solver *s) {
  if (!s) return AHT_EINVAL;
  int rc = set_device(s);
  if (rc) return rc;
  const aht::Grid &g = s->g;
  HIP_TRY(s, hipMemsetAsync(s->acc_surf, 0, 3 * (size_t)g.NXY * sizeof(double), s->st));
  HIP_TRY(s, hipMemsetAsync(s->acc_face, 0, face_total(g.nx, g.ny, g.nz) * sizeof(double), s->st));
  HIP_TRY(s, hipMemsetAsync(s->acc_content, 0, s->content_cap * sizeof(double), s->st));
  s->n_acc = 0;
  return AHT_OK;
}

int aht_read_face_mean(aht_solver *s, int32_t face, double *out) {
  if (!s || !out || face < 0 || face > 3) return s ? fail(s, AHT_EINVAL, "face must be 0..3") : AHT_EINVAL;
  int rc = set_device(s);
  if (rc) return rc;
  const aht::Grid &g = s->g;
  const size_t len = aht::face_len(face, g.nx, g.ny, g.nz), off = aht::face_offset(face, g.nx, g.ny, g.nz);
  HIP_TRY(s, hipMemcpyAsync(out, s->acc_face + off, len * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(s, hipStreamSynchronize(s->st));
  const double inv = s->n_acc > 0 ? 1.0 / (double)s->n_acc : 0.0;
  for (size_t k = 0; k < len; k++) out[k] *= inv;
  return AHT_OK;
}

int aht_read_mean(aht_solver *s, double *surf3, aht_budget *b) {
  if (!s) return AHT_EINVAL;
  int rc = set_device(s);
  if (rc) return rc;
  const aht::Grid &g = s->g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t NXY = g.NXY;
  const double inv = s->n_acc > 0 ? 1.0 / (double)s->n_acc : 0.0;
  if (surf3) {
    std::vector<double> h(3 * NXY);
    HIP_TRY(s, hipMemcpyAsync(h.data(), s->acc_surf, 3 * NXY * sizeof(double), hipMemcpyDeviceToHost, s->st));
    HIP_TRY(s, hipStreamSynchronize(s->st));
    for (int comp = 0; comp < 3; comp++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++)
          surf3[comp * NXY + (size_t)(ny - 1 - j) * nx + i] = h[comp * NXY + (size_t)j * nx + i] * inv;
  }
  if (b) {
    memset(b, 0, sizeof *b);
    b->n_acc = s->n_acc;
    double hin = 0.0;
    for (float x : s->h_q) hin += (double)x;
    b->heat_in = hin;
    HIP_TRY(s, aht::launch_sum(s->acc_content, s->content_cap, s->red_out, s->st));
    double content = 0.0;
    HIP_TRY(s, hipMemcpyAsync(&content, s->red_out, sizeof(double), hipMemcpyDeviceToHost, s->st));
    std::vector<double> fc(face_total(nx, ny, nz));
    HIP_TRY(s, hipMemcpyAsync(fc.data(), s->acc_face, fc.size() * sizeof(double), hipMemcpyDeviceToHost, s->st));
    HIP_TRY(s, hipStreamSynchronize(s->st));
    b->content = content * inv;
    // heat_diag on the mean face fields: out = sum_E th*u - sum_W th*u + sum_N th*v - sum_S th*v
    double hf[4] = {0, 0, 0, 0}, qf[4] = {0, 0, 0, 0};
    for (int f = 0; f < 4; f++) {
      const size_t len = aht::face_len(f, nx, ny, nz) / 2, off = aht::face_offset(f, nx, ny, nz);
      for (size_t k = 0; k < len; k++) {
        const double th = fc[off + k] * inv, un = fc[off + len + k] * inv;
        hf[f] += th * un;
        qf[f] += un;
      }
    }
    b->heat_out = hf[1] - hf[0] + hf[3] - hf[2];
    b->q_w = qf[0]; b->q_e = qf[1]; b->q_s = qf[2]; b->q_n = qf[3];
    const double thru = std::fabs(qf[0]) + std::fabs(qf[1]) + std::fabs(qf[2]) + std::fabs(qf[3]);
    b->mass_imbalance = std::fabs(qf[0] - qf[1] + qf[2] - qf[3]) / std::max(thru, 1e-300);
  }
  return AHT_OK;
}

// state exchange through one staging plane: storage <-> f32 / f64, one population at a time
static int state_io(aht_solver *s, void *f, void *g, int f64, int write) {
  if (!s) return AHT_EINVAL;
  int rc = set_device(s);
  if (rc) return rc;
  const size_t N = s->g.N;
  const int p = s->parity;
  const size_t hb = f64 ? 8 : 4;
  // f32 storage in the reference's [q][cell] layout and f32 host format: plain copies
  if (s->storage == AHT_STORE_F32 && s->g.lb == 0 && !f64) {
    if (write) {
      if (f) HIP_TRY(s, hipMemcpyAsync(s->f[p], f, 19 * N * 4, hipMemcpyHostToDevice, s->st));
      if (g) HIP_TRY(s, hipMemcpyAsync(s->gq[p], g, 7 * N * 4, hipMemcpyHostToDevice, s->st));
    } else {
      if (f) HIP_TRY(s, hipMemcpyAsync(f, s->f[p], 19 * N * 4, hipMemcpyDeviceToHost, s->st));
      if (g) HIP_TRY(s, hipMemcpyAsync(g, s->gq[p], 7 * N * 4, hipMemcpyDeviceToHost, s->st));
    }
    HIP_TRY(s, hipStreamSynchronize(s->st));
    return AHT_OK;
  }
  if (!s->stage) HIP_TRY(s, hipMalloc(&s->stage, N * 8));
  const float w0 = (float)(1.0 / 3.0), w1 = (float)(1.0 / 18.0), w2 = (float)(1.0 / 36.0);
  const float W19[19] = {w0, w1, w1, w1, w1, w1, w1, w2, w2, w2, w2, w2, w2, w2, w2, w2, w2, w2, w2};
  const float W7[7] = {0.25f, 0.125f, 0.125f, 0.125f, 0.125f, 0.125f, 0.125f};
  const bool thref = aht::storage_theta_ref(s->storage);
  for (int is_f = 1; is_f >= 0; is_f--) {
    char *host = (char *)(is_f ? f : g);
    if (!host) continue;
    char *dev = (char *)(is_f ? s->f[p] : s->gq[p]);
    const float *th = (!is_f && thref) ? (const float *)(dev + s->g.th_off) : nullptr;
    if (th && write) {
      // theta plane of the lattice: theta = sum_q g_q of the state being written, then the deviations against it
      std::vector<float> t(N);
      for (size_t n = 0; n < N; n++) {
        double a = 0.0;
        for (int q = 0; q < 7; q++)
          a += f64 ? ((const double *)host)[(size_t)q * N + n] : (double)((const float *)host)[(size_t)q * N + n];
        t[n] = (float)a;
      }
      HIP_TRY(s, hipMemcpyAsync((void *)th, t.data(), N * 4, hipMemcpyHostToDevice, s->st));
      HIP_TRY(s, hipStreamSynchronize(s->st));
    }
    const int nq = is_f ? 19 : 7;
    for (int q = 0; q < nq; q++) {
      const float w = is_f ? W19[q] : W7[q];
      if (write) {
        HIP_TRY(s, hipMemcpyAsync(s->stage, host + (size_t)q * N * hb, N * hb, hipMemcpyHostToDevice, s->st));
        HIP_TRY(s, aht::launch_encode(s->storage, s->g.lb, s->stage, f64, dev, q, (uint32_t)N, w, is_f, th, s->st));
      } else {
        HIP_TRY(s, aht::launch_decode(s->storage, s->g.lb, dev, q, s->stage, f64, (uint32_t)N, w, is_f, th, s->st));
        HIP_TRY(s, hipMemcpyAsync(host + (size_t)q * N * hb, s->stage, N * hb, hipMemcpyDeviceToHost, s->st));
      }
    }
  }
  HIP_TRY(s, hipStreamSynchronize(s->st));
  return AHT_OK;
}

int aht_read_state(aht_solver *s, float *f, float *g) { return state_io(s, f, g, 0, 0); }
int aht_write_state(aht_solver *s, const float *f, const float *g) {
  return state_io(s, (void *)f, (void *)g, 0, 1);
}
int aht_read_state_f64(aht_solver *s, double *f, double *g) { return state_io(s, f, g, 1, 0); }
int aht_write_state_f64(aht_solver *s, const double *f, const double *g) {
  return state_io(s, (void *)f, (void *)g, 1, 1);
}

// ---- batches ------------------------------------------------------------------------------------------------------
struct aht_batch {
  int device = 0, storage = 0;
  hipStream_t st = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<aht_solver *> mem;       // add order
  aht::Member *dmem = nullptr;         // device table, ordered by the chunk's step counts (descending)
  size_t dcap = 0;
  std::vector<aht::Member> hmem;
  std::vector<int> order;              // table row -> member (add order)
  uint32_t total = 0;                  // tiles of every member, at the block size of the last chunk
  int version = 0;                     // bumps on membership change
  int grid = 0;                        // fixed grid (blocks), 0 = tuned
  int block = 0;                       // fixed block size (threads), 0 = tuned
  bool use_graph = true;
  // the captured chunk: valid while the block size and every launch's (members, tiles, grid) are unchanged (a change
  // of a member's columns or levels changes its tile count, and with it the launches)
  hipGraphExec_t gexec = nullptr;
  int g_block = -1, g_grid = -1, g_waves = -1;
  std::vector<uint64_t> g_key;
  // online tuner: (block, grid, wavefronts-per-SIMD cap) candidates timed on real chunks as device time per cell-step
  // (chunks with different members compare), tuned again when the chunk's cell count moves by more than 2x from where
  // the last tuning started
  struct Cand {
    int block, grid, waves;
  };
  std::vector<Cand> cand;
  std::vector<double> cand_ms;
  int tune_next = 0;
  double tune_cells = 0.0;
  int last_block = 0, last_grid = 0, last_waves = 0;   // used by the most recent chunk
  uint32_t lds_cu = 0, lds_block = 0;                  // the device's LDS per CU and per block (bytes)
  int64_t steps = 0;
  std::string err;
};

namespace {
int bfail(aht_batch *b, int code, const char *fmt, ...) {
  if (b) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    b->err = buf;
  }
  return code;
}
#define HIP_TRYB(b, call)                                                                                   \
  do {                                                                                                      \
    hipError_t e_ = (call);                                                                                 \
    if (e_ != hipSuccess) return bfail((b), e_ == hipErrorOutOfMemory ? AHT_ENOMEM : AHT_EDEVICE, "%s: %s", \
                                       #call, hipGetErrorString(e_));                                       \
  } while (0)

int cu_count(int device) {
  int n = 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || n <= 0) n = 256;
  return n;
}

constexpr int ONE_BLOCK_PER_TILE = 1 << 30;

// candidates of the batch tuner: fixed values or, when tuned, block sizes 256 / 128 / 64 and grids of one block per
// tile or 32 persistent blocks per CU; one block per tile also with its wavefronts capped at 5 and 4 per SIMD
void reset_tuner(aht_batch *b) {
  std::vector<int> blocks = b->block ? std::vector<int>{b->block} : std::vector<int>{256, 128, 64};
  std::vector<int> grids = b->grid ? std::vector<int>{b->grid} : std::vector<int>{ONE_BLOCK_PER_TILE, 32 * cu_count(b->device)};
  b->cand.clear();
  for (int g : grids)
    for (int bl : blocks)
      for (int wv : {0, 5, 4}) {
        if (wv && (g != ONE_BLOCK_PER_TILE || !aht::occupancy_lds(bl, wv, b->lds_cu, b->lds_block))) continue;
        b->cand.push_back(aht_batch::Cand{bl, g, wv});
      }
  b->cand_ms.assign(b->cand.size(), 0.0);
  b->tune_next = 0;
  b->tune_cells = 0.0;
}

size_t best_candidate(const aht_batch *b) {
  size_t best = 0;
  for (size_t k = 1; k < b->cand.size(); k++)
    if (b->cand_ms[k] < b->cand_ms[best]) best = k;
  return best;
}

// tiles of one step of solver s with blocks of `block` columns, one level per thread (batched launches step every
// member so, whatever its own launch shape: a kernel with a per-member level loop held 30-40 more VGPRs)
void member_tiles(const aht_solver *s, int block, aht::Member &M) {
  const aht::Grid &g = s->g;
  M.bx = (g.NXY + (uint32_t)block - 1) / (uint32_t)block;
  M.inv_bx = 1.0f / (float)M.bx;
  M.tiles = g.kmin < g.nz ? M.bx * (uint32_t)(g.nz - g.kmin) : 0;
}
}  // namespace

int aht_batch_create(int32_t device, int32_t storage, aht_batch **out) {
  if (!out) return AHT_EINVAL;
  *out = nullptr;
  if (storage < AHT_STORE_F32 || storage > AHT_STORE_F16S) return AHT_EINVAL;
  aht_batch *b = new aht_batch();
  *out = b;
  b->device = device;
  b->storage = storage;
  HIP_TRYB(b, hipSetDevice(device));
  HIP_TRYB(b, hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking));
  HIP_TRYB(b, hipEventCreate(&b->ev0));
  HIP_TRYB(b, hipEventCreate(&b->ev1));
  int cu = 0, blk = 0;
  HIP_TRYB(b, hipDeviceGetAttribute(&cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device));
  HIP_TRYB(b, hipDeviceGetAttribute(&blk, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
  b->lds_cu = (uint32_t)std::max(cu, 0);
  b->lds_block = (uint32_t)std::max(blk, 0);
  reset_tuner(b);
  return AHT_OK;
}

void aht_batch_destroy(aht_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->st) (void)hipStreamSynchronize(b->st);
  for (aht_solver *s : b->mem) {
    s->batch = nullptr;
    s->batch_pending = false;
  }
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->dmem) (void)hipFree(b->dmem);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
}

const char *aht_batch_last_error(const aht_batch *b) { return b ? b->err.c_str() : "null batch"; }

int aht_batch_add(aht_batch *b, aht_solver *s) {
  if (!b || !s) return AHT_EINVAL;
  if (s->batch) return bfail(b, AHT_ESTATE, "solver already belongs to a batch");
  if (s->device != b->device || s->storage != b->storage)
    return bfail(b, AHT_EINVAL, "solver device/storage (%d/%d) differ from the batch's (%d/%d)", s->device, s->storage,
                 b->device, b->storage);
  if (!s->has_cols || !s->has_params) return bfail(b, AHT_ESTATE, "solver needs columns and parameters first");
  if (!b->mem.empty() && b->mem[0]->g.lb != s->g.lb)
    return bfail(b, AHT_EINVAL, "batch members share one population layout (lattice pairs of 4 GiB and more use the "
                                "[q][cell] layout, smaller ones the blocked one)");
  b->mem.push_back(s);
  s->batch = b;
  b->version++;
  return AHT_OK;
}

int aht_batch_remove(aht_batch *b, aht_solver *s) {
  if (!b || !s) return AHT_EINVAL;
  for (size_t k = 0; k < b->mem.size(); k++) {
    if (b->mem[k] == s) {
      (void)hipSetDevice(b->device);
      (void)hipStreamSynchronize(b->st);       // the member's last chunk has finished
      b->mem.erase(b->mem.begin() + (long)k);
      s->batch = nullptr;
      s->batch_pending = false;
      b->version++;
      return AHT_OK;
    }
  }
  return bfail(b, AHT_EINVAL, "solver is not a member of this batch");
}

int aht_batch_size(const aht_batch *b) { return b ? (int)b->mem.size() : -1; }

int aht_batch_set_grid(aht_batch *b, int32_t blocks, int32_t use_graph) {
  if (!b || blocks < 0) return AHT_EINVAL;
  b->grid = blocks;
  b->use_graph = use_graph != 0;
  reset_tuner(b);
  return AHT_OK;
}

int aht_batch_set_block(aht_batch *b, int32_t block) {
  if (!b) return AHT_EINVAL;
  if (block != 0 && block != 64 && block != 128 && block != 256)
    return bfail(b, AHT_EINVAL, "batch block must be 0 (tuned), 64, 128 or 256");
  b->block = block;
  reset_tuner(b);
  return AHT_OK;
}

int aht_batch_step(aht_batch *b, int32_t n, const int32_t *nsteps, const int32_t *acc_from, double *ms) {
  if (!b || n < 0) return AHT_EINVAL;
  if (ms) *ms = 0.0;
  if (n == 0 || b->mem.empty()) return AHT_OK;
  HIP_TRYB(b, hipSetDevice(b->device));
  const size_t nm = b->mem.size();
  for (size_t d = 0; d < nm; d++) {
    const int ns = nsteps ? nsteps[d] : n;
    if (ns < 0 || ns > n) return bfail(b, AHT_EINVAL, "member %zu: nsteps %d outside 0..%d", d, ns, n);
  }
  // (block, grid): fixed, or the online tuner's next candidate / its pick
  double cells = 0.0;                                 // cell-steps of this chunk (all cells of the fluid levels)
  for (size_t d = 0; d < nm; d++) {
    const aht::Grid &g = b->mem[d]->g;
    cells += (double)g.NXY * (double)std::max(0, g.nz - g.kmin) * (double)(nsteps ? nsteps[d] : n);
  }
  int cand_idx = -1;
  size_t pick;
  if (b->cand.size() > 1) {
    if (b->tune_cells == 0.0 || cells > 2 * b->tune_cells || 2 * cells < b->tune_cells) {
      std::fill(b->cand_ms.begin(), b->cand_ms.end(), 0.0);
      b->tune_next = 0;
      b->tune_cells = std::max(cells, 1.0);
    }
    if (b->tune_next < (int)b->cand.size()) cand_idx = b->tune_next++;
    pick = cand_idx >= 0 ? (size_t)cand_idx : best_candidate(b);
  } else {
    pick = 0;
  }
  const int block = b->cand[pick].block, grid = b->cand[pick].grid, waves = b->cand[pick].waves;
  const uint32_t lds = aht::occupancy_lds(block, waves, b->lds_cu, b->lds_block);
  b->last_block = block;
  b->last_grid = grid;
  b->last_waves = waves;
  // member table, longest-stepping members first: the members that step in launch `it` are then a prefix of it
  b->order.resize(nm);
  for (size_t d = 0; d < nm; d++) b->order[d] = (int)d;
  std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) {
    return (nsteps ? nsteps[x] : n) > (nsteps ? nsteps[y] : n);
  });
  b->hmem.resize(nm);
  uint32_t base = 0;
  for (size_t r = 0; r < nm; r++) {
    const int d = b->order[r];
    aht_solver *s = b->mem[d];
    aht::Member &M = b->hmem[r];
    M.g = s->g;
    M.f[0] = s->f[0]; M.f[1] = s->f[1]; M.gq[0] = s->gq[0]; M.gq[1] = s->gq[1];
    M.acc = aht::Acc{s->acc_surf, s->acc_face, s->acc_content};
    member_tiles(s, block, M);
    M.tile_base = base;
    base += M.tiles;
    M.parity0 = s->parity;
    M.nsteps = nsteps ? nsteps[d] : n;
    M.acc_from = acc_from ? acc_from[d] : n;
    M.step0 = (uint32_t)s->steps;
    for (int q = 0; q < 2; q++) M.s[q] = aht::make_streams(s->storage, s->g, s->f[q], s->f[1 - q], s->gq[q], s->gq[1 - q]);
    // the batch stream waits for work the member enqueued on its own stream
    HIP_TRYB(b, hipEventRecord(s->own, s->st));
    HIP_TRYB(b, hipStreamWaitEvent(b->st, s->own, 0));
  }
  b->total = base;
  // launch it covers the first m members (those with nsteps > it) and their tiles
  std::vector<uint64_t> key((size_t)n);
  std::vector<std::pair<int, uint32_t>> launches((size_t)n);
  {
    size_t m = nm;
    for (int it = 0; it < n; it++) {
      while (m > 0 && b->hmem[m - 1].nsteps <= it) m--;
      const uint32_t tiles = m < nm ? b->hmem[m].tile_base : base;
      launches[(size_t)it] = {(int)m, tiles};
      key[(size_t)it] = ((uint64_t)m << 32) | (uint64_t)tiles;
    }
  }
  if (b->dcap < nm) {
    HIP_TRYB(b, hipStreamSynchronize(b->st));        // a captured chunk may still read the old table
    if (b->dmem) HIP_TRYB(b, hipFree(b->dmem));
    b->dmem = nullptr;
    b->dcap = nm + 16;
    HIP_TRYB(b, hipMalloc(&b->dmem, b->dcap * sizeof(aht::Member)));
    if (b->gexec) HIP_TRYB(b, hipGraphExecDestroy(b->gexec));
    b->gexec = nullptr;
  }
  HIP_TRYB(b, hipMemcpyAsync(b->dmem, b->hmem.data(), nm * sizeof(aht::Member), hipMemcpyHostToDevice, b->st));
  const int lb = b->mem[0]->g.lb;
  auto enqueue = [&](int it) {
    const auto &L = launches[(size_t)it];
    return aht::launch_batch_step(b->storage, lb, block, b->dmem, L.first, L.second, it, grid, b->st, lds);
  };
  if (b->use_graph) {
    if (!b->gexec || b->g_block != block || b->g_grid != grid || b->g_waves != waves || b->g_key != key) {
      if (b->gexec) HIP_TRYB(b, hipGraphExecDestroy(b->gexec));
      b->gexec = nullptr;
      hipGraph_t graph;
      HIP_TRYB(b, hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
      for (int it = 0; it < n; it++) {
        const hipError_t e = enqueue(it);
        if (e != hipSuccess) {
          hipGraph_t dummy;
          (void)hipStreamEndCapture(b->st, &dummy);
          return bfail(b, AHT_EDEVICE, "batch launch during capture: %s", hipGetErrorString(e));
        }
      }
      HIP_TRYB(b, hipStreamEndCapture(b->st, &graph));
      HIP_TRYB(b, hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0));
      HIP_TRYB(b, hipGraphDestroy(graph));
      b->g_block = block;
      b->g_grid = grid;
      b->g_waves = waves;
      b->g_key = key;
    }
    HIP_TRYB(b, hipEventRecord(b->ev0, b->st));       // after the (host-side) capture: times the device work only
    HIP_TRYB(b, hipGraphLaunch(b->gexec, b->st));
  } else {
    HIP_TRYB(b, hipEventRecord(b->ev0, b->st));
    for (int it = 0; it < n; it++) HIP_TRYB(b, enqueue(it));
  }
  HIP_TRYB(b, hipEventRecord(b->ev1, b->st));
  // member bookkeeping; their next own-stream work waits for this chunk
  for (size_t d = 0; d < nm; d++) {
    aht_solver *s = b->mem[d];
    const int ns = nsteps ? nsteps[d] : n, af = acc_from ? acc_from[d] : n;
    s->parity ^= (ns & 1);
    s->steps += ns;
    s->n_acc += std::max(0, ns - std::max(0, af));
    HIP_TRYB(b, hipEventRecord(s->after_batch, b->st));
    s->batch_pending = true;
  }
  b->steps += n;
  if (cand_idx >= 0 || ms) {
    HIP_TRYB(b, hipEventSynchronize(b->ev1));
    float t = 0.0f;
    HIP_TRYB(b, hipEventElapsedTime(&t, b->ev0, b->ev1));
    if (cand_idx >= 0) b->cand_ms[cand_idx] = t / std::max(cells, 1.0);
    if (ms) *ms = t;
  }
  return AHT_OK;
}

int aht_batch_sync(aht_batch *b) {
  if (!b) return AHT_EINVAL;
  HIP_TRYB(b, hipSetDevice(b->device));
  HIP_TRYB(b, hipStreamSynchronize(b->st));
  return AHT_OK;
}

int aht_batch_grid(const aht_batch *b) {
  if (!b) return -1;
  return b->last_grid ? (int)std::min<uint32_t>((uint32_t)b->last_grid, std::max<uint32_t>(b->total, 1)) : 0;
}

int aht_batch_block(const aht_batch *b) { return b ? b->last_block : -1; }

// ---- launch-shape autotuner for one solver ------------------------------------------------------------------------
namespace {
int check_launch(aht_solver *s, const aht_launch &l) {
  if (l.block < 64 || l.block > 256 || l.block % 64) return fail(s, AHT_EINVAL, "block must be 64, 128, 192 or 256");
  if (l.levels != 1 && l.levels != 2 && l.levels != 4) return fail(s, AHT_EINVAL, "levels must be 1, 2 or 4");
  if (l.window != -1 && (l.window < 0 || l.window > 4096 || (l.window & (l.window - 1))))
    return fail(s, AHT_EINVAL, "window must be -1 (plane slabs), 0 (slabs) or a power of two up to 4096");
  if (l.sweep < -1023 || l.sweep > 1023) return fail(s, AHT_EINVAL, "sweep chunk must be at most 1023 levels");
  if (l.waves < 0 || l.waves > 8) return fail(s, AHT_EINVAL, "waves must be 0 (no cap) or 1 to 8");
  if (l.waves && !aht::occupancy_lds(l.block, l.waves, s->lds_cu, s->lds_block))
    return fail(s, AHT_EINVAL, "a cap of %d wavefronts per SIMD cannot be set for blocks of %d threads on this device",
                l.waves, l.block);
  return AHT_OK;
}
void apply_launch(aht_solver *s, const aht_launch &l) {
  s->block = l.block;
  s->levels = l.levels;
  s->sweep = l.sweep;
  s->waves = l.waves;
  s->g.win = (uint32_t)l.window;
  s->gdirty = true;
}
}  // namespace

int aht_autotune(aht_solver *s, int32_t steps, aht_launch *best) {
  int rc = check_ready(s);
  if (rc) return rc;
  if (steps < 1) return fail(s, AHT_EINVAL, "autotune needs steps >= 1");
  if ((rc = set_device(s))) return rc;
  const size_t fb = s->fbytes, gb = s->gbytes;
  const int p0 = s->parity;
  const int64_t steps0 = s->steps, nacc0 = s->n_acc;
  void *bf = nullptr, *bg = nullptr;   // gb = 0 for a blocked pair: fb covers both lattices
  HIP_TRY(s, hipMalloc(&bf, fb));
  if (gb) {
    const hipError_t eg = hipMalloc(&bg, gb);
    if (eg != hipSuccess) { (void)hipFree(bf); return fail(s, AHT_ENOMEM, "autotune backup: %s", hipGetErrorString(eg)); }
  }
  HIP_TRY(s, hipMemcpyAsync(bf, s->f[p0], fb, hipMemcpyDeviceToDevice, s->st));
  if (gb) HIP_TRY(s, hipMemcpyAsync(bg, s->gq[p0], gb, hipMemcpyDeviceToDevice, s->st));
  // every shape over `steps` timed steps (after one untimed step that also uploads the changed grid), then the six
  // fastest again in three interleaved rounds over 10 x `steps`, best of the three: neighbouring shapes differ by
  // ~1 %, about the noise of a short timing (with the four fastest twice over 3 x `steps` the 600 x 600 x 64 f32 step
  // once kept 256,2,1 at 0.829 ms over 256,1,1 at 0.819 ms)
  const int cb[3] = {64, 128, 256}, cl[3] = {1, 2, 4}, cw[7] = {1, 2, 4, 8, 16, 0, -1};
  struct Cand {
    aht_launch l;
    double ms;
  };
  std::vector<Cand> cand;
  aht_launch keep{s->block, s->levels, (int32_t)s->g.win, s->sweep, s->waves};
  auto time_shape = [&](const aht_launch &l, int32_t n, double &ms) {
    apply_launch(s, l);
    int r = aht_step(s, 1, 0);
    if (!r) r = aht_benchmark(s, n, &ms);
    return r;
  };
  for (int b : cb)
    for (int l : cl)
      for (int w : cw) {
        if (rc) break;
        double ms = 0;
        rc = time_shape(aht_launch{b, l, w, 0, 0}, steps, ms);
        cand.push_back(Cand{aht_launch{b, l, w, 0, 0}, ms / steps});
      }
  // the three fastest shapes again with their wavefronts capped at 5 and 4 per SIMD
  if (!rc) {
    std::sort(cand.begin(), cand.end(), [](const Cand &a, const Cand &b) { return a.ms < b.ms; });
    const size_t top = std::min<size_t>(cand.size(), 3);
    for (size_t t = 0; t < top && !rc; t++)
      for (int wv : {5, 4}) {
        aht_launch l = cand[t].l;
        l.waves = wv;
        if (!aht::occupancy_lds(l.block, wv, s->lds_cu, s->lds_block)) continue;
        double ms = 0;
        if ((rc = time_shape(l, steps, ms))) break;
        cand.push_back(Cand{l, ms / steps});
      }
  }
  if (!rc) {
    std::sort(cand.begin(), cand.end(), [](const Cand &a, const Cand &b) { return a.ms < b.ms; });
    cand.resize(std::min<size_t>(cand.size(), 6));
    for (Cand &c : cand) c.ms = 1e30;
    for (int round = 0; round < 3 && !rc; round++)
      for (Cand &c : cand) {
        double ms = 0;
        if ((rc = time_shape(c.l, 10 * steps, ms))) break;
        c.ms = std::min(c.ms, ms / (10 * steps));
      }
    if (!rc) keep = std::min_element(cand.begin(), cand.end(), [](const Cand &a, const Cand &b) { return a.ms < b.ms; })->l;
  }
  apply_launch(s, keep);
  // restore the state and the counters: tuning leaves the simulation where it was
  (void)hipMemcpyAsync(s->f[p0], bf, fb, hipMemcpyDeviceToDevice, s->st);
  if (gb) (void)hipMemcpyAsync(s->gq[p0], bg, gb, hipMemcpyDeviceToDevice, s->st);
  (void)hipStreamSynchronize(s->st);
  (void)hipFree(bf);
  if (bg) (void)hipFree(bg);
  s->parity = p0; s->steps = steps0; s->n_acc = nacc0;
  if (rc) return rc;
  if (best) *best = keep;
  return AHT_OK;
}

int aht_set_launch(aht_solver *s, const aht_launch *l) {
  if (!s) return AHT_EINVAL;
  if (!l) return fail(s, AHT_EINVAL, "null launch shape");
  const int rc = check_launch(s, *l);
  if (rc) return rc;
  apply_launch(s, *l);
  return AHT_OK;
}

int aht_get_launch(const aht_solver *s, aht_launch *l) {
  if (!s || !l) return AHT_EINVAL;
  *l = aht_launch{s->block, s->levels, (int32_t)s->g.win, s->sweep, s->waves};
  return AHT_OK;
}

}  // extern "C"
