// A CPU stand-in of libaht.so's C-ABI (include/aht.h) for testing the N-API addon's threading and lifetime rules
// without a GPU. Test infrastructure only: it computes nothing, it checks how the addon calls the library.
//
// It models the two hazards the addon must avoid, and aborts with a message naming the one that happened:
//  * a runtime with per-thread state: every thread that calls a device entry point gets thread-local state whose
//    destructor needs the runtime, and the runtime itself is a static object destroyed at process exit (as a GPU
//    runtime's is). A thread that makes device calls and exits after that destructor (libuv's pool threads are joined
//    by a library destructor at exit) tears its state down against a finished runtime;
//  * the C-ABI's one-caller-per-object rule: two calls running at once on one solver or batch, or a solver or batch
//    destroyed while a call on it is still running.
#include <unistd.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "aht.h"

namespace {

struct Runtime {
  std::atomic<int> alive{1};
  ~Runtime() { alive = 0; }
};
Runtime runtime;

[[noreturn]] void fail(const char *what) {
  std::fprintf(stderr, "STANDIN FAILURE: %s\n", what);
  std::fflush(stderr);
  std::abort();
}

struct ThreadState {
  int used = 0;
  ~ThreadState() {
    if (used && !runtime.alive) fail("a thread that made device calls exited after the runtime was torn down");
  }
};
thread_local ThreadState tstate;

void device_call() {
  if (!runtime.alive) fail("device call after the runtime was torn down");
  tstate.used = 1;
}

struct Busy {          // one call at a time per object, and never destroyed while a call runs
  std::atomic<int> running{0};
  int destroyed = 0;
};

struct Guard {
  Busy &b;
  explicit Guard(Busy &x) : b(x) {
    if (b.destroyed) fail("call on a destroyed object");
    if (b.running.fetch_add(1)) fail("two calls at once on one object");
  }
  ~Guard() {
    if (b.destroyed) fail("object destroyed while a call on it was running");
    b.running.fetch_sub(1);
  }
};

void work(int n) { usleep(1000u * (unsigned)(n < 50 ? n : 50)); }

}  // namespace

struct aht_solver {
  Busy busy;
  aht_dims d;
  int64_t steps = 0;
  aht_launch launch{256, 1, 1, 0};
};

struct aht_batch {
  Busy busy;
  int n = 0;
  int block = 0, grid = 0;
};

extern "C" {

int aht_create(const aht_dims *dims, aht_solver **out) {
  device_call();
  if (!dims || !out) return AHT_EINVAL;
  *out = new aht_solver();
  (*out)->d = *dims;
  return AHT_OK;
}

void aht_destroy(aht_solver *s) {
  device_call();
  if (!s) return;
  if (s->busy.running) fail("solver destroyed while a call on it was running");
  s->busy.destroyed = 1;   // kept (not freed), so later misuse is detected rather than undefined
}

const char *aht_last_error(const aht_solver *) { return "stand-in"; }
const char *aht_error_string(int code) { return code ? "error" : "ok"; }

#define SOLVER_CALL(s) \
  device_call();       \
  if (!(s)) return AHT_EINVAL; \
  Guard g_((s)->busy)

int aht_set_columns(aht_solver *s, const uint32_t *, const uint32_t *, const float *, const float *) { SOLVER_CALL(s); return AHT_OK; }
int aht_set_params(aht_solver *s, const aht_params *, const float *) { SOLVER_CALL(s); return AHT_OK; }
int aht_init_flow(aht_solver *s) { SOLVER_CALL(s); return AHT_OK; }
int aht_reset_heat(aht_solver *s) { SOLVER_CALL(s); return AHT_OK; }
int aht_reinit_cells(aht_solver *s, const uint32_t *) { SOLVER_CALL(s); return AHT_OK; }
int aht_step(aht_solver *s, int32_t n, uint32_t) { SOLVER_CALL(s); work(n); s->steps += n; return AHT_OK; }
int aht_sync(aht_solver *s) { SOLVER_CALL(s); return AHT_OK; }
int64_t aht_steps_done(const aht_solver *s) { return s ? s->steps : 0; }
int aht_benchmark(aht_solver *s, int32_t n, double *ms) { SOLVER_CALL(s); work(n); s->steps += n; *ms = n; return AHT_OK; }
int aht_read_surface(aht_solver *s, float, float *, float *, float *) { SOLVER_CALL(s); return AHT_OK; }
int aht_read_volume(aht_solver *s, int32_t, float, uint8_t *, float *tm) { SOLVER_CALL(s); *tm = 0; return AHT_OK; }
int aht_stats(aht_solver *s, aht_diag *o) { SOLVER_CALL(s); std::memset(o, 0, sizeof *o); return AHT_OK; }
int aht_stats_noalloc(aht_solver *s, aht_diag *o) { return aht_stats(s, o); }
int aht_read_out(aht_solver *s, float *) { SOLVER_CALL(s); return AHT_OK; }
int aht_mean_reset(aht_solver *s) { SOLVER_CALL(s); return AHT_OK; }
int aht_read_mean(aht_solver *s, double *, aht_budget *b) { SOLVER_CALL(s); if (b) std::memset(b, 0, sizeof *b); return AHT_OK; }
int aht_read_face_mean(aht_solver *s, int32_t, double *) { SOLVER_CALL(s); return AHT_OK; }
int aht_read_state(aht_solver *s, float *, float *) { SOLVER_CALL(s); return AHT_OK; }
int aht_write_state(aht_solver *s, const float *, const float *) { SOLVER_CALL(s); return AHT_OK; }
int aht_read_state_f64(aht_solver *s, double *, double *) { SOLVER_CALL(s); return AHT_OK; }
int aht_write_state_f64(aht_solver *s, const double *, const double *) { SOLVER_CALL(s); return AHT_OK; }
int aht_autotune(aht_solver *s, int32_t, aht_launch *best) { SOLVER_CALL(s); if (best) *best = s->launch; return AHT_OK; }
int aht_set_launch(aht_solver *s, const aht_launch *l) { SOLVER_CALL(s); s->launch = *l; return AHT_OK; }
int aht_get_launch(const aht_solver *s, aht_launch *l) { if (!s) return AHT_EINVAL; *l = s->launch; return AHT_OK; }
int64_t aht_footprint_bytes(const aht_dims *d) { return d ? (int64_t)d->nx * d->ny * d->nz * 208 : 0; }

int aht_batch_create(int32_t, int32_t, aht_batch **out) { device_call(); *out = new aht_batch(); return AHT_OK; }
void aht_batch_destroy(aht_batch *b) {
  device_call();
  if (!b) return;
  if (b->busy.running) fail("batch destroyed while a call on it was running");
  b->busy.destroyed = 1;
}
const char *aht_batch_last_error(const aht_batch *) { return "stand-in"; }
#define BATCH_CALL(b) \
  device_call();      \
  if (!(b)) return AHT_EINVAL; \
  Guard g_((b)->busy)
int aht_batch_add(aht_batch *b, aht_solver *s) { BATCH_CALL(b); if (s->busy.destroyed) fail("destroyed solver added"); b->n++; return AHT_OK; }
int aht_batch_remove(aht_batch *b, aht_solver *) { BATCH_CALL(b); b->n--; return AHT_OK; }
int aht_batch_size(const aht_batch *b) { return b ? b->n : 0; }
int aht_batch_step(aht_batch *b, int32_t n, const int32_t *, const int32_t *, double *ms) {
  BATCH_CALL(b);
  work(n);
  if (ms) *ms = n;
  return AHT_OK;
}
int aht_batch_sync(aht_batch *b) { BATCH_CALL(b); return AHT_OK; }
int aht_batch_set_grid(aht_batch *b, int32_t blocks, int32_t) { BATCH_CALL(b); b->grid = blocks; return AHT_OK; }
int aht_batch_set_block(aht_batch *b, int32_t block) { BATCH_CALL(b); b->block = block; return AHT_OK; }
int aht_batch_grid(const aht_batch *b) { return b ? b->grid : 0; }
int aht_batch_block(const aht_batch *b) { return b ? b->block : 0; }

}  // extern "C"
