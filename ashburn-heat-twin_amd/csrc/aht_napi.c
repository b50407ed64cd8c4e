/*
 * N-API addon: the thin boundary between the Node host (js/solver.js) and the C-ABI of libaht.so (include/aht.h).
 * One JS function per aht_* entry point, typed arrays passed by pointer, no copies beyond what the C-ABI does.
 * step(), benchmark() and batchStep() run aht_step + aht_sync (or the batch's) on the addon's device thread and return
 * a Promise (the reference's step(n) resolves on queue completion, js/live/solver.js:401-406). Errors throw with
 * aht_last_error's message.
 *
 * The device thread: one thread per process takes the asynchronous jobs in the order they were issued (one queue, as
 * the reference's device queue), so two jobs never run the C-ABI on one solver at once, and a Promise settles on the
 * main thread through a thread-safe function. It is not libuv's thread pool on purpose: libuv joins its pool threads
 * in a library destructor at process exit, after the static destructors of the libraries the addon loaded (the HIP
 * runtime among them) have run, so a pool thread that made HIP calls would tear its per-thread HIP state down against
 * a finished runtime. The device thread is joined in the environment's cleanup hook instead, before those
 * destructors; after process.exit() it is never joined and keeps its state until the process ends.
 *
 * The solver handle is a JS external. The C-ABI takes one caller per solver at a time: the synchronous calls on a
 * solver (or batch) throw while an asynchronous job that uses it is queued or running; steps queue behind each other.
 *
 * Lifetime: a step, benchmark or batch step runs on the device thread while the main thread goes on. destroy() of a
 * solver (or batch) with such work in flight does not free it under the job: the handle is marked dead at once
 * (every later call on it throws) and the library object is destroyed by the completion callback of the last job
 * that uses it, on the main thread. A solver that is a member of a batch counts the batch's in-flight steps as its
 * own. Batch membership cannot change while a batch step is in flight.
 */
#define NAPI_VERSION 4
#include <node_api.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "aht.h"

#define CHECK_NAPI(call)                                                         \
  do {                                                                           \
    if ((call) != napi_ok) {                                                     \
      napi_throw_error(env, NULL, "aht addon: N-API call failed: " #call);       \
      return NULL;                                                               \
    }                                                                            \
  } while (0)

typedef struct bhandle bhandle_t;

typedef struct handle {
  aht_solver *s;
  int inflight;        /* step / benchmark jobs queued or running on the device thread            */
  int dead;            /* destroy() was called; the solver goes when no job uses it any more       */
  int released;        /* the JS external was finalized (the struct is freed once nothing uses it) */
  bhandle_t *batch;    /* the batch this solver belongs to (NULL: none)                            */
} handle_t;

struct bhandle {
  aht_batch *b;
  int inflight, dead, released;
  handle_t **mem;      /* member handles, in add order */
  int nmem, cap;
};

static int solver_busy(const handle_t *h) { return h->inflight > 0 || (h->batch && h->batch->inflight > 0); }

static void batch_forget(bhandle_t *b, handle_t *h) {
  for (int k = 0; k < b->nmem; k++)
    if (b->mem[k] == h) {
      memmove(b->mem + k, b->mem + k + 1, (size_t)(b->nmem - k - 1) * sizeof *b->mem);
      b->nmem--;
      break;
    }
  h->batch = NULL;
}

/* the library solver goes now (nothing uses it); frees the struct too once JS has let go of it */
static void solver_release(handle_t *h) {
  if (h->s) {
    if (h->batch) batch_forget(h->batch, h);   /* aht_destroy leaves the library batch itself */
    aht_destroy(h->s);
    h->s = NULL;
  }
  if (h->released) free(h);
}

static void batch_release(bhandle_t *b) {
  if (b->b) {
    for (int k = 0; k < b->nmem; k++) b->mem[k]->batch = NULL;
    b->nmem = 0;
    aht_batch_destroy(b->b);
    b->b = NULL;
  }
  if (b->released) {
    free(b->mem);
    free(b);
  }
}

/* after a job finished: release what was destroyed while it ran */
static void solver_settle(handle_t *h) {
  if ((h->dead || h->released) && !solver_busy(h)) solver_release(h);
}

static void finalize_handle(napi_env env, void *data, void *hint) {
  (void)env; (void)hint;
  handle_t *h = (handle_t *)data;
  h->released = 1;
  if (!solver_busy(h)) solver_release(h);   /* else the last job's completion releases it */
}

static napi_value throw_rc(napi_env env, aht_solver *s, int rc) {
  char buf[600];
  snprintf(buf, sizeof buf, "aht: %s (%s)", s ? aht_last_error(s) : "", aht_error_string(rc));
  napi_throw_error(env, NULL, buf);
  return NULL;
}

static napi_value undef(napi_env env) {
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}

static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value *argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok) return 0;
  if (argc < want) {
    napi_throw_type_error(env, NULL, "aht addon: too few arguments");
    return 0;
  }
  return 1;
}

static handle_t *get_handle(napi_env env, napi_value v) {
  void *p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p || !((handle_t *)p)->s || ((handle_t *)p)->dead) {
    napi_throw_error(env, NULL, "aht addon: invalid or destroyed solver handle");
    return NULL;
  }
  return (handle_t *)p;
}

/* the handle of a solver that no asynchronous job uses (its own, or its batch's): the synchronous calls run the C-ABI
 * on the main thread, and the C-ABI takes one caller per solver at a time */
static handle_t *get_idle_handle(napi_env env, napi_value v) {
  handle_t *h = get_handle(env, v);
  if (h && solver_busy(h)) {
    napi_throw_error(env, NULL, "aht addon: the solver has a step in flight (await it first)");
    return NULL;
  }
  return h;
}

/* pointer and byte length of a typed array; NULL for null/undefined */
static int get_buf(napi_env env, napi_value v, void **ptr, size_t *bytes) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_null || t == napi_undefined) { *ptr = NULL; *bytes = 0; return 1; }
  bool is_ta = false;
  napi_is_typedarray(env, v, &is_ta);
  if (!is_ta) { napi_throw_type_error(env, NULL, "aht addon: expected a typed array"); return 0; }
  napi_typedarray_type ty;
  size_t len, off;
  napi_value ab;
  void *data;
  if (napi_get_typedarray_info(env, v, &ty, &len, &data, &ab, &off) != napi_ok) return 0;
  size_t el = (ty == napi_int8_array || ty == napi_uint8_array || ty == napi_uint8_clamped_array) ? 1
            : (ty == napi_int16_array || ty == napi_uint16_array) ? 2
            : (ty == napi_float64_array || ty == napi_bigint64_array || ty == napi_biguint64_array) ? 8 : 4;
  *ptr = data;
  *bytes = len * el;
  return 1;
}

static int get_i32(napi_env env, napi_value v, int32_t *out) {
  if (napi_get_value_int32(env, v, out) != napi_ok) { napi_throw_type_error(env, NULL, "aht addon: expected a number"); return 0; }
  return 1;
}

static int get_f64(napi_env env, napi_value v, double *out) {
  if (napi_get_value_double(env, v, out) != napi_ok) { napi_throw_type_error(env, NULL, "aht addon: expected a number"); return 0; }
  return 1;
}

static int get_prop_i32(napi_env env, napi_value o, const char *k, int32_t *out, int32_t dflt) {
  bool has = false;
  napi_has_named_property(env, o, k, &has);
  if (!has) { *out = dflt; return 1; }
  napi_value v;
  napi_get_named_property(env, o, k, &v);
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_undefined || t == napi_null) { *out = dflt; return 1; }
  return get_i32(env, v, out);
}

static int get_prop_f64(napi_env env, napi_value o, const char *k, double *out) {
  napi_value v;
  if (napi_get_named_property(env, o, k, &v) != napi_ok) return 0;
  return get_f64(env, v, out);
}

static void set_num(napi_env env, napi_value o, const char *k, double x) {
  napi_value v;
  napi_create_double(env, x, &v);
  napi_set_named_property(env, o, k, v);
}

/* create({nx, ny, nz, storage, device, block}) -> handle */
static napi_value js_create(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  aht_dims d;
  memset(&d, 0, sizeof d);
  if (!get_prop_i32(env, argv[0], "nx", &d.nx, 0) || !get_prop_i32(env, argv[0], "ny", &d.ny, 0) ||
      !get_prop_i32(env, argv[0], "nz", &d.nz, 0) || !get_prop_i32(env, argv[0], "storage", &d.storage, 0) ||
      !get_prop_i32(env, argv[0], "device", &d.device, 0) || !get_prop_i32(env, argv[0], "block", &d.block, 0) ||
      !get_prop_i32(env, argv[0], "layout", &d.layout, 0))
    return NULL;
  aht_solver *s = NULL;
  int rc = aht_create(&d, &s);
  if (rc != AHT_OK) {
    napi_value r = throw_rc(env, s, rc);
    if (s) aht_destroy(s);
    return r;
  }
  handle_t *h = (handle_t *)calloc(1, sizeof *h);
  if (!h) {
    aht_destroy(s);
    napi_throw_error(env, NULL, "aht: out of host memory");
    return NULL;
  }
  h->s = s;
  napi_value ext;
  CHECK_NAPI(napi_create_external(env, h, finalize_handle, NULL, &ext));
  return ext;
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  void *p = NULL;
  napi_get_value_external(env, argv[0], &p);
  handle_t *h = (handle_t *)p;
  if (h && h->s && !h->dead) {
    h->dead = 1;                                /* later calls throw from here on */
    if (!solver_busy(h)) solver_release(h);     /* else the last job using it releases it */
  }
  return undef(env);
}

/* setColumns(h, kt u32, ktop u32, cd f32, q f32) */
static napi_value js_set_columns(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  if (!get_args(env, info, 5, argv)) return NULL;
  handle_t *h = get_idle_handle(env, argv[0]);
  if (!h) return NULL;
  void *p[4];
  size_t b[4];
  for (int k = 0; k < 4; k++) if (!get_buf(env, argv[1 + k], &p[k], &b[k])) return NULL;
  int rc = aht_set_columns(h->s, (const uint32_t *)p[0], (const uint32_t *)p[1], (const float *)p[2], (const float *)p[3]);
  return rc ? throw_rc(env, h->s, rc) : undef(env);
}

/* setParams(h, {tau0..sigmax, inflow_w.., nsp, nsw, oclip}, prof f32[nz*4]) */
static napi_value js_set_params(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  handle_t *h = get_idle_handle(env, argv[0]);
  if (!h) return NULL;
  aht_params P;
  memset(&P, 0, sizeof P);
  static const char *fk[8] = {"tau0", "cs2smag", "gbeta", "prt", "kfloor", "momfac", "srcscale", "sigmax"};
  static const char *ik[7] = {"inflow_w", "inflow_e", "inflow_s", "inflow_n", "nsp", "nsw", "oclip"};
  float *pf = &P.tau0;
  int32_t *pi = &P.inflow_w;
  for (int k = 0; k < 8; k++) { double x; if (!get_prop_f64(env, argv[1], fk[k], &x)) return NULL; pf[k] = (float)x; }
  for (int k = 0; k < 7; k++) if (!get_prop_i32(env, argv[1], ik[k], &pi[k], 0)) return NULL;
  void *pr;
  size_t b;
  if (!get_buf(env, argv[2], &pr, &b)) return NULL;
  int rc = aht_set_params(h->s, &P, (const float *)pr);
  return rc ? throw_rc(env, h->s, rc) : undef(env);
}

#define SIMPLE(name, call)                                               \
  static napi_value name(napi_env env, napi_callback_info info) {        \
    napi_value argv[1];                                                  \
    if (!get_args(env, info, 1, argv)) return NULL;                      \
    handle_t *h = get_idle_handle(env, argv[0]);                              \
    if (!h) return NULL;                                                 \
    int rc = call(h->s);                                                 \
    return rc ? throw_rc(env, h->s, rc) : undef(env);                    \
  }
SIMPLE(js_init_flow, aht_init_flow)
SIMPLE(js_reset_heat, aht_reset_heat)
SIMPLE(js_mean_reset, aht_mean_reset)
SIMPLE(js_sync, aht_sync)

static napi_value js_reinit_cells(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  handle_t *h = get_idle_handle(env, argv[0]);
  if (!h) return NULL;
  void *p;
  size_t b;
  if (!get_buf(env, argv[1], &p, &b)) return NULL;
  int rc = aht_reinit_cells(h->s, (const uint32_t *)p);
  return rc ? throw_rc(env, h->s, rc) : undef(env);
}

/* ---- asynchronous jobs on the device thread ------------------------------------------------------------------ */
typedef struct job job_t;
struct job {
  job_t *next;
  napi_threadsafe_function done;               /* settles the job on the main thread                  */
  napi_deferred deferred;
  napi_ref keep;                               /* keeps the handle's external alive while the job runs */
  void (*run)(job_t *);                        /* on the device thread                                 */
  void (*complete)(napi_env, job_t *);         /* on the main thread; env NULL: environment torn down  */
  int rc;
  double ms;
  char err[512];
};

static struct {
  pthread_mutex_t mu;
  pthread_cond_t cv;
  job_t *head, *tail;
  int started, stop;
  pthread_t th;
} dq = {PTHREAD_MUTEX_INITIALIZER, PTHREAD_COND_INITIALIZER, NULL, NULL, 0, 0};

static void *device_loop(void *arg) {
  (void)arg;
  for (;;) {
    pthread_mutex_lock(&dq.mu);
    while (!dq.head && !dq.stop) pthread_cond_wait(&dq.cv, &dq.mu);
    job_t *j = dq.head;
    if (!j) {                                  /* stop requested and nothing queued */
      pthread_mutex_unlock(&dq.mu);
      return NULL;
    }
    dq.head = j->next;
    if (!dq.head) dq.tail = NULL;
    pthread_mutex_unlock(&dq.mu);
    j->run(j);
    const napi_threadsafe_function tf = j->done;   /* j may be freed on the main thread once the call is queued */
    if (napi_call_threadsafe_function(tf, j, napi_tsfn_blocking) != napi_ok) j->complete(NULL, j);
    napi_release_threadsafe_function(tf, napi_tsfn_release);
  }
}

/* the environment's cleanup hook: runs before the static destructors of the loaded libraries at a normal exit */
static void device_thread_stop(void *arg) {
  (void)arg;
  pthread_mutex_lock(&dq.mu);
  const int started = dq.started;
  dq.stop = 1;
  pthread_cond_signal(&dq.cv);
  pthread_mutex_unlock(&dq.mu);
  if (started) pthread_join(dq.th, NULL);
  pthread_mutex_lock(&dq.mu);
  dq.started = 0;
  dq.stop = 0;
  pthread_mutex_unlock(&dq.mu);
}

static void job_settle(napi_env env, napi_value cb, void *ctx, void *data) {
  (void)cb; (void)ctx;
  job_t *j = (job_t *)data;
  j->complete(env, j);
}

/* resolves (value) or rejects (j->err) the job's Promise and drops its handle reference */
static void job_finish(napi_env env, job_t *j, int resolve_ms) {
  if (j->rc) {
    napi_value msg, e;
    napi_create_string_utf8(env, j->err, NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, NULL, msg, &e);
    napi_reject_deferred(env, j->deferred, e);
  } else {
    napi_value v;
    if (resolve_ms) napi_create_double(env, j->ms, &v);
    else napi_get_undefined(env, &v);
    napi_resolve_deferred(env, j->deferred, v);
  }
  napi_delete_reference(env, j->keep);
}

/* queues j on the device thread (started on first use); returns the job's Promise, NULL with an exception pending */
static napi_value queue_job(napi_env env, job_t *j, napi_value keep) {
  napi_value promise, name;
  CHECK_NAPI(napi_create_string_utf8(env, "aht_device_job", NAPI_AUTO_LENGTH, &name));
  CHECK_NAPI(napi_create_threadsafe_function(env, NULL, NULL, name, 0, 1, NULL, NULL, NULL, job_settle, &j->done));
  const char *fail = NULL;
  if (napi_create_reference(env, keep, 1, &j->keep) != napi_ok) fail = "aht addon: N-API call failed: napi_create_reference";
  else if (napi_create_promise(env, &j->deferred, &promise) != napi_ok) {
    napi_delete_reference(env, j->keep);
    fail = "aht addon: N-API call failed: napi_create_promise";
  }
  if (fail) {
    napi_release_threadsafe_function(j->done, napi_tsfn_abort);
    napi_throw_error(env, NULL, fail);
    return NULL;
  }
  pthread_mutex_lock(&dq.mu);
  if (!dq.started) {
    if (pthread_create(&dq.th, NULL, device_loop, NULL) != 0) {
      pthread_mutex_unlock(&dq.mu);
      napi_release_threadsafe_function(j->done, napi_tsfn_abort);
      napi_delete_reference(env, j->keep);
      napi_value err, msg;           /* the Promise exists: reject it rather than leave it pending */
      napi_create_string_utf8(env, "aht addon: cannot start the device thread", NAPI_AUTO_LENGTH, &msg);
      napi_create_error(env, NULL, msg, &err);
      napi_reject_deferred(env, j->deferred, err);
      napi_throw(env, err);
      return NULL;
    }
    dq.started = 1;
    napi_add_env_cleanup_hook(env, device_thread_stop, NULL);
  }
  j->next = NULL;
  if (dq.tail) dq.tail->next = j;
  else dq.head = j;
  dq.tail = j;
  pthread_cond_signal(&dq.cv);
  pthread_mutex_unlock(&dq.mu);
  return promise;
}

typedef struct {
  job_t base;
  handle_t *h;
  int32_t n;
  uint32_t flags;
  int bench;
} step_job;

static void step_run(job_t *jb) {
  step_job *j = (step_job *)jb;
  if (j->bench) {
    jb->rc = aht_benchmark(j->h->s, j->n, &jb->ms);
  } else {
    jb->rc = aht_step(j->h->s, j->n, j->flags);
    if (jb->rc == AHT_OK) jb->rc = aht_sync(j->h->s);
  }
  if (jb->rc) snprintf(jb->err, sizeof jb->err, "aht: %s (%s)", aht_last_error(j->h->s), aht_error_string(jb->rc));
}

static void step_complete(napi_env env, job_t *jb) {
  step_job *j = (step_job *)jb;
  if (env) job_finish(env, jb, j->bench);
  handle_t *h = j->h;
  free(j);
  h->inflight--;
  if (env) solver_settle(h);
}

static napi_value start_job(napi_env env, napi_value hv, handle_t *h, int32_t n, uint32_t flags, int bench) {
  step_job *j = (step_job *)calloc(1, sizeof *j);
  if (!j) {
    napi_throw_error(env, NULL, "aht: out of host memory");
    return NULL;
  }
  j->base.run = step_run;
  j->base.complete = step_complete;
  j->h = h; j->n = n; j->flags = flags; j->bench = bench;
  h->inflight++;
  napi_value promise = queue_job(env, &j->base, hv);
  if (!promise) {
    h->inflight--;
    free(j);
  }
  return promise;
}

/* step(h, n, flags) -> Promise<void> */
static napi_value js_step(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  handle_t *h = get_handle(env, argv[0]);
  if (!h) return NULL;
  int32_t n, fl;
  if (!get_i32(env, argv[1], &n) || !get_i32(env, argv[2], &fl)) return NULL;
  if (h->batch && h->batch->inflight) {
    napi_throw_error(env, NULL, "aht: the solver's batch has a step in flight");
    return NULL;
  }
  return start_job(env, argv[0], h, n, (uint32_t)fl, 0);
}

/* benchmark(h, n) -> Promise<ms> */
static napi_value js_benchmark(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  handle_t *h = get_handle(env, argv[0]);
  if (!h) return NULL;
  int32_t n;
  if (!get_i32(env, argv[1], &n)) return NULL;
  if (h->batch && h->batch->inflight) {
    napi_throw_error(env, NULL, "aht: the solver's batch has a step in flight");
    return NULL;
  }
  return start_job(env, argv[0], h, n, 0, 1);
}

/* readSurface(h, conv, th f32, u f32|null, v f32|null) */
static napi_value js_read_surface(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  if (!get_args(env, info, 5, argv)) return NULL;
  handle_t *h = get_idle_handle(env, argv[0]);
  if (!h) return NULL;
  double conv;
  void *p[3];
  size_t b[3];
  if (!get_f64(env, argv[1], &conv)) return NULL;
  for (int k = 0; k < 3; k++) if (!get_buf(env, argv[2 + k], &p[k], &b[k])) return NULL;
  int rc = aht_read_surface(h->s, (float)conv, (float *)p[0], (float *)p[1], (float *)p[2]);
  return rc ? throw_rc(env, h->s, rc) : undef(env);
}

/* readVolume(h, nzv, volmax, vol u8) -> thetaMax */
static napi_value js_read_volume(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return NULL;
  handle_t *h = get_idle_handle(env, argv[0]);
  if (!h) return NULL;
  int32_t nzv;
  double vm;
  void *p;
  size_t b;
  if (!get_i32(env, argv[1], &nzv) || !get_f64(env, argv[2], &vm) || !get_buf(env, argv[3], &p, &b)) return NULL;
  float tmax = 0.0f;
  int rc = aht_read_volume(h->s, nzv, (float)vm, (uint8_t *)p, &tmax);
  if (rc) return throw_rc(env, h->s, rc);
  napi_value r;
  napi_create_double(env, tmax, &r);
  return r;
}

/* stats(h) -> {sum_theta, ...} */
static napi_value js_stats(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  handle_t *h = get_idle_handle(env, argv[0]);
  if (!h) return NULL;
  aht_diag d;
  int rc = aht_stats(h->s, &d);
  if (rc) return throw_rc(env, h->s, rc);
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "sum_theta", d.sum_theta);
  set_num(env, o, "sum_ke", d.sum_ke);
  set_num(env, o, "sum_drho", d.sum_drho);
  set_num(env, o, "n_fluid", d.n_fluid);
  set_num(env, o, "theta_max", d.theta_max);
  set_num(env, o, "umax", d.umax);
  set_num(env, o, "non_finite", (double)d.non_finite);
  return o;
}

/* readMean(h, surf3 f64[3*NXY] | null) -> budget object */
static napi_value js_read_mean(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  handle_t *h = get_idle_handle(env, argv[0]);
  if (!h) return NULL;
  void *p;
  size_t b;
  if (!get_buf(env, argv[1], &p, &b)) return NULL;
  aht_budget bg;
  int rc = aht_read_mean(h->s, (double *)p, &bg);
  if (rc) return throw_rc(env, h->s, rc);
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "n_acc", (double)bg.n_acc);
  set_num(env, o, "heat_in", bg.heat_in);
  set_num(env, o, "heat_out", bg.heat_out);
  set_num(env, o, "content", bg.content);
  set_num(env, o, "q_w", bg.q_w);
  set_num(env, o, "q_e", bg.q_e);
  set_num(env, o, "q_s", bg.q_s);
  set_num(env, o, "q_n", bg.q_n);
  set_num(env, o, "mass_imbalance", bg.mass_imbalance);
  return o;
}

/* readState(h, f f32|null, g f32|null), writeState(h, f|null, g|null) */
static napi_value js_state(napi_env env, napi_callback_info info, int write) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  handle_t *h = get_idle_handle(env, argv[0]);
  if (!h) return NULL;
  void *f, *g;
  size_t bf, bg;
  if (!get_buf(env, argv[1], &f, &bf) || !get_buf(env, argv[2], &g, &bg)) return NULL;
  int rc = write ? aht_write_state(h->s, (const float *)f, (const float *)g) : aht_read_state(h->s, (float *)f, (float *)g);
  return rc ? throw_rc(env, h->s, rc) : undef(env);
}
static napi_value js_read_state(napi_env env, napi_callback_info info) { return js_state(env, info, 0); }
static napi_value js_write_state(napi_env env, napi_callback_info info) { return js_state(env, info, 1); }

static napi_value js_read_out(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  handle_t *h = get_idle_handle(env, argv[0]);
  if (!h) return NULL;
  void *p;
  size_t b;
  if (!get_buf(env, argv[1], &p, &b)) return NULL;
  int rc = aht_read_out(h->s, (float *)p);
  return rc ? throw_rc(env, h->s, rc) : undef(env);
}

/* footprint({nx, ny, nz, storage}) -> bytes */
static napi_value js_footprint(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  aht_dims d;
  memset(&d, 0, sizeof d);
  if (!get_prop_i32(env, argv[0], "nx", &d.nx, 0) || !get_prop_i32(env, argv[0], "ny", &d.ny, 0) ||
      !get_prop_i32(env, argv[0], "nz", &d.nz, 0) || !get_prop_i32(env, argv[0], "storage", &d.storage, 0) ||
      !get_prop_i32(env, argv[0], "layout", &d.layout, 0))
    return NULL;
  napi_value r;
  napi_create_double(env, (double)aht_footprint_bytes(&d), &r);
  return r;
}

static napi_value launch_obj(napi_env env, const aht_launch *l) {
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "block", l->block);
  set_num(env, o, "levels", l->levels);
  set_num(env, o, "window", l->window);
  set_num(env, o, "sweep", l->sweep);
  set_num(env, o, "waves", l->waves);
  return o;
}

/* autotune(h, steps) -> {block, levels, window, sweep}: times every launch-shape candidate for `steps` steps, keeps the
 * fastest, leaves the state and step counters as they were (blocks the calling thread for the tuning steps) */
static napi_value js_autotune(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  handle_t *h = get_idle_handle(env, argv[0]);
  int32_t steps;
  if (!h || !get_i32(env, argv[1], &steps)) return NULL;
  aht_launch best;
  const int rc = aht_autotune(h->s, steps, &best);
  if (rc) return throw_rc(env, h->s, rc);
  return launch_obj(env, &best);
}

/* setLaunch(h, {block, levels, window, sweep}); getLaunch(h) -> {block, levels, window, sweep} */
static napi_value js_set_launch(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  handle_t *h = get_idle_handle(env, argv[0]);
  if (!h) return NULL;
  aht_launch l;
  memset(&l, 0, sizeof l);
  if (!get_prop_i32(env, argv[1], "block", &l.block, 256) || !get_prop_i32(env, argv[1], "levels", &l.levels, 1) ||
      !get_prop_i32(env, argv[1], "window", &l.window, 1) || !get_prop_i32(env, argv[1], "sweep", &l.sweep, 0) ||
      !get_prop_i32(env, argv[1], "waves", &l.waves, 0))
    return NULL;
  const int rc = aht_set_launch(h->s, &l);
  if (rc) return throw_rc(env, h->s, rc);
  return undef(env);
}

static napi_value js_get_launch(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  handle_t *h = get_handle(env, argv[0]);
  if (!h) return NULL;
  aht_launch l;
  const int rc = aht_get_launch(h->s, &l);
  if (rc) return throw_rc(env, h->s, rc);
  return launch_obj(env, &l);
}

static napi_value js_steps_done(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  handle_t *h = get_handle(env, argv[0]);
  if (!h) return NULL;
  napi_value r;
  napi_create_double(env, (double)aht_steps_done(h->s), &r);
  return r;
}

/* ---- batches -------------------------------------------------------------------------------------------------- */
static void finalize_batch(napi_env env, void *data, void *hint) {
  (void)env; (void)hint;
  bhandle_t *b = (bhandle_t *)data;
  b->released = 1;
  if (!b->inflight) batch_release(b);
}

static bhandle_t *get_bhandle(napi_env env, napi_value v) {
  void *p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p || !((bhandle_t *)p)->b || ((bhandle_t *)p)->dead) {
    napi_throw_error(env, NULL, "aht addon: invalid or destroyed batch handle");
    return NULL;
  }
  return (bhandle_t *)p;
}

static napi_value throw_brc(napi_env env, aht_batch *b, int rc) {
  char buf[600];
  snprintf(buf, sizeof buf, "aht batch: %s (%s)", b ? aht_batch_last_error(b) : "", aht_error_string(rc));
  napi_throw_error(env, NULL, buf);
  return NULL;
}

/* batchCreate(device, storage) -> handle */
static napi_value js_batch_create(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  int32_t dev, stg;
  if (!get_i32(env, argv[0], &dev) || !get_i32(env, argv[1], &stg)) return NULL;
  aht_batch *b = NULL;
  int rc = aht_batch_create(dev, stg, &b);
  if (rc) {
    napi_value r = throw_brc(env, b, rc);
    if (b) aht_batch_destroy(b);
    return r;
  }
  bhandle_t *h = (bhandle_t *)calloc(1, sizeof *h);
  if (!h) {
    aht_batch_destroy(b);
    napi_throw_error(env, NULL, "aht batch: out of host memory");
    return NULL;
  }
  h->b = b;
  napi_value ext;
  CHECK_NAPI(napi_create_external(env, h, finalize_batch, NULL, &ext));
  return ext;
}

static napi_value js_batch_destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  void *p = NULL;
  napi_get_value_external(env, argv[0], &p);
  bhandle_t *b = (bhandle_t *)p;
  if (b && b->b && !b->dead) {
    b->dead = 1;
    if (!b->inflight) batch_release(b);
  }
  return undef(env);
}

/* batchAdd(b, solver), batchRemove(b, solver) */
static napi_value js_batch_member(napi_env env, napi_callback_info info, int add) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  bhandle_t *b = get_bhandle(env, argv[0]);
  if (!b) return NULL;
  handle_t *h = get_handle(env, argv[1]);
  if (!h) return NULL;
  if (b->inflight || h->inflight) {
    napi_throw_error(env, NULL, "aht batch: membership cannot change while a step of the batch or the solver is in flight");
    return NULL;
  }
  int rc = add ? aht_batch_add(b->b, h->s) : aht_batch_remove(b->b, h->s);
  if (rc) return throw_brc(env, b->b, rc);
  if (add) {
    if (b->nmem == b->cap) {
      const int cap = b->cap ? 2 * b->cap : 16;
      handle_t **m = (handle_t **)realloc(b->mem, (size_t)cap * sizeof *m);
      if (!m) {
        (void)aht_batch_remove(b->b, h->s);
        napi_throw_error(env, NULL, "aht batch: out of host memory");
        return NULL;
      }
      b->mem = m;
      b->cap = cap;
    }
    b->mem[b->nmem++] = h;
    h->batch = b;
  } else {
    batch_forget(b, h);
  }
  return undef(env);
}
static napi_value js_batch_add(napi_env env, napi_callback_info info) { return js_batch_member(env, info, 1); }
static napi_value js_batch_remove(napi_env env, napi_callback_info info) { return js_batch_member(env, info, 0); }

/* batchSetGrid(b, blocks, useGraph), batchGrid(b) -> blocks */
static napi_value js_batch_set_grid(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  bhandle_t *b = get_bhandle(env, argv[0]);
  if (!b) return NULL;
  if (b->inflight) {
    napi_throw_error(env, NULL, "aht batch: the batch has a step in flight (await it first)");
    return NULL;
  }
  int32_t g, ug;
  if (!get_i32(env, argv[1], &g) || !get_i32(env, argv[2], &ug)) return NULL;
  int rc = aht_batch_set_grid(b->b, g, ug);
  return rc ? throw_brc(env, b->b, rc) : undef(env);
}
static napi_value js_batch_grid(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  bhandle_t *b = get_bhandle(env, argv[0]);
  if (!b) return NULL;
  napi_value r;
  napi_create_int32(env, aht_batch_grid(b->b), &r);
  return r;
}

/* batchSetBlock(b, threads: 0 | 64 | 128 | 256), batchBlock(b) -> threads (0 while tuning) */
static napi_value js_batch_set_block(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  bhandle_t *b = get_bhandle(env, argv[0]);
  if (!b) return NULL;
  if (b->inflight) {
    napi_throw_error(env, NULL, "aht batch: the batch has a step in flight (await it first)");
    return NULL;
  }
  int32_t bl;
  if (!get_i32(env, argv[1], &bl)) return NULL;
  int rc = aht_batch_set_block(b->b, bl);
  return rc ? throw_brc(env, b->b, rc) : undef(env);
}
static napi_value js_batch_block(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  bhandle_t *b = get_bhandle(env, argv[0]);
  if (!b) return NULL;
  napi_value r;
  napi_create_int32(env, aht_batch_block(b->b), &r);
  return r;
}

typedef struct {
  job_t base;
  bhandle_t *b;
  int32_t n;
  int32_t *ns, *af;
} bstep_job;

static void bstep_run(job_t *jb) {
  bstep_job *j = (bstep_job *)jb;
  jb->rc = aht_batch_step(j->b->b, j->n, j->ns, j->af, &jb->ms);
  if (jb->rc == AHT_OK) jb->rc = aht_batch_sync(j->b->b);
  if (jb->rc)
    snprintf(jb->err, sizeof jb->err, "aht batch: %s (%s)", aht_batch_last_error(j->b->b), aht_error_string(jb->rc));
}

static void bstep_complete(napi_env env, job_t *jb) {
  bstep_job *j = (bstep_job *)jb;
  if (env) job_finish(env, jb, 1);
  bhandle_t *b = j->b;
  free(j->ns);
  free(j->af);
  free(j);
  b->inflight--;
  if (env && !b->inflight) {
    /* members destroyed while the step ran go now (solver_release edits the member list: walk it backwards) */
    for (int k = b->nmem - 1; k >= 0; k--) {
      handle_t *h = b->mem[k];
      if ((h->dead || h->released) && !h->inflight) solver_release(h);
    }
    if (b->dead || b->released) batch_release(b);
  }
}

/* batchStep(b, n, nsteps Int32Array|null, accFrom Int32Array|null) -> Promise<ms> (the arrays are copied) */
static napi_value js_batch_step(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return NULL;
  bhandle_t *b = get_bhandle(env, argv[0]);
  if (!b) return NULL;
  int32_t n;
  if (!get_i32(env, argv[1], &n)) return NULL;
  void *pn, *pa;
  size_t bn, ba;
  if (!get_buf(env, argv[2], &pn, &bn) || !get_buf(env, argv[3], &pa, &ba)) return NULL;
  const size_t m = (size_t)aht_batch_size(b->b);
  for (int k = 0; k < b->nmem; k++)
    if (b->mem[k]->inflight) {
      napi_throw_error(env, NULL, "aht batch: a member solver has its own step in flight");
      return NULL;
    }
  if ((pn && bn < m * 4) || (pa && ba < m * 4)) {
    napi_throw_range_error(env, NULL, "aht batch: nsteps/accFrom need one entry per member");
    return NULL;
  }
  bstep_job *j = (bstep_job *)calloc(1, sizeof *j);
  if (j && pn && !(j->ns = (int32_t *)malloc(m * 4 + 4))) { free(j); j = NULL; }
  if (j && pa && !(j->af = (int32_t *)malloc(m * 4 + 4))) { free(j->ns); free(j); j = NULL; }
  if (!j) {
    napi_throw_error(env, NULL, "aht batch: out of host memory");
    return NULL;
  }
  if (pn) memcpy(j->ns, pn, m * 4);
  if (pa) memcpy(j->af, pa, m * 4);
  j->base.run = bstep_run;
  j->base.complete = bstep_complete;
  j->b = b; j->n = n;
  b->inflight++;
  napi_value promise = queue_job(env, &j->base, argv[0]);
  if (!promise) {
    b->inflight--;
    free(j->ns);
    free(j->af);
    free(j);
  }
  return promise;
}

static napi_value init(napi_env env, napi_value exports) {
  struct { const char *name; napi_callback cb; } fns[] = {
    {"create", js_create}, {"destroy", js_destroy}, {"setColumns", js_set_columns}, {"setParams", js_set_params},
    {"initFlow", js_init_flow}, {"resetHeat", js_reset_heat}, {"reinitCells", js_reinit_cells},
    {"step", js_step}, {"sync", js_sync}, {"benchmark", js_benchmark}, {"readSurface", js_read_surface},
    {"readVolume", js_read_volume}, {"stats", js_stats}, {"meanReset", js_mean_reset}, {"readMean", js_read_mean},
    {"readState", js_read_state}, {"writeState", js_write_state}, {"readOut", js_read_out},
    {"footprint", js_footprint}, {"stepsDone", js_steps_done},
    {"autotune", js_autotune}, {"setLaunch", js_set_launch}, {"getLaunch", js_get_launch},
    {"batchCreate", js_batch_create}, {"batchDestroy", js_batch_destroy}, {"batchAdd", js_batch_add},
    {"batchRemove", js_batch_remove}, {"batchSetGrid", js_batch_set_grid}, {"batchGrid", js_batch_grid},
    {"batchStep", js_batch_step}, {"batchSetBlock", js_batch_set_block}, {"batchBlock", js_batch_block},
  };
  for (size_t k = 0; k < sizeof fns / sizeof fns[0]; k++) {
    napi_value f;
    CHECK_NAPI(napi_create_function(env, fns[k].name, NAPI_AUTO_LENGTH, fns[k].cb, NULL, &f));
    CHECK_NAPI(napi_set_named_property(env, exports, fns[k].name, f));
  }
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
