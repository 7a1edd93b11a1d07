/*
 * shx.node: the N-API addon that binds libshx (include/shx.h) for the reference's TypeScript
 * simulation API (src/sim.ts `Simulation`, src/amr/amr_sim.ts `AmrSim`). A thin layer: every
 * export converts JS values to the C ABI's plain pointers and sizes, calls one shx_* entry, and
 * throws a JS Error carrying shx_last_error() on a failure status (the reference throws Error for
 * refusals, src/sim.ts:2289-2311). Handles are napi externals whose finalizer destroys the sim if
 * JS drops it without destroy(). State arrays are Float32Array (precision 32) or Float64Array
 * (precision 64), copied in and out. readStateAsync / diagnosticsAsync enqueue the readback's
 * copies on the sim's stream when they are called (shx_readback_begin: the state as the steps
 * enqueued so far leave it, into host memory the readback owns), wait for them on a worker
 * thread (napi_create_async_work) and resolve a Promise, as the reference's mapAsync readbacks
 * do (src/sim.ts:3598-3640, 3790-3807). The worker touches only the readback, so the sim may
 * step, re-mesh or be destroyed while one is pending.
 */
#include <node_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/shx.h"

#define NAPI_CALL(env, call)                                                   \
  do {                                                                         \
    if ((call) != napi_ok) {                                                   \
      napi_throw_error((env), NULL, "N-API call failed: " #call);              \
      return NULL;                                                             \
    }                                                                          \
  } while (0)

static napi_value throw_status(napi_env env, int rc) {
  const char* msg = shx_last_error();
  char buf[600];
  snprintf(buf, sizeof buf, "libshx error %d: %s", rc, msg ? msg : "");
  napi_throw_error(env, NULL, buf);
  return NULL;
}
#define SHX_CALL(env, call)                   \
  do {                                        \
    int rc_ = (call);                         \
    if (rc_ != SHX_OK) return throw_status((env), rc_); \
  } while (0)

/* ---------------- handles ---------------- */
typedef struct {
  int amr;        /* 0: shx_sim, 1: shx_amr */
  void* p;
  int precision;
  size_t ncell;
  int nblocks;
  int cells;      /* per block: 4096, or 256 on a 2D block grid */
} Handle;

static void finalize_handle(napi_env env, void* data, void* hint) {
  (void)env; (void)hint;
  Handle* h = (Handle*)data;
  if (h->p) {
    if (h->amr) shx_amr_destroy((shx_amr*)h->p);
    else shx_destroy((shx_sim*)h->p);
  }
  free(h);
}

static Handle* get_handle(napi_env env, napi_value v, int amr) {
  void* data = NULL;
  if (napi_get_value_external(env, v, &data) != napi_ok || !data) {
    napi_throw_type_error(env, NULL, "expected a shx handle");
    return NULL;
  }
  Handle* h = (Handle*)data;
  if (amr >= 0 && h->amr != amr) {
    napi_throw_type_error(env, NULL, amr ? "expected a block-grid handle" : "expected a uniform-grid handle");
    return NULL;
  }
  if (!h->p) {
    napi_throw_error(env, NULL, "the simulation was destroyed");
    return NULL;
  }
  return h;
}

static napi_value wrap_handle(napi_env env, Handle* h) {
  napi_value out;
  if (napi_create_external(env, h, finalize_handle, NULL, &out) != napi_ok) {
    finalize_handle(env, h, NULL);
    napi_throw_error(env, NULL, "napi_create_external failed");
    return NULL;
  }
  return out;
}

/* ---------------- argument helpers ---------------- */
static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok) return 0;
  if (argc < want) {
    napi_throw_type_error(env, NULL, "too few arguments");
    return 0;
  }
  return 1;
}

static double num_prop(napi_env env, napi_value obj, const char* key, double dflt) {
  bool has = false;
  napi_value v;
  double d = dflt;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return dflt;
  if (napi_get_named_property(env, obj, key, &v) != napi_ok) return dflt;
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_boolean) {
    bool b;
    napi_get_value_bool(env, v, &b);
    return b ? 1 : 0;
  }
  if (t != napi_number) return dflt;
  napi_get_value_double(env, v, &d);
  return d;
}

static void int3_prop(napi_env env, napi_value obj, const char* key, int out[3]) {
  bool has = false;
  napi_value v;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return;
  if (napi_get_named_property(env, obj, key, &v) != napi_ok) return;
  for (uint32_t i = 0; i < 3; i++) {
    napi_value e;
    double d;
    if (napi_get_element(env, v, i, &e) == napi_ok && napi_get_value_double(env, e, &d) == napi_ok) out[i] = (int)d;
  }
}

static void dbl3_prop(napi_env env, napi_value obj, const char* key, double out[3]) {
  bool has = false;
  napi_value v;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return;
  if (napi_get_named_property(env, obj, key, &v) != napi_ok) return;
  for (uint32_t i = 0; i < 3; i++) {
    napi_value e;
    double d;
    if (napi_get_element(env, v, i, &e) == napi_ok && napi_get_value_double(env, e, &d) == napi_ok) out[i] = d;
  }
}

/* The shx_config from a plain object whose fields are already resolved by the JS facade
 * (configFor's defaults, src/sim.ts:645-736): n, ndim, bcLo, bcHi, limiter, riemann (codes),
 * isentropic, precision, path, device, dx, gamma, cfl, rhoFloor, pFloor, etaDual, shockThresh,
 * tEnd, fluff, tile. */
static void config_from(napi_env env, napi_value o, shx_config* c) {
  int n[3] = {1, 1, 1};
  int3_prop(env, o, "n", n);
  shx_config_default(c, n[0], n[1], n[2]);
  c->ndim = (int)num_prop(env, o, "ndim", c->ndim);
  int3_prop(env, o, "bcLo", c->bc_lo);
  int3_prop(env, o, "bcHi", c->bc_hi);
  c->limiter = (int)num_prop(env, o, "limiter", c->limiter);
  c->riemann = (int)num_prop(env, o, "riemann", c->riemann);
  c->isentropic = (int)num_prop(env, o, "isentropic", c->isentropic);
  c->precision = (int)num_prop(env, o, "precision", c->precision);
  c->path = (int)num_prop(env, o, "path", c->path);
  c->device = (int)num_prop(env, o, "device", c->device);
  c->dx = num_prop(env, o, "dx", c->dx);
  c->gamma = num_prop(env, o, "gamma", c->gamma);
  c->cfl = num_prop(env, o, "cfl", c->cfl);
  c->rho_floor = num_prop(env, o, "rhoFloor", c->rho_floor);
  c->p_floor = num_prop(env, o, "pFloor", c->p_floor);
  c->eta_dual = num_prop(env, o, "etaDual", c->eta_dual);
  c->shock_thresh = num_prop(env, o, "shockThresh", c->shock_thresh);
  c->t_end = num_prop(env, o, "tEnd", c->t_end);
  dbl3_prop(env, o, "fluff", c->fluff);
  int3_prop(env, o, "tile", c->tile);
  dbl3_prop(env, o, "org", c->org);
}

/* A typed array's data pointer, checked against the element type and length wanted. */
static void* typed_data(napi_env env, napi_value v, int precision, size_t want) {
  bool is = false;
  if (napi_is_typedarray(env, v, &is) != napi_ok || !is) {
    napi_throw_type_error(env, NULL, "expected a Float32Array or Float64Array");
    return NULL;
  }
  napi_typedarray_type t;
  size_t len;
  void* data;
  napi_value ab;
  size_t off;
  napi_get_typedarray_info(env, v, &t, &len, &data, &ab, &off);
  napi_typedarray_type want_t = precision == 64 ? napi_float64_array : napi_float32_array;
  if (t != want_t) {
    napi_throw_type_error(env, NULL, precision == 64 ? "precision 64 state is a Float64Array" : "precision 32 state is a Float32Array");
    return NULL;
  }
  if (len != want) {
    char buf[160];
    snprintf(buf, sizeof buf, "state array holds %zu values, %zu expected", len, want);
    napi_throw_range_error(env, NULL, buf);
    return NULL;
  }
  return data;
}

static napi_value new_typed(napi_env env, int precision, size_t n, void** data) {
  napi_value ab, arr;
  size_t es = precision == 64 ? 8 : 4;
  if (napi_create_arraybuffer(env, n * es, data, &ab) != napi_ok) return NULL;
  if (napi_create_typedarray(env, precision == 64 ? napi_float64_array : napi_float32_array, n, ab, 0, &arr) != napi_ok)
    return NULL;
  return arr;
}

static napi_value f64_array(napi_env env, const double* v, size_t n) {
  void* data;
  napi_value arr = new_typed(env, 64, n, &data);
  if (arr) memcpy(data, v, n * sizeof(double));
  return arr;
}

static napi_value undefined(napi_env env) {
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}

/* ---------------- uniform grid ---------------- */
static napi_value js_version(napi_env env, napi_callback_info info) {
  (void)info;
  napi_value s;
  NAPI_CALL(env, napi_create_string_utf8(env, shx_version(), NAPI_AUTO_LENGTH, &s));
  return s;
}

static napi_value js_create(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  shx_config c;
  config_from(env, argv[0], &c);
  shx_sim* sim = NULL;
  SHX_CALL(env, shx_create(&c, &sim));
  Handle* h = (Handle*)calloc(1, sizeof(Handle));
  h->amr = 0;
  h->p = sim;
  h->precision = c.precision;
  h->ncell = (size_t)c.n[0] * c.n[1] * c.n[2];
  return wrap_handle(env, h);
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  void* data = NULL;
  NAPI_CALL(env, napi_get_value_external(env, argv[0], &data));
  Handle* h = (Handle*)data;
  if (h && h->p) {
    if (h->amr) shx_amr_destroy((shx_amr*)h->p);
    else shx_destroy((shx_sim*)h->p);
    h->p = NULL;
  }
  return undefined(env);
}

static napi_value js_set_state(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  size_t argc = 5;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 3) { napi_throw_type_error(env, NULL, "setState(handle, u1, u2[, time, steps])"); return NULL; }
  Handle* h = get_handle(env, argv[0], 0);
  if (!h) return NULL;
  void* u1 = typed_data(env, argv[1], h->precision, 4 * h->ncell);
  if (!u1) return NULL;
  void* u2 = typed_data(env, argv[2], h->precision, 4 * h->ncell);
  if (!u2) return NULL;
  double time = 0, steps = 0;
  if (argc > 3) napi_get_value_double(env, argv[3], &time);
  if (argc > 4) napi_get_value_double(env, argv[4], &steps);
  if (h->precision == 64) SHX_CALL(env, shx_set_state_f64((shx_sim*)h->p, u1, u2, h->ncell, time, (int)steps));
  else SHX_CALL(env, shx_set_state((shx_sim*)h->p, u1, u2, h->ncell, time, (int)steps));
  return undefined(env);
}

static napi_value state_pair(napi_env env, napi_value a, napi_value b) {
  napi_value o;
  NAPI_CALL(env, napi_create_object(env, &o));
  NAPI_CALL(env, napi_set_named_property(env, o, "u1", a));
  NAPI_CALL(env, napi_set_named_property(env, o, "u2", b));
  return o;
}

static napi_value js_read_state(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Handle* h = get_handle(env, argv[0], 0);
  if (!h) return NULL;
  void *d1, *d2;
  napi_value a = new_typed(env, h->precision, 4 * h->ncell, &d1);
  napi_value b = new_typed(env, h->precision, 4 * h->ncell, &d2);
  if (!a || !b) { napi_throw_error(env, NULL, "out of memory"); return NULL; }
  if (h->precision == 64) SHX_CALL(env, shx_read_state_f64((shx_sim*)h->p, d1, d2));
  else SHX_CALL(env, shx_read_state((shx_sim*)h->p, d1, d2));
  return state_pair(env, a, b);
}

static napi_value js_step(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  Handle* h = get_handle(env, argv[0], 0);
  if (!h) return NULL;
  int32_t count;
  NAPI_CALL(env, napi_get_value_int32(env, argv[1], &count));
  int taken = 0;
  SHX_CALL(env, shx_step((shx_sim*)h->p, count, &taken));
  napi_value out;
  NAPI_CALL(env, napi_create_int32(env, taken, &out));
  return out;
}

/* writeParams / buildHydro (src/sim.ts:2680-2700, 2434-2560; src/amr/amr_sim.ts:1677, :1107): the
 * facade's current config, on either grid. */
static napi_value reconfigure(napi_env env, napi_callback_info info, int scheme) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  Handle* h = get_handle(env, argv[0], -1);
  if (!h) return NULL;
  shx_config c;
  config_from(env, argv[1], &c);
  if (h->amr) {
    if (scheme) SHX_CALL(env, shx_amr_build_hydro((shx_amr*)h->p, &c));
    else SHX_CALL(env, shx_amr_write_params((shx_amr*)h->p, &c));
  } else {
    if (scheme) SHX_CALL(env, shx_build_hydro((shx_sim*)h->p, &c));
    else SHX_CALL(env, shx_write_params((shx_sim*)h->p, &c));
  }
  return undefined(env);
}
static napi_value js_write_params(napi_env env, napi_callback_info info) { return reconfigure(env, info, 0); }
static napi_value js_build_hydro(napi_env env, napi_callback_info info) { return reconfigure(env, info, 1); }

static napi_value js_sync(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Handle* h = get_handle(env, argv[0], 0);
  if (!h) return NULL;
  SHX_CALL(env, shx_sync((shx_sim*)h->p));
  return undefined(env);
}

/* After steps on a block grid that re-meshes by itself (layout regrid_every): the slot count
   of the current tree, which sizes the table and the state. */
static int amr_resync(Handle* h) {
  int st[5];
  int rc = shx_amr_stats((shx_amr*)h->p, st);
  if (rc != SHX_OK) return rc;
  h->nblocks = st[0];
  h->ncell = (size_t)h->nblocks * h->cells;
  return SHX_OK;
}

static napi_value js_run_to_end(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  Handle* h = get_handle(env, argv[0], -1);
  if (!h) return NULL;
  int32_t batch, max_steps;
  NAPI_CALL(env, napi_get_value_int32(env, argv[1], &batch));
  NAPI_CALL(env, napi_get_value_int32(env, argv[2], &max_steps));
  int steps = 0;
  if (h->amr) {
    SHX_CALL(env, shx_amr_run_to_end((shx_amr*)h->p, batch, max_steps, &steps));
    SHX_CALL(env, amr_resync(h));
  } else {
    SHX_CALL(env, shx_run_to_end((shx_sim*)h->p, batch, max_steps, &steps));
  }
  napi_value out;
  NAPI_CALL(env, napi_create_int32(env, steps, &out));
  return out;
}

static napi_value js_read_clock(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Handle* h = get_handle(env, argv[0], 0);
  if (!h) return NULL;
  double out[11];
  SHX_CALL(env, shx_read_clock((shx_sim*)h->p, out));
  return f64_array(env, out, 11);
}

static napi_value js_diagnostics(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Handle* h = get_handle(env, argv[0], 0);
  if (!h) return NULL;
  double out[24];
  SHX_CALL(env, shx_diagnostics((shx_sim*)h->p, out));
  return f64_array(env, out, 24);
}

static napi_value js_flux_mismatch(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Handle* h = get_handle(env, argv[0], 0);
  if (!h) return NULL;
  int64_t mis = 0, faces = 0;
  SHX_CALL(env, shx_flux_mismatch((shx_sim*)h->p, &mis, &faces));
  double v[2] = {(double)mis, (double)faces};
  return f64_array(env, v, 2);
}

static napi_value js_profile(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  Handle* h = get_handle(env, argv[0], 0);
  if (!h) return NULL;
  int32_t steps;
  NAPI_CALL(env, napi_get_value_int32(env, argv[1], &steps));
  double ms[5];
  SHX_CALL(env, shx_profile((shx_sim*)h->p, steps, ms));
  return f64_array(env, ms, 5);
}

/* ---------------- async readbacks ---------------- */
typedef struct {
  napi_async_work work;
  napi_deferred deferred;
  shx_readback* rb;    /* begun on the JS thread; the worker only waits for it */
  int kind;            /* 0 state, 1 diagnostics, 2 the clock alone */
  int rc;
  char err[512];
} Async;

static void async_execute(napi_env env, void* data) {
  (void)env;
  Async* a = (Async*)data;
  a->rc = shx_readback_wait(a->rb);
  if (a->rc != SHX_OK) snprintf(a->err, sizeof a->err, "libshx error %d: %s", a->rc, shx_last_error());
}

/* {u1, u2, clock} or {diag, clock}: the clock read with the data, after the same steps. */
static napi_value readback_value(napi_env env, Async* a) {
  const void *u1, *u2;
  const double *clock, *diag;
  size_t ncell;
  int precision;
  if (shx_readback_data(a->rb, &u1, &u2, &ncell, &precision, &clock, &diag) != SHX_OK) return NULL;
  napi_value o, ck;
  if (napi_create_object(env, &o) != napi_ok) return NULL;
  ck = f64_array(env, clock, 11);
  if (!ck || napi_set_named_property(env, o, "clock", ck) != napi_ok) return NULL;
  if (a->kind == 0) {
    size_t n = 4 * ncell, es = precision == 64 ? 8 : 4;
    void *p1, *p2;
    napi_value x = new_typed(env, precision, n, &p1);
    napi_value y = new_typed(env, precision, n, &p2);
    if (!x || !y) return NULL;
    memcpy(p1, u1, n * es);
    memcpy(p2, u2, n * es);
    if (napi_set_named_property(env, o, "u1", x) != napi_ok || napi_set_named_property(env, o, "u2", y) != napi_ok)
      return NULL;
  } else if (a->kind == 1) {
    napi_value d = f64_array(env, diag, 24);
    if (!d || napi_set_named_property(env, o, "diag", d) != napi_ok) return NULL;
  }
  return o;
}

static void async_complete(napi_env env, napi_status status, void* data) {
  Async* a = (Async*)data;
  napi_value result = NULL;
  if (status == napi_ok && a->rc == SHX_OK) result = readback_value(env, a);
  if (result) {
    napi_resolve_deferred(env, a->deferred, result);
  } else {
    napi_value msg, err;
    const char* m = a->rc != SHX_OK ? a->err : status != napi_ok ? "async readback cancelled" : "out of memory";
    napi_create_string_utf8(env, m, NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, NULL, msg, &err);
    napi_reject_deferred(env, a->deferred, err);
  }
  napi_delete_async_work(env, a->work);
  shx_readback_free(a->rb);
  free(a);
}

static napi_value start_async(napi_env env, Handle* h, int kind) {
  Async* a = (Async*)calloc(1, sizeof(Async));
  if (!a) { napi_throw_error(env, NULL, "out of memory"); return NULL; }
  a->kind = kind;
  const int what = kind == 0 ? SHX_READ_STATE : kind == 1 ? SHX_READ_DIAG : 0;
  int rc = h->amr ? shx_amr_readback_begin((shx_amr*)h->p, what, &a->rb)
                  : shx_readback_begin((shx_sim*)h->p, what, &a->rb);
  if (rc != SHX_OK) { free(a); return throw_status(env, rc); }
  napi_value promise, name;
  if (napi_create_promise(env, &a->deferred, &promise) != napi_ok ||
      napi_create_string_utf8(env, "shx.readback", NAPI_AUTO_LENGTH, &name) != napi_ok ||
      napi_create_async_work(env, NULL, name, async_execute, async_complete, a, &a->work) != napi_ok ||
      napi_queue_async_work(env, a->work) != napi_ok) {
    shx_readback_free(a->rb);
    free(a);
    napi_throw_error(env, NULL, "N-API: could not queue the readback");
    return NULL;
  }
  return promise;
}

static napi_value js_read_state_async(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  void* data = NULL;
  NAPI_CALL(env, napi_get_value_external(env, argv[0], &data));
  Handle* h = (Handle*)data;
  if (!h || !h->p) { napi_throw_error(env, NULL, "the simulation was destroyed"); return NULL; }
  return start_async(env, h, 0);
}

static napi_value js_diagnostics_async(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  void* data = NULL;
  NAPI_CALL(env, napi_get_value_external(env, argv[0], &data));
  Handle* h = (Handle*)data;
  if (!h || !h->p) { napi_throw_error(env, NULL, "the simulation was destroyed"); return NULL; }
  return start_async(env, h, 1);
}

/* The clock alone, read back as the run loop's clockCopy does (src/ui/api.ts:503-534). */
static napi_value js_clock_async(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  void* data = NULL;
  NAPI_CALL(env, napi_get_value_external(env, argv[0], &data));
  Handle* h = (Handle*)data;
  if (!h || !h->p) { napi_throw_error(env, NULL, "the simulation was destroyed"); return NULL; }
  return start_async(env, h, 2);
}

/* ---------------- block grid ---------------- */
static napi_value js_amr_create(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  shx_config c;
  config_from(env, argv[0], &c);
  shx_amr_layout lay;
  memset(&lay, 0, sizeof lay);
  lay.root[0] = lay.root[1] = lay.root[2] = 1;
  int3_prop(env, argv[1], "root", lay.root);
  lay.path = (int)num_prop(env, argv[1], "path", SHX_AMR_PATH_FUSED);
  lay.compensated = (int)num_prop(env, argv[1], "compensated", 1);
  lay.time_bins = (int)num_prop(env, argv[1], "timeBins", 0);
  lay.regrid_every = (int)num_prop(env, argv[1], "regridEvery", 0);
  lay.max_level = (int)num_prop(env, argv[1], "maxLevel", 0);
  lay.refine_above = num_prop(env, argv[1], "refineAbove", 0.8);
  lay.coarsen_below = num_prop(env, argv[1], "coarsenBelow", 0.2);
  bool is = false;
  napi_is_typedarray(env, argv[2], &is);
  if (!is) { napi_throw_type_error(env, NULL, "blocks is an Int32Array of (level, ix, iy, iz) rows"); return NULL; }
  napi_typedarray_type t;
  size_t len;
  void* data;
  napi_value ab;
  size_t off;
  NAPI_CALL(env, napi_get_typedarray_info(env, argv[2], &t, &len, &data, &ab, &off));
  if (t != napi_int32_array || len % 4 != 0) { napi_throw_type_error(env, NULL, "blocks is an Int32Array of (level, ix, iy, iz) rows"); return NULL; }
  shx_amr* sim = NULL;
  SHX_CALL(env, shx_amr_create(&c, &lay, (const int*)data, (int)(len / 4), &sim));
  Handle* h = (Handle*)calloc(1, sizeof(Handle));
  h->amr = 1;
  h->p = sim;
  h->precision = c.precision;
  h->nblocks = (int)(len / 4);
  h->cells = c.ndim == 2 ? 256 : 4096;
  h->ncell = (size_t)h->nblocks * h->cells;
  return wrap_handle(env, h);
}

static napi_value js_amr_table(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Handle* h = get_handle(env, argv[0], 1);
  if (!h) return NULL;
  size_t n = (size_t)h->nblocks * 32;
  void* data;
  napi_value ab, arr;
  NAPI_CALL(env, napi_create_arraybuffer(env, n * sizeof(int), &data, &ab));
  NAPI_CALL(env, napi_create_typedarray(env, napi_int32_array, n, ab, 0, &arr));
  SHX_CALL(env, shx_amr_table((shx_amr*)h->p, (int*)data, n));
  return arr;
}

static napi_value js_amr_set_state(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  size_t argc = 5;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 3) { napi_throw_type_error(env, NULL, "amrSetState(handle, u1, u2[, time, steps])"); return NULL; }
  Handle* h = get_handle(env, argv[0], 1);
  if (!h) return NULL;
  void* u1 = typed_data(env, argv[1], h->precision, 4 * h->ncell);
  if (!u1) return NULL;
  void* u2 = typed_data(env, argv[2], h->precision, 4 * h->ncell);
  if (!u2) return NULL;
  double time = 0, steps = 0;
  if (argc > 3) napi_get_value_double(env, argv[3], &time);
  if (argc > 4) napi_get_value_double(env, argv[4], &steps);
  if (h->precision == 64) SHX_CALL(env, shx_amr_set_state_f64((shx_amr*)h->p, u1, u2, h->ncell, time, (int)steps));
  else SHX_CALL(env, shx_amr_set_state((shx_amr*)h->p, u1, u2, h->ncell, time, (int)steps));
  return undefined(env);
}

static napi_value js_amr_read_state(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Handle* h = get_handle(env, argv[0], 1);
  if (!h) return NULL;
  void *d1, *d2;
  napi_value a = new_typed(env, h->precision, 4 * h->ncell, &d1);
  napi_value b = new_typed(env, h->precision, 4 * h->ncell, &d2);
  if (!a || !b) { napi_throw_error(env, NULL, "out of memory"); return NULL; }
  if (h->precision == 64) SHX_CALL(env, shx_amr_read_state_f64((shx_amr*)h->p, d1, d2));
  else SHX_CALL(env, shx_amr_read_state((shx_amr*)h->p, d1, d2));
  return state_pair(env, a, b);
}

static napi_value js_amr_step(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  Handle* h = get_handle(env, argv[0], 1);
  if (!h) return NULL;
  int32_t count;
  NAPI_CALL(env, napi_get_value_int32(env, argv[1], &count));
  int taken = 0;
  SHX_CALL(env, shx_amr_step((shx_amr*)h->p, count, &taken));
  SHX_CALL(env, amr_resync(h));
  napi_value out;
  NAPI_CALL(env, napi_create_int32(env, taken, &out));
  return out;
}

static napi_value js_amr_sync(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Handle* h = get_handle(env, argv[0], 1);
  if (!h) return NULL;
  SHX_CALL(env, shx_amr_sync((shx_amr*)h->p));
  return undefined(env);
}

static napi_value js_amr_read_clock(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Handle* h = get_handle(env, argv[0], 1);
  if (!h) return NULL;
  double out[11];
  SHX_CALL(env, shx_amr_read_clock((shx_amr*)h->p, out));
  return f64_array(env, out, 11);
}

static napi_value js_amr_diagnostics(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Handle* h = get_handle(env, argv[0], 1);
  if (!h) return NULL;
  double out[24];
  SHX_CALL(env, shx_amr_diagnostics((shx_amr*)h->p, out));
  return f64_array(env, out, 24);
}

static napi_value js_amr_profile(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  Handle* h = get_handle(env, argv[0], 1);
  if (!h) return NULL;
  int32_t steps;
  NAPI_CALL(env, napi_get_value_int32(env, argv[1], &steps));
  double ms[5];
  SHX_CALL(env, shx_amr_profile((shx_amr*)h->p, steps, ms));
  return f64_array(env, ms, 5);
}

static napi_value js_amr_stats(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Handle* h = get_handle(env, argv[0], 1);
  if (!h) return NULL;
  int out[5];
  SHX_CALL(env, shx_amr_stats((shx_amr*)h->p, out));
  double v[5];
  for (int i = 0; i < 5; i++) v[i] = out[i];
  return f64_array(env, v, 5);
}

static napi_value js_amr_indicator(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Handle* h = get_handle(env, argv[0], 1);
  if (!h) return NULL;
  double* v = (double*)malloc((size_t)h->nblocks * sizeof(double));
  if (!v) { napi_throw_error(env, NULL, "out of memory"); return NULL; }
  int rc = shx_amr_indicator((shx_amr*)h->p, v);
  if (rc != SHX_OK) { free(v); return throw_status(env, rc); }
  napi_value out = f64_array(env, v, h->nblocks);
  free(v);
  return out;
}

static napi_value js_amr_regrid(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  Handle* h = get_handle(env, argv[0], 1);
  if (!h) return NULL;
  bool is = false;
  napi_is_typedarray(env, argv[1], &is);
  napi_typedarray_type t;
  size_t len = 0;
  void* data = NULL;
  napi_value ab;
  size_t off;
  if (is) NAPI_CALL(env, napi_get_typedarray_info(env, argv[1], &t, &len, &data, &ab, &off));
  if (!is || t != napi_int32_array || len % 4 != 0) {
    napi_throw_type_error(env, NULL, "blocks is an Int32Array of (level, ix, iy, iz) rows");
    return NULL;
  }
  SHX_CALL(env, shx_amr_regrid((shx_amr*)h->p, (const int*)data, (int)(len / 4)));
  h->nblocks = (int)(len / 4);
  h->ncell = (size_t)h->nblocks * h->cells;
  return undefined(env);
}

/* ---------------- module ---------------- */
#define EXPORT(name, fn) {name, NULL, fn, NULL, NULL, NULL, (napi_property_attributes)(napi_enumerable | napi_writable), NULL}

static napi_value init(napi_env env, napi_value exports) {
  napi_property_descriptor props[] = {
      EXPORT("version", js_version),
      EXPORT("create", js_create),
      EXPORT("destroy", js_destroy),
      EXPORT("setState", js_set_state),
      EXPORT("readState", js_read_state),
      EXPORT("readStateAsync", js_read_state_async),
      EXPORT("step", js_step),
      EXPORT("writeParams", js_write_params),
      EXPORT("buildHydro", js_build_hydro),
      EXPORT("sync", js_sync),
      EXPORT("runToEnd", js_run_to_end),
      EXPORT("readClock", js_read_clock),
      EXPORT("diagnostics", js_diagnostics),
      EXPORT("diagnosticsAsync", js_diagnostics_async),
      EXPORT("clockAsync", js_clock_async),
      EXPORT("fluxMismatch", js_flux_mismatch),
      EXPORT("profile", js_profile),
      EXPORT("amrCreate", js_amr_create),
      EXPORT("amrTable", js_amr_table),
      EXPORT("amrSetState", js_amr_set_state),
      EXPORT("amrReadState", js_amr_read_state),
      EXPORT("amrStep", js_amr_step),
      EXPORT("amrSync", js_amr_sync),
      EXPORT("amrReadClock", js_amr_read_clock),
      EXPORT("amrDiagnostics", js_amr_diagnostics),
      EXPORT("amrProfile", js_amr_profile),
      EXPORT("amrStats", js_amr_stats),
      EXPORT("amrIndicator", js_amr_indicator),
      EXPORT("amrRegrid", js_amr_regrid),
  };
  if (napi_define_properties(env, exports, sizeof props / sizeof props[0], props) != napi_ok) return NULL;
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
