/*
 * Node.js N-API addon over the engine's C ABI (include/mpm_engine.h).
 *
 * Exposes the C ABI one to one as functions on an external engine handle;
 * node/index.js builds the reference's `MPM3` class (src/three/mpm.js:7-136)
 * on top of it. Typed arrays cross the boundary without copies on the JS side:
 * Float64Array (or Float32Array) views are passed straight to
 * mpm_upload/mpm_download.
 */
#define NAPI_VERSION 4
#include <node_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mpm_engine.h"

#define CHECK_NAPI(env, call)                                   \
  do {                                                          \
    if ((call) != napi_ok) {                                    \
      napi_throw_error((env), NULL, "N-API call failed: " #call); \
      return NULL;                                              \
    }                                                           \
  } while (0)

static napi_value throw_mpm(napi_env env, int rc, mpm_engine* e) {
  char code[16], msg[600];
  snprintf(code, sizeof code, "MPM%d", rc);
  snprintf(msg, sizeof msg, "mpm error %d: %s", rc, e ? mpm_last_error(e) : "");
  napi_throw_error(env, code, msg);
  return NULL;
}

static void finalize_engine(napi_env env, void* data, void* hint) {
  (void)env;
  (void)hint;
  mpm_engine** slot = (mpm_engine**)data;
  if (*slot) mpm_destroy(*slot);
  free(slot);
}

static mpm_engine* get_engine(napi_env env, napi_value v) {
  void* data = NULL;
  if (napi_get_value_external(env, v, &data) != napi_ok || !data) {
    napi_throw_type_error(env, NULL, "expected an engine handle");
    return NULL;
  }
  mpm_engine* e = *(mpm_engine**)data;
  if (!e) napi_throw_error(env, NULL, "engine was destroyed");
  return e;
}

static double get_num(napi_env env, napi_value obj, const char* key, double dflt) {
  bool has = false;
  napi_value v;
  double out = dflt;
  if (napi_has_named_property(env, obj, key, &has) == napi_ok && has &&
      napi_get_named_property(env, obj, key, &v) == napi_ok) {
    napi_valuetype t;
    napi_typeof(env, v, &t);
    if (t == napi_number) napi_get_value_double(env, v, &out);
    if (t == napi_boolean) {
      bool b;
      napi_get_value_bool(env, v, &b);
      out = b ? 1 : 0;
    }
  }
  return out;
}

static void get_vec(napi_env env, napi_value obj, const char* key, double* out, int n) {
  bool has = false;
  napi_value arr, el;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return;
  napi_get_named_property(env, obj, key, &arr);
  for (int i = 0; i < n; ++i) {
    if (napi_get_element(env, arr, (uint32_t)i, &el) == napi_ok) {
      double d;
      if (napi_get_value_double(env, el, &d) == napi_ok) out[i] = d;
    }
  }
}

/* create({lower, upper, cellSize, capacity, wallMode, gravity,
 *         gravityOnParticles, massEps, sortEvery, strictDomain}) */
static napi_value js_create(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1) {
    napi_throw_type_error(env, NULL, "create(config) needs a config object");
    return NULL;
  }
  mpm_config c;
  memset(&c, 0, sizeof c);
  get_vec(env, argv[0], "lower", c.lower, 3);
  get_vec(env, argv[0], "upper", c.upper, 3);
  c.cell_size = get_num(env, argv[0], "cellSize", 0);
  c.capacity = (uint32_t)get_num(env, argv[0], "capacity", 0);
  double walls[6] = {0, 0, 0, 0, 0, 0};
  get_vec(env, argv[0], "wallMode", walls, 6);
  for (int f = 0; f < 6; ++f) c.wall_mode[f] = (int32_t)walls[f];
  get_vec(env, argv[0], "gravity", c.gravity, 3);
  c.gravity_on_particles = (int32_t)get_num(env, argv[0], "gravityOnParticles", 0);
  c.mass_eps = get_num(env, argv[0], "massEps", 1e-12);
  c.sort_every = (int32_t)get_num(env, argv[0], "sortEvery", 8);
  c.strict_domain = (int32_t)get_num(env, argv[0], "strictDomain", 0);
  c.guard_off = (int32_t)get_num(env, argv[0], "guardOff", 0);
  c.cfl_number = get_num(env, argv[0], "cflNumber", 0);
  c.cfl_cap = get_num(env, argv[0], "cflCap", 0);
  c.grad_number = get_num(env, argv[0], "gradNumber", 0);
  mpm_engine* e = NULL;
  int rc = mpm_create(&c, &e);
  if (rc) return throw_mpm(env, rc, NULL);
  mpm_engine** slot = (mpm_engine**)malloc(sizeof *slot);
  *slot = e;
  napi_value ext;
  CHECK_NAPI(env, napi_create_external(env, slot, finalize_engine, NULL, &ext));
  return ext;
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  void* data = NULL;
  if (argc < 1 || napi_get_value_external(env, argv[0], &data) != napi_ok || !data) return NULL;
  mpm_engine** slot = (mpm_engine**)data;
  if (*slot) mpm_destroy(*slot);
  *slot = NULL;
  return NULL;
}

/* setMaterials(h, [{model, E, mu, lambda, yieldStrain, friction, thermal:{...},
 * plastic:{...}, pressure:{...}, fracture:{enabled, onset, failure}}]) */
static napi_value js_set_materials(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm_engine* e = argc >= 2 ? get_engine(env, argv[0]) : NULL;
  if (!e) return NULL;
  uint32_t n = 0;
  CHECK_NAPI(env, napi_get_array_length(env, argv[1], &n));
  if (n < 1 || n > 64) {
    napi_throw_range_error(env, NULL, "1..64 materials required");
    return NULL;
  }
  mpm_material mats[64];
  memset(mats, 0, sizeof mats);
  for (uint32_t i = 0; i < n; ++i) {
    napi_value m, th, pl, pr, fr;
    bool has;
    CHECK_NAPI(env, napi_get_element(env, argv[1], i, &m));
    mpm_material* o = &mats[i];
    /* the reference's defaults for every record, then the given fields */
    mpm_material_defaults((int32_t)get_num(env, m, "model", 0), (float)get_num(env, m, "E", 1), o);
    o->mu = (float)get_num(env, m, "mu", o->mu);
    o->lambda = (float)get_num(env, m, "lambda", o->lambda);
    o->yield_strain = (float)get_num(env, m, "yieldStrain", o->yield_strain);
    o->friction = (float)get_num(env, m, "friction", o->friction);
    if (napi_has_named_property(env, m, "pressure", &has) == napi_ok && has &&
        napi_get_named_property(env, m, "pressure", &pr) == napi_ok) {
      o->pr_cohesion = (float)get_num(env, pr, "cohesion", o->pr_cohesion);
      o->pr_slope = (float)get_num(env, pr, "slope", o->pr_slope);
      o->pr_dilatancy = (float)get_num(env, pr, "dilatancy", o->pr_dilatancy);
      o->pr_tension = (float)get_num(env, pr, "tension", o->pr_tension);
      o->pr_alpha = (float)get_num(env, pr, "alpha", o->pr_alpha);
      o->pr_plateau = (float)get_num(env, pr, "plateau", o->pr_plateau);
      o->pr_gamma = (float)get_num(env, pr, "gamma", o->pr_gamma);
      o->pr_densify = (float)get_num(env, pr, "densify", o->pr_densify);
      o->pr_beta = (float)get_num(env, pr, "beta", o->pr_beta);
      o->pr_epsD = (float)get_num(env, pr, "epsD", o->pr_epsD);
    }
    if (napi_has_named_property(env, m, "fracture", &has) == napi_ok && has &&
        napi_get_named_property(env, m, "fracture", &fr) == napi_ok) {
      bool on = false;
      napi_value v;
      if (napi_get_named_property(env, fr, "enabled", &v) == napi_ok) napi_get_value_bool(env, v, &on);
      o->fracture = on ? 1 : 0;
      o->fr_onset = (float)get_num(env, fr, "onset", o->fr_onset);
      o->fr_failure = (float)get_num(env, fr, "failure", o->fr_failure);
    }
    if (napi_has_named_property(env, m, "thermal", &has) == napi_ok && has &&
        napi_get_named_property(env, m, "thermal", &th) == napi_ok) {
      o->th_Tref = (float)get_num(env, th, "Tref", o->th_Tref);
      o->th_Thot = (float)get_num(env, th, "Thot", o->th_Thot);
      o->th_hotRatio = (float)get_num(env, th, "hotRatio", o->th_hotRatio);
      o->th_A = (float)get_num(env, th, "A", o->th_A);
      o->th_H = (float)get_num(env, th, "H", o->th_H);
      o->th_rate = (float)get_num(env, th, "rate", o->th_rate);
      o->th_Q = (float)get_num(env, th, "Q", o->th_Q);
      o->th_n = (float)get_num(env, th, "n", o->th_n);
      o->th_stress = (float)get_num(env, th, "stress", o->th_stress);
      o->th_kelvin_offset = (float)get_num(env, th, "kelvinOffset", o->th_kelvin_offset);
    }
    if (napi_has_named_property(env, m, "plastic", &has) == napi_ok && has &&
        napi_get_named_property(env, m, "plastic", &pl) == napi_ok) {
      o->pl_A = (float)get_num(env, pl, "A", o->pl_A);
      o->pl_B = (float)get_num(env, pl, "B", o->pl_B);
      o->pl_n = (float)get_num(env, pl, "n", o->pl_n);
      o->pl_C = (float)get_num(env, pl, "C", o->pl_C);
      o->pl_rate0 = (float)get_num(env, pl, "rate0", o->pl_rate0);
      o->pl_Tref = (float)get_num(env, pl, "Tref", o->pl_Tref);
      o->pl_Tm = (float)get_num(env, pl, "Tm", o->pl_Tm);
      o->pl_m = (float)get_num(env, pl, "m", o->pl_m);
    }
  }
  int rc = mpm_set_materials(e, mats, (int32_t)n);
  if (rc) return throw_mpm(env, rc, e);
  return NULL;
}

/* Fills an mpm_particles view from {count, x, y, z, vx, vy, vz, F, C, mass,
 * volume, temperature, plastic, cachedTau, material, pinned, cacheValid}:
 * floating arrays all Float64Array or all Float32Array, flags Uint8Array. */
static int particle_view(napi_env env, napi_value obj, mpm_particles* p) {
  memset(p, 0, sizeof *p);
  p->count = (uint32_t)get_num(env, obj, "count", 0);
  p->dtype = -1;
  struct {
    const char* key;
    void** dst;
    int flag;
  } fields[] = {{"x", &p->x, 0}, {"y", &p->y, 0}, {"z", &p->z, 0}, {"vx", &p->vx, 0}, {"vy", &p->vy, 0},
                {"vz", &p->vz, 0}, {"F", &p->F, 0}, {"C", &p->C, 0}, {"mass", &p->mass, 0},
                {"volume", &p->volume, 0}, {"temperature", &p->temperature, 0}, {"plastic", &p->plastic, 0},
                {"cachedTau", &p->cached_tau, 0}, {"material", (void**)&p->material, 1},
                {"pinned", (void**)&p->pinned, 1}, {"cacheValid", (void**)&p->cache_valid, 1},
                {"jp", &p->jp, 0}, {"dilation", &p->dilation, 0}, {"damage", &p->damage, 0}};
  for (size_t k = 0; k < sizeof fields / sizeof fields[0]; ++k) {
    bool has = false;
    napi_value v;
    if (napi_has_named_property(env, obj, fields[k].key, &has) != napi_ok || !has) continue;
    napi_get_named_property(env, obj, fields[k].key, &v);
    bool is_ta = false;
    napi_is_typedarray(env, v, &is_ta);
    if (!is_ta) continue;
    napi_typedarray_type type;
    size_t length, offset;
    void* data;
    napi_value ab;
    napi_get_typedarray_info(env, v, &type, &length, &data, &ab, &offset);
    const size_t need = (fields[k].key[0] == 'F' || fields[k].key[0] == 'C' || !strcmp(fields[k].key, "cachedTau"))
                            ? (size_t)p->count * 9
                            : (size_t)p->count;
    if (length < need) {
      napi_throw_range_error(env, NULL, fields[k].key);
      return -1;
    }
    if (fields[k].flag) {
      if (type != napi_uint8_array) {
        napi_throw_type_error(env, NULL, "flag arrays must be Uint8Array");
        return -1;
      }
    } else {
      const int dt = type == napi_float64_array ? MPM_F64 : type == napi_float32_array ? MPM_F32 : -2;
      if (dt < 0 || (p->dtype >= 0 && p->dtype != dt)) {
        napi_throw_type_error(env, NULL, "floating arrays must all be Float64Array or all Float32Array");
        return -1;
      }
      p->dtype = dt;
    }
    *fields[k].dst = data;
  }
  if (p->dtype < 0) p->dtype = MPM_F64;
  return 0;
}

static napi_value js_upload(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm_engine* e = argc >= 2 ? get_engine(env, argv[0]) : NULL;
  if (!e) return NULL;
  mpm_particles p;
  if (particle_view(env, argv[1], &p)) return NULL;
  int rc = mpm_upload(e, &p);
  if (rc) return throw_mpm(env, rc, e);
  return NULL;
}

static napi_value js_download(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm_engine* e = argc >= 2 ? get_engine(env, argv[0]) : NULL;
  if (!e) return NULL;
  mpm_particles p;
  if (particle_view(env, argv[1], &p)) return NULL;
  int rc = mpm_download(e, &p);
  if (rc) return throw_mpm(env, rc, e);
  return NULL;
}

static napi_value js_step(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm_engine* e = argc >= 2 ? get_engine(env, argv[0]) : NULL;
  if (!e) return NULL;
  double dt = 0;
  int32_t n = 1;
  napi_get_value_double(env, argv[1], &dt);
  if (argc >= 3) napi_get_value_int32(env, argv[2], &n);
  int rc = mpm_step(e, dt, n);
  if (rc) return throw_mpm(env, rc, e);
  return NULL;
}

static napi_value js_synchronize(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm_engine* e = argc >= 1 ? get_engine(env, argv[0]) : NULL;
  if (!e) return NULL;
  int rc = mpm_synchronize(e);
  if (rc) return throw_mpm(env, rc, e);
  return NULL;
}

static napi_value js_stable_dt(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm_engine* e = argc >= 1 ? get_engine(env, argv[0]) : NULL;
  if (!e) return NULL;
  double cap = 0, cfl = 0, out = 0;
  if (argc >= 2) napi_get_value_double(env, argv[1], &cap);
  if (argc >= 3) napi_get_value_double(env, argv[2], &cfl);
  int rc = mpm_stable_dt(e, cap, cfl, &out);
  if (rc) return throw_mpm(env, rc, e);
  napi_value r;
  CHECK_NAPI(env, napi_create_double(env, out, &r));
  return r;
}

/* advanceFrame(engine, dt, settings) -> {substeps, advanced}: one frame of
 * Simulation3.step for an all-solid scene (mpm_advance_frame). settings:
 * timeStepScale, maxSubsteps, velocityLimit, boundsLo, boundsHi, antiClump,
 * volumeUnit, gravityMode (0 none, 1 uniform, 2 attract, 3 repel),
 * gravityStrength, gravityDirection, gravityCenter, gravitySoftening; missing
 * fields keep the reference defaults. */
static napi_value js_advance_frame(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm_engine* e = argc >= 2 ? get_engine(env, argv[0]) : NULL;
  if (!e) return NULL;
  double dt = 0;
  napi_get_value_double(env, argv[1], &dt);
  mpm_frame_config c;
  int rc = mpm_frame_defaults(e, &c);
  if (rc) return throw_mpm(env, rc, e);
  if (argc >= 3) {
    napi_value o = argv[2];
    c.time_step_scale = get_num(env, o, "timeStepScale", c.time_step_scale);
    c.max_substeps = (int32_t)get_num(env, o, "maxSubsteps", c.max_substeps);
    c.velocity_limit = get_num(env, o, "velocityLimit", c.velocity_limit);
    get_vec(env, o, "boundsLo", c.bounds_lo, 3);
    get_vec(env, o, "boundsHi", c.bounds_hi, 3);
    c.anti_clump = (int32_t)get_num(env, o, "antiClump", c.anti_clump);
    c.volume_unit = get_num(env, o, "volumeUnit", c.volume_unit);
    c.gravity_mode = (int32_t)get_num(env, o, "gravityMode", c.gravity_mode);
    c.gravity_strength = get_num(env, o, "gravityStrength", c.gravity_strength);
    get_vec(env, o, "gravityDirection", c.gravity_direction, 3);
    get_vec(env, o, "gravityCenter", c.gravity_center, 3);
    c.gravity_softening = get_num(env, o, "gravitySoftening", c.gravity_softening);
  }
  mpm_frame_report r;
  rc = mpm_advance_frame(e, &c, dt, &r);
  if (rc) return throw_mpm(env, rc, e);
  napi_value out, x;
  CHECK_NAPI(env, napi_create_object(env, &out));
  CHECK_NAPI(env, napi_create_int32(env, r.substeps, &x));
  CHECK_NAPI(env, napi_set_named_property(env, out, "substeps", x));
  CHECK_NAPI(env, napi_create_double(env, r.advanced, &x));
  CHECK_NAPI(env, napi_set_named_property(env, out, "advanced", x));
  return out;
}

static void set_num(napi_env env, napi_value obj, const char* key, double v) {
  napi_value x;
  napi_create_double(env, v, &x);
  napi_set_named_property(env, obj, key, x);
}

static void set_vec(napi_env env, napi_value obj, const char* key, const double* v, int n) {
  napi_value arr, x;
  napi_create_array_with_length(env, (size_t)n, &arr);
  for (int i = 0; i < n; ++i) {
    napi_create_double(env, v[i], &x);
    napi_set_element(env, arr, (uint32_t)i, x);
  }
  napi_set_named_property(env, obj, key, arr);
}

static napi_value js_conservation(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm_engine* e = argc >= 1 ? get_engine(env, argv[0]) : NULL;
  if (!e) return NULL;
  mpm_conservation c;
  int rc = mpm_conservation_report(e, &c);
  if (rc) return throw_mpm(env, rc, e);
  napi_value o, b;
  CHECK_NAPI(env, napi_create_object(env, &o));
  set_num(env, o, "particleMass", c.particle_mass);
  set_num(env, o, "gridMass", c.grid_mass);
  /* fixed-point totals as BigInt-free strings keep all 64 bits */
  char buf[32];
  snprintf(buf, sizeof buf, "%lld", (long long)c.particle_mass_q);
  napi_create_string_utf8(env, buf, NAPI_AUTO_LENGTH, &b);
  napi_set_named_property(env, o, "particleMassQ", b);
  snprintf(buf, sizeof buf, "%lld", (long long)c.grid_mass_q);
  napi_create_string_utf8(env, buf, NAPI_AUTO_LENGTH, &b);
  napi_set_named_property(env, o, "gridMassQ", b);
  set_num(env, o, "massScale", c.mass_scale);
  set_vec(env, o, "momentum", c.momentum, 3);
  set_vec(env, o, "gridMomentum", c.grid_momentum, 3);
  set_vec(env, o, "angular", c.angular, 3);
  set_num(env, o, "kinetic", c.kinetic);
  set_num(env, o, "affine", c.affine);
  set_num(env, o, "minJ", c.min_J);
  set_num(env, o, "maxSpeed", c.max_speed);
  set_num(env, o, "nonfinite", (double)c.nonfinite);
  return o;
}

/* snapshot(h, previewLimit, codes|null) -> {count, renderCount, sampleStride,
 * solids, liquids, models, mass, min, max, minJ, maxStress, maxPlastic,
 * nonfinite, data: Float32Array(renderCount * 8)} */
static napi_value js_snapshot(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm_engine* e = argc >= 2 ? get_engine(env, argv[0]) : NULL;
  if (!e) return NULL;
  double limit = 0;
  if (napi_get_value_double(env, argv[1], &limit) != napi_ok || !(limit >= 1) || limit > 4294967295.0) {
    napi_throw_range_error(env, NULL, "previewLimit must be a positive integer");
    return NULL;
  }
  int32_t codes[64];
  int32_t ncodes = 0;
  bool is_array = false;
  if (argc >= 3) napi_is_array(env, argv[2], &is_array);
  if (is_array) {
    uint32_t n = 0;
    napi_get_array_length(env, argv[2], &n);
    if (n > 64) n = 64;
    for (uint32_t k = 0; k < n; ++k) {
      napi_value v;
      double d = 11;
      napi_get_element(env, argv[2], k, &v);
      napi_get_value_double(env, v, &d);
      codes[k] = (int32_t)d;
    }
    ncodes = (int32_t)n;
  }
  mpm_snapshot_stats st;
  int rc = mpm_snapshot(e, (uint32_t)limit, is_array ? codes : NULL, ncodes, NULL, &st);
  if (rc) return throw_mpm(env, rc, e);
  napi_value ab, data, o;
  void* raw = NULL;
  CHECK_NAPI(env, napi_create_arraybuffer(env, (size_t)st.render_count * 32, &raw, &ab));
  rc = mpm_snapshot(e, (uint32_t)limit, is_array ? codes : NULL, ncodes, (float*)raw, &st);
  if (rc) return throw_mpm(env, rc, e);
  CHECK_NAPI(env, napi_create_typedarray(env, napi_float32_array, (size_t)st.render_count * 8, ab, 0, &data));
  CHECK_NAPI(env, napi_create_object(env, &o));
  set_num(env, o, "count", st.count);
  set_num(env, o, "renderCount", st.render_count);
  set_num(env, o, "sampleStride", st.sample_stride);
  set_num(env, o, "solids", (double)st.solids);
  set_num(env, o, "liquids", (double)st.liquids);
  double models[12];
  for (int k = 0; k < 12; ++k) models[k] = (double)st.models[k];
  set_vec(env, o, "models", models, 12);
  set_num(env, o, "mass", st.mass);
  set_num(env, o, "min", st.min_temperature);
  set_num(env, o, "max", st.max_temperature);
  set_num(env, o, "minJ", st.min_J);
  set_num(env, o, "maxStress", st.max_stress);
  set_num(env, o, "maxPlastic", st.max_plastic);
  set_num(env, o, "nonfinite", (double)st.nonfinite);
  napi_set_named_property(env, o, "data", data);
  return o;
}

/* framePacket(headerString, Float32Array rows*8) -> ArrayBuffer */
static napi_value js_frame_packet(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  size_t len = 0;
  if (argc < 2 || napi_get_value_string_utf8(env, argv[0], NULL, 0, &len) != napi_ok) {
    napi_throw_type_error(env, NULL, "framePacket(header string, Float32Array)");
    return NULL;
  }
  char* head = (char*)malloc(len + 1);
  if (!head) return NULL;
  napi_get_value_string_utf8(env, argv[0], head, len + 1, &len);
  napi_typedarray_type type;
  size_t n = 0, offset = 0;
  void* p = NULL;
  napi_value buf;
  bool is_ta = false;
  napi_is_typedarray(env, argv[1], &is_ta);
  if (!is_ta || napi_get_typedarray_info(env, argv[1], &type, &n, &p, &buf, &offset) != napi_ok ||
      type != napi_float32_array || n % 8) {
    free(head);
    napi_throw_type_error(env, NULL, "rows must be a Float32Array of 8 values per row");
    return NULL;
  }
  uint64_t need = 0;
  mpm_frame_packet(head, (uint32_t)len, (const float*)p, n / 8, NULL, 0, &need);
  void* out = NULL;
  napi_value ab;
  if (napi_create_arraybuffer(env, (size_t)need, &out, &ab) != napi_ok) {
    free(head);
    return NULL;
  }
  int rc = mpm_frame_packet(head, (uint32_t)len, (const float*)p, n / 8, (uint8_t*)out, need, &need);
  free(head);
  if (rc) return throw_mpm(env, rc, NULL);
  return ab;
}

/* ------------------------------------------------------------ split step */
static int typed(napi_env env, napi_value v, napi_typedarray_type want, size_t min_len, void** data) {
  bool is_ta = false;
  napi_typedarray_type type;
  size_t length, offset;
  napi_value ab;
  *data = NULL;
  napi_is_typedarray(env, v, &is_ta);
  if (!is_ta || napi_get_typedarray_info(env, v, &type, &length, data, &ab, &offset) != napi_ok || type != want ||
      length < min_len) {
    napi_throw_type_error(env, NULL, "typed array of the wrong type or length");
    return -1;
  }
  return 0;
}

static size_t grid_nodes_of(mpm_engine* e) {
  int32_t N = 0;
  mpm_debug_p2g(e, 0.0, NULL, NULL, &N);
  return (size_t)N * N * N;
}

/* stepToGrid(h, dt) */
static napi_value js_step_to_grid(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm_engine* e = argc >= 2 ? get_engine(env, argv[0]) : NULL;
  if (!e) return NULL;
  double dt = 0;
  napi_get_value_double(env, argv[1], &dt);
  int rc = mpm_step_to_grid(e, dt);
  if (rc) return throw_mpm(env, rc, e);
  return NULL;
}

/* gridDownload(h, mass Float64Array N^3, v Float64Array 3N^3, fixed Uint8Array N^3) */
static napi_value js_grid_download(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm_engine* e = argc >= 4 ? get_engine(env, argv[0]) : NULL;
  if (!e) return NULL;
  const size_t nn = grid_nodes_of(e);
  void *m, *v, *f;
  if (typed(env, argv[1], napi_float64_array, nn, &m) || typed(env, argv[2], napi_float64_array, 3 * nn, &v) ||
      typed(env, argv[3], napi_uint8_array, nn, &f))
    return NULL;
  int rc = mpm_grid_download(e, (double*)m, (double*)v, (uint8_t*)f);
  if (rc) return throw_mpm(env, rc, e);
  return NULL;
}

/* gridUploadVelocity(h, v Float64Array 3N^3) */
static napi_value js_grid_upload_velocity(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm_engine* e = argc >= 2 ? get_engine(env, argv[0]) : NULL;
  if (!e) return NULL;
  void* v;
  if (typed(env, argv[1], napi_float64_array, 3 * grid_nodes_of(e), &v)) return NULL;
  int rc = mpm_grid_upload_velocity(e, (const double*)v);
  if (rc) return throw_mpm(env, rc, e);
  return NULL;
}

/* stepFromGrid(h, dt, slots Int32Array|null, deltas Float64Array|null, nslots) */
static napi_value js_step_from_grid(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value argv[5];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm_engine* e = argc >= 2 ? get_engine(env, argv[0]) : NULL;
  if (!e) return NULL;
  double dt = 0;
  int32_t nslots = 0;
  napi_get_value_double(env, argv[1], &dt);
  if (argc >= 5) napi_get_value_int32(env, argv[4], &nslots);
  void *slots = NULL, *deltas = NULL;
  if (nslots > 0) {
    if (typed(env, argv[2], napi_int32_array, 64, &slots) ||
        typed(env, argv[3], napi_float64_array, (size_t)nslots * 3 * grid_nodes_of(e), &deltas))
      return NULL;
  }
  int rc = mpm_step_from_grid(e, dt, (const int32_t*)slots, nslots > 0 ? 64 : 0, (const double*)deltas, nslots);
  if (rc) return throw_mpm(env, rc, e);
  return NULL;
}

/* ------------------------------------------------------------------ 2D */
static void finalize_engine2(napi_env env, void* data, void* hint) {
  (void)env;
  (void)hint;
  mpm2_engine** slot = (mpm2_engine**)data;
  if (*slot) mpm2_destroy(*slot);
  free(slot);
}

static mpm2_engine* get_engine2(napi_env env, napi_value v) {
  void* data = NULL;
  if (napi_get_value_external(env, v, &data) != napi_ok || !data) {
    napi_throw_type_error(env, NULL, "expected a 2D engine handle");
    return NULL;
  }
  mpm2_engine* e = *(mpm2_engine**)data;
  if (!e) napi_throw_error(env, NULL, "engine was destroyed");
  return e;
}

static napi_value throw_mpm2(napi_env env, int rc, mpm2_engine* e) {
  char code[16], msg[600];
  snprintf(code, sizeof code, "MPM%d", rc);
  snprintf(msg, sizeof msg, "mpm error %d: %s", rc, e ? mpm2_last_error(e) : "");
  napi_throw_error(env, code, msg);
  return NULL;
}

/* create2(capacity) */
static napi_value js_create2(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  double cap = 0;
  if (argc < 1 || napi_get_value_double(env, argv[0], &cap) != napi_ok || !(cap >= 1) || cap > 4294967295.0) {
    napi_throw_range_error(env, NULL, "capacity must be a positive integer");
    return NULL;
  }
  mpm2_engine* e = NULL;
  int rc = mpm2_create((uint32_t)cap, &e);
  if (rc) return throw_mpm2(env, rc, NULL);
  mpm2_engine** slot = (mpm2_engine**)malloc(sizeof *slot);
  *slot = e;
  napi_value ext;
  CHECK_NAPI(env, napi_create_external(env, slot, finalize_engine2, NULL, &ext));
  return ext;
}

static napi_value js_destroy2(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  void* data = NULL;
  if (argc < 1 || napi_get_value_external(env, argv[0], &data) != napi_ok || !data) return NULL;
  mpm2_engine** slot = (mpm2_engine**)data;
  if (*slot) mpm2_destroy(*slot);
  *slot = NULL;
  return NULL;
}

/* {count, x, y, vx, vy, F, C, mass, volume, pinned}: floats all Float64Array
 * or all Float32Array (F and C 4 per particle), pinned a Uint8Array. */
static int particle_view2(napi_env env, napi_value obj, mpm2_particles* p) {
  memset(p, 0, sizeof *p);
  p->count = (uint32_t)get_num(env, obj, "count", 0);
  p->dtype = -1;
  struct {
    const char* key;
    void** dst;
    int per, flag;
  } fields[] = {{"x", &p->x, 1, 0},       {"y", &p->y, 1, 0},       {"vx", &p->vx, 1, 0},
                {"vy", &p->vy, 1, 0},     {"F", &p->F, 4, 0},       {"C", &p->C, 4, 0},
                {"mass", &p->mass, 1, 0}, {"volume", &p->volume, 1, 0}, {"pinned", (void**)&p->pinned, 1, 1}};
  for (size_t k = 0; k < sizeof fields / sizeof fields[0]; ++k) {
    bool has = false, is_ta = false;
    napi_value v, ab;
    if (napi_has_named_property(env, obj, fields[k].key, &has) != napi_ok || !has) continue;
    napi_get_named_property(env, obj, fields[k].key, &v);
    napi_is_typedarray(env, v, &is_ta);
    if (!is_ta) continue;
    napi_typedarray_type type;
    size_t length, offset;
    void* data;
    napi_get_typedarray_info(env, v, &type, &length, &data, &ab, &offset);
    if (length < (size_t)p->count * fields[k].per) {
      napi_throw_range_error(env, NULL, fields[k].key);
      return -1;
    }
    if (fields[k].flag) {
      if (type != napi_uint8_array) {
        napi_throw_type_error(env, NULL, "pinned must be a Uint8Array");
        return -1;
      }
    } else {
      const int dt = type == napi_float64_array ? MPM_F64 : type == napi_float32_array ? MPM_F32 : -2;
      if (dt < 0 || (p->dtype >= 0 && p->dtype != dt)) {
        napi_throw_type_error(env, NULL, "floating arrays must all be Float64Array or all Float32Array");
        return -1;
      }
      p->dtype = dt;
    }
    *fields[k].dst = data;
  }
  if (p->dtype < 0) p->dtype = MPM_F64;
  return 0;
}

/* settings object {stiffness, walls, bound, floor, top} */
static void settings2(napi_env env, napi_value obj, mpm2_settings* s) {
  mpm2_settings_defaults(NULL, s);
  s->stiffness = (float)get_num(env, obj, "stiffness", s->stiffness);
  s->walls = (int32_t)get_num(env, obj, "walls", s->walls);
  s->bound = (float)get_num(env, obj, "bound", s->bound);
  s->floor = (float)get_num(env, obj, "floor", s->floor);
  s->top = (float)get_num(env, obj, "top", s->top);
}

static napi_value js_upload2(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm2_engine* e = argc >= 2 ? get_engine2(env, argv[0]) : NULL;
  if (!e) return NULL;
  mpm2_particles p;
  if (particle_view2(env, argv[1], &p)) return NULL;
  int rc = mpm2_upload(e, &p);
  if (rc) return throw_mpm2(env, rc, e);
  return NULL;
}

static napi_value js_download2(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm2_engine* e = argc >= 2 ? get_engine2(env, argv[0]) : NULL;
  if (!e) return NULL;
  mpm2_particles p;
  if (particle_view2(env, argv[1], &p)) return NULL;
  int rc = mpm2_download(e, &p);
  if (rc) return throw_mpm2(env, rc, e);
  return NULL;
}

/* step2(h, settings, dt, n) */
static napi_value js_step2(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm2_engine* e = argc >= 3 ? get_engine2(env, argv[0]) : NULL;
  if (!e) return NULL;
  mpm2_settings s;
  settings2(env, argv[1], &s);
  double dt = 0;
  int32_t n = 1;
  napi_get_value_double(env, argv[2], &dt);
  if (argc >= 4) napi_get_value_int32(env, argv[3], &n);
  int rc = mpm2_step(e, &s, dt, n);
  if (!rc) rc = mpm2_synchronize(e);
  if (rc) return throw_mpm2(env, rc, e);
  return NULL;
}

/* stableDt2(h, settings) */
static napi_value js_stable_dt2(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm2_engine* e = argc >= 2 ? get_engine2(env, argv[0]) : NULL;
  if (!e) return NULL;
  mpm2_settings s;
  settings2(env, argv[1], &s);
  double out = 0;
  int rc = mpm2_stable_dt(e, &s, &out);
  if (rc) return throw_mpm2(env, rc, e);
  napi_value v;
  napi_create_double(env, out, &v);
  return v;
}

/* advance2(h, settings, dt, gravityMode, strength, maxSubsteps) -> {substeps, advanced} */
static napi_value js_advance2(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value argv[6];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm2_engine* e = argc >= 3 ? get_engine2(env, argv[0]) : NULL;
  if (!e) return NULL;
  mpm2_settings s;
  settings2(env, argv[1], &s);
  double dt = 0, strength = 1;
  int32_t mode = 0, maxs = 96;
  napi_get_value_double(env, argv[2], &dt);
  if (argc >= 4) napi_get_value_int32(env, argv[3], &mode);
  if (argc >= 5) napi_get_value_double(env, argv[4], &strength);
  if (argc >= 6) napi_get_value_int32(env, argv[5], &maxs);
  int32_t n = 0;
  double adv = 0;
  int rc = mpm2_advance(e, &s, dt, mode, (float)strength, maxs, &n, &adv);
  if (rc) return throw_mpm2(env, rc, e);
  napi_value o;
  CHECK_NAPI(env, napi_create_object(env, &o));
  set_num(env, o, "substeps", n);
  set_num(env, o, "advanced", adv);
  return o;
}

/* stats2(h) -> {count, minJ, maxJ, strain, nodes, substeps, gridMassQ, particleMassQ} */
static napi_value js_stats2(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm2_engine* e = argc >= 1 ? get_engine2(env, argv[0]) : NULL;
  if (!e) return NULL;
  mpm2_stats st;
  int rc = mpm2_statistics(e, &st);
  if (rc) return throw_mpm2(env, rc, e);
  napi_value o, b;
  CHECK_NAPI(env, napi_create_object(env, &o));
  set_num(env, o, "count", st.count);
  set_num(env, o, "minJ", st.min_J);
  set_num(env, o, "maxJ", st.max_J);
  set_num(env, o, "strain", st.strain);
  set_num(env, o, "nodes", st.nodes);
  set_num(env, o, "substeps", st.substeps);
  char buf[32];
  snprintf(buf, sizeof buf, "%lld", (long long)st.grid_mass_q);
  napi_create_string_utf8(env, buf, NAPI_AUTO_LENGTH, &b);
  napi_set_named_property(env, o, "gridMassQ", b);
  snprintf(buf, sizeof buf, "%lld", (long long)st.particle_mass_q);
  napi_create_string_utf8(env, buf, NAPI_AUTO_LENGTH, &b);
  napi_set_named_property(env, o, "particleMassQ", b);
  return o;
}

static napi_value js_refresh_diagnostics(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm_engine* e = argc >= 2 ? get_engine(env, argv[0]) : NULL;
  if (!e) return NULL;
  napi_typedarray_type type;
  size_t length, offset;
  void* data;
  napi_value ab;
  bool is_ta = false;
  napi_is_typedarray(env, argv[1], &is_ta);
  if (!is_ta || napi_get_typedarray_info(env, argv[1], &type, &length, &data, &ab, &offset) != napi_ok ||
      type != napi_float64_array) {
    napi_throw_type_error(env, NULL, "refreshDiagnostics needs a Float64Array");
    return NULL;
  }
  uint32_t n = 0;
  mpm_count(e, &n);
  if (length < n) {
    napi_throw_range_error(env, NULL, "stress array too short");
    return NULL;
  }
  int rc = mpm_refresh_diagnostics(e, (double*)data);
  if (rc) return throw_mpm(env, rc, e);
  return NULL;
}

/* stressTensors(engine, tau Float64Array [9 count]): MPM3.tau of the resident
 * state (mpm_stress_tensors). */
static napi_value js_stress_tensors(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm_engine* e = argc >= 2 ? get_engine(env, argv[0]) : NULL;
  if (!e) return NULL;
  uint32_t n = 0;
  mpm_count(e, &n);
  void* data = NULL;
  if (typed(env, argv[1], napi_float64_array, 9 * (size_t)n, &data)) return NULL;
  int rc = mpm_stress_tensors(e, (double*)data);
  if (rc) return throw_mpm(env, rc, e);
  return NULL;
}

/* checkState(engine) -> 0, 7 or 8: waits for the engine and returns the state
 * guard's finding as the GPU service's status code (mpm_synchronize); other
 * failures throw. */
static napi_value js_check_state(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1], r;
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm_engine* e = argc >= 1 ? get_engine(env, argv[0]) : NULL;
  if (!e) return NULL;
  const int rc = mpm_synchronize(e);
  if (rc && rc != MPM_ERR_CFL && rc != MPM_ERR_INVALID_STATE) return throw_mpm(env, rc, e);
  CHECK_NAPI(env, napi_create_int32(env, rc, &r));
  return r;
}

/* profile(engine, enable, reset): per-kernel timing on or off (mpm_profile). */
static napi_value js_profile(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm_engine* e = argc >= 1 ? get_engine(env, argv[0]) : NULL;
  if (!e) return NULL;
  int32_t enable = 1, reset = 1;
  if (argc >= 2) napi_get_value_int32(env, argv[1], &enable);
  if (argc >= 3) napi_get_value_int32(env, argv[2], &reset);
  int rc = mpm_profile(e, enable, reset);
  if (rc) return throw_mpm(env, rc, e);
  return NULL;
}

/* kernelTimes(engine) -> {name: milliseconds} accumulated while profiling. */
static napi_value js_kernel_times(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1], out;
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm_engine* e = argc >= 1 ? get_engine(env, argv[0]) : NULL;
  if (!e) return NULL;
  mpm_kernel_time t[32];
  int32_t n = 0;
  int rc = mpm_kernel_times(e, t, 32, &n);
  if (rc) return throw_mpm(env, rc, e);
  CHECK_NAPI(env, napi_create_object(env, &out));
  for (int32_t k = 0; k < n && k < 32; ++k) set_num(env, out, t[k].name, t[k].total_ms);
  return out;
}

static napi_value js_last_error(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1], r;
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mpm_engine* e = argc >= 1 ? get_engine(env, argv[0]) : NULL;
  if (!e) return NULL;
  CHECK_NAPI(env, napi_create_string_utf8(env, mpm_last_error(e), NAPI_AUTO_LENGTH, &r));
  return r;
}

static napi_value init(napi_env env, napi_value exports) {
  struct {
    const char* name;
    napi_callback fn;
  } fns[] = {{"create", js_create},
             {"destroy", js_destroy},
             {"setMaterials", js_set_materials},
             {"upload", js_upload},
             {"download", js_download},
             {"step", js_step},
             {"advanceFrame", js_advance_frame},
             {"synchronize", js_synchronize},
             {"stableDt", js_stable_dt},
             {"conservation", js_conservation},
             {"refreshDiagnostics", js_refresh_diagnostics},
             {"stressTensors", js_stress_tensors},
             {"checkState", js_check_state},
             {"profile", js_profile},
             {"kernelTimes", js_kernel_times},
             {"snapshot", js_snapshot},
             {"framePacket", js_frame_packet},
             {"stepToGrid", js_step_to_grid},
             {"gridDownload", js_grid_download},
             {"gridUploadVelocity", js_grid_upload_velocity},
             {"stepFromGrid", js_step_from_grid},
             {"create2", js_create2},
             {"destroy2", js_destroy2},
             {"upload2", js_upload2},
             {"download2", js_download2},
             {"step2", js_step2},
             {"stableDt2", js_stable_dt2},
             {"advance2", js_advance2},
             {"stats2", js_stats2},
             {"lastError", js_last_error}};
  for (size_t i = 0; i < sizeof fns / sizeof fns[0]; ++i) {
    napi_value f;
    CHECK_NAPI(env, napi_create_function(env, fns[i].name, NAPI_AUTO_LENGTH, fns[i].fn, NULL, &f));
    CHECK_NAPI(env, napi_set_named_property(env, exports, fns[i].name, f));
  }
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
