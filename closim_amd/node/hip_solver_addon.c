/* hip_solver_addon.c — Node-API binding of the cloth step's C ABI
 * (include/cloth_hip.h), the native half of the JS `HipSolver`
 * (hip-solver.js) that stands behind the reference's ClothSolver interface
 * (src/backend/solver.ts:5-13) in place of CpuSolver
 * (src/backend/cpu-solver.ts:7-49).
 *
 * Every entry checks its arguments (types, typed-array kinds and lengths)
 * before it reaches the library and throws a JS TypeError on a mismatch; a
 * non-zero status from the library becomes `Error(cs_last_error())`, as the
 * reference throws `Error`s.  Solver handles are externals whose finalizer
 * calls cs_destroy. */
#include <node_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cloth_hip.h"

#define CALL(env, expr)                                                   \
  do {                                                                    \
    if ((expr) != napi_ok) {                                              \
      const napi_extended_error_info* ei_ = NULL;                         \
      napi_get_last_error_info((env), &ei_);                              \
      napi_throw_error((env), NULL, ei_ && ei_->error_message ? ei_->error_message : "napi call failed"); \
      return NULL;                                                        \
    }                                                                     \
  } while (0)

static napi_value throw_status(napi_env env) {
  napi_throw_error(env, NULL, cs_last_error());
  return NULL;
}

static napi_value throw_type(napi_env env, const char* what, const char* name) {
  char buf[256];
  snprintf(buf, sizeof buf, "%s: %s", name, what);
  napi_throw_type_error(env, NULL, buf);
  return NULL;
}

static int is_nullish(napi_env env, napi_value v) {
  napi_valuetype t;
  if (napi_typeof(env, v, &t) != napi_ok) return 1;
  return t == napi_undefined || t == napi_null;
}

static const char* kind_name(napi_typedarray_type t) {
  switch (t) {
    case napi_float64_array: return "Float64Array";
    case napi_float32_array: return "Float32Array";
    case napi_uint32_array: return "Uint32Array";
    case napi_int32_array: return "Int32Array";
    default: return "typed array";
  }
}

/* A typed array of kind `want` and exactly `want_len` elements (any length
 * when want_len == (size_t)-1); 0 on success, -1 after throwing. */
static int typed_of(napi_env env, napi_value v, const char* name, napi_typedarray_type want, size_t want_len,
                    void** data, size_t* len) {
  bool is = false;
  if (napi_is_typedarray(env, v, &is) != napi_ok || !is) {
    char msg[96];
    snprintf(msg, sizeof msg, "expected a %s", kind_name(want));
    throw_type(env, msg, name);
    return -1;
  }
  napi_typedarray_type t;
  size_t n = 0;
  void* p = NULL;
  if (napi_get_typedarray_info(env, v, &t, &n, &p, NULL, NULL) != napi_ok || t != want) {
    char msg[96];
    snprintf(msg, sizeof msg, "expected a %s", kind_name(want));
    throw_type(env, msg, name);
    return -1;
  }
  if (want_len != (size_t)-1 && n != want_len) {
    char msg[128];
    snprintf(msg, sizeof msg, "length %zu, expected %zu", n, want_len);
    throw_type(env, msg, name);
    return -1;
  }
  *data = p;
  if (len) *len = n;
  return 0;
}

static int prop(napi_env env, napi_value obj, const char* name, napi_value* out) {
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) {
    throw_type(env, "missing", name);
    return -1;
  }
  return napi_get_named_property(env, obj, name, out) == napi_ok ? 0 : -1;
}

static int opt_prop(napi_env env, napi_value obj, const char* name, napi_value* out) {
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return 1;
  if (napi_get_named_property(env, obj, name, out) != napi_ok) return -1;
  return is_nullish(env, *out) ? 1 : 0;
}

static int num_of(napi_env env, napi_value v, const char* name, double* out) {
  napi_valuetype t;
  if (napi_typeof(env, v, &t) != napi_ok || t != napi_number) {
    throw_type(env, "expected a number", name);
    return -1;
  }
  return napi_get_value_double(env, v, out) == napi_ok ? 0 : -1;
}

static int num_prop(napi_env env, napi_value obj, const char* name, double* out) {
  napi_value v;
  if (prop(env, obj, name, &v)) return -1;
  return num_of(env, v, name, out);
}

static int u32_prop(napi_env env, napi_value obj, const char* name, uint32_t* out) {
  double d;
  if (num_prop(env, obj, name, &d)) return -1;
  if (!(d >= 0) || d > 4294967295.0 || d != (double)(uint32_t)d) {
    throw_type(env, "expected a non-negative integer", name);
    return -1;
  }
  *out = (uint32_t)d;
  return 0;
}

static int material_of(napi_env env, napi_value m, cs_material* out) {
  static const char* names[9] = {"arealDensityKgM2", "thickness", "stretchWarp", "stretchWeft", "stretchCoupling",
                                 "shear",           "bendWarp",  "bendWeft",    "damping"};
  double* dst[9] = {&out->areal_density, &out->thickness, &out->stretch_warp, &out->stretch_weft,
                    &out->stretch_coupling, &out->shear, &out->bend_warp, &out->bend_weft, &out->damping};
  napi_valuetype t;
  if (napi_typeof(env, m, &t) != napi_ok || t != napi_object) {
    throw_type(env, "expected an object", "material");
    return -1;
  }
  for (int k = 0; k < 9; k++)
    if (num_prop(env, m, names[k], dst[k])) return -1;
  return 0;
}

/* The packed scene (see packScene in hip-solver.js) into a cs_scene_desc whose
 * pointers alias the JS typed arrays (valid for the duration of the call). */
static int scene_of(napi_env env, napi_value o, cs_scene_desc* d) {
  memset(d, 0, sizeof *d);
  napi_valuetype t;
  if (napi_typeof(env, o, &t) != napi_ok || t != napi_object) {
    throw_type(env, "expected an object", "scene");
    return -1;
  }
  if (u32_prop(env, o, "nVerts", &d->n_verts) || u32_prop(env, o, "nTris", &d->n_tris) ||
      u32_prop(env, o, "nHinges", &d->n_hinges))
    return -1;
  const size_t n = d->n_verts, m = d->n_tris, h = d->n_hinges;
  napi_value v;
  void* p;
#define TYPED(field, name, kind, count, type)                        \
  if (prop(env, o, name, &v) || typed_of(env, v, name, kind, count, &p, NULL)) return -1; \
  d->field = (type)p;
  TYPED(positions, "positions", napi_float64_array, 3 * n, const double*);
  TYPED(velocities, "velocities", napi_float64_array, 3 * n, const double*);
  TYPED(indices, "indices", napi_uint32_array, 3 * m, const uint32_t*);
  TYPED(inv_dm, "invDm", napi_float32_array, 4 * m, const float*);
  TYPED(areas, "areas", napi_float32_array, m, const float*);
  TYPED(masses, "masses", napi_float32_array, n, const float*);
  TYPED(hinge_verts, "hingeVerts", napi_uint32_array, 4 * h, const uint32_t*);
  TYPED(hinge_rest_angle, "hingeRestAngle", napi_float64_array, h, const double*);
  TYPED(hinge_edge_len, "hingeEdgeLen", napi_float64_array, h, const double*);
  TYPED(hinge_area_sum, "hingeAreaSum", napi_float64_array, h, const double*);
#undef TYPED
  int r = opt_prop(env, o, "restPositions", &v);
  if (r < 0) return -1;
  if (r == 0) {
    if (typed_of(env, v, "restPositions", napi_float32_array, 3 * n, &p, NULL)) return -1;
    d->rest_positions = (const float*)p;
  }
  if (prop(env, o, "material", &v) || material_of(env, v, &d->material)) return -1;
  if (prop(env, o, "gravity", &v)) return -1;
  for (uint32_t k = 0; k < 3; k++) {
    napi_value e;
    if (napi_get_element(env, v, k, &e) != napi_ok || num_of(env, e, "gravity", &d->gravity[k])) return -1;
  }
  size_t np = 0;
  if (prop(env, o, "pinIds", &v) || typed_of(env, v, "pinIds", napi_uint32_array, (size_t)-1, &p, &np)) return -1;
  d->pin_ids = (const uint32_t*)p;
  d->n_pins = (uint32_t)np;
  if (prop(env, o, "pinPos", &v) || typed_of(env, v, "pinPos", napi_float64_array, 3 * np, &p, NULL)) return -1;
  d->pin_pos = (const double*)p;
  napi_value c;
  r = opt_prop(env, o, "contact", &c);
  if (r < 0) return -1;
  if (r == 0) {
    d->contact_enabled = 1;
    napi_value pr;
    if (prop(env, c, "params", &pr)) return -1;
    if (num_prop(env, pr, "dHatM", &d->contact.d_hat) || num_prop(env, pr, "kappaJ", &d->contact.kappa) ||
        num_prop(env, pr, "dMinM", &d->contact.d_min) || num_prop(env, pr, "frictionMu", &d->contact.friction_mu) ||
        num_prop(env, pr, "frictionEpsM", &d->contact.friction_eps))
      return -1;
    r = opt_prop(env, c, "floorY", &v);
    if (r < 0) return -1;
    if (r == 0) {
      d->has_floor = 1;
      if (num_of(env, v, "floorY", &d->floor_y)) return -1;
    }
    napi_value sp, si;
    const int rp = opt_prop(env, c, "staticPos", &sp), ri = opt_prop(env, c, "staticIdx", &si);
    if (rp < 0 || ri < 0) return -1;
    if (rp == 0 && ri == 0) {
      size_t ns = 0, ms = 0;
      if (typed_of(env, sp, "staticPos", napi_float32_array, (size_t)-1, &p, &ns)) return -1;
      d->static_pos = (const float*)p;
      if (typed_of(env, si, "staticIdx", napi_uint32_array, (size_t)-1, &p, &ms)) return -1;
      d->static_idx = (const uint32_t*)p;
      if (ns % 3 || ms % 3) {
        throw_type(env, "static mesh arrays must hold whole vertices and triangles", "contact");
        return -1;
      }
      d->n_static_verts = (uint32_t)(ns / 3);
      d->n_static_tris = (uint32_t)(ms / 3);
    }
  }
  return 0;
}

static int str_eq(napi_env env, napi_value v, const char* s) {
  char buf[32];
  size_t len = 0;
  if (napi_get_value_string_utf8(env, v, buf, sizeof buf, &len) != napi_ok) return 0;
  return strcmp(buf, s) == 0;
}

/* { precision: "fp64" | "mixed", preconditioner: "jacobi" | "block-jacobi" |
 *   "schwarz", schwarz: [target, min, max], hvp: "fd" | "analytic",
 *   coarse: "none" | "mas2" | "c0" | "c1", coarseIters, coarseOmega,
 *   newtonIters, pcgMaxIters, lineSearchIters, gradTol, flags, device } */
static int options_of(napi_env env, napi_value o, cs_options* opt) {
  cs_options_default(opt);
  if (is_nullish(env, o)) return 0;
  napi_value v;
  int r;
  if ((r = opt_prop(env, o, "precision", &v)) < 0) return -1;
  if (r == 0) {
    if (str_eq(env, v, "mixed")) opt->precision = CS_PREC_MIXED;
    else if (str_eq(env, v, "fp64")) opt->precision = CS_PREC_FP64;
    else { throw_type(env, "expected \"fp64\" or \"mixed\"", "precision"); return -1; }
  }
  if ((r = opt_prop(env, o, "preconditioner", &v)) < 0) return -1;
  if (r == 0) {
    if (str_eq(env, v, "jacobi")) opt->precond = CS_PRECOND_JACOBI;
    else if (str_eq(env, v, "block-jacobi")) opt->precond = CS_PRECOND_BLOCK_JACOBI;
    else if (str_eq(env, v, "schwarz")) opt->precond = CS_PRECOND_SCHWARZ;
    else { throw_type(env, "expected \"jacobi\", \"block-jacobi\" or \"schwarz\"", "preconditioner"); return -1; }
  }
  if ((r = opt_prop(env, o, "schwarz", &v)) < 0) return -1;
  if (r == 0) {
    int32_t* dst[3] = {&opt->schwarz_target, &opt->schwarz_min, &opt->schwarz_max};
    for (uint32_t k = 0; k < 3; k++) {
      napi_value e;
      double x;
      if (napi_get_element(env, v, k, &e) != napi_ok || num_of(env, e, "schwarz", &x)) return -1;
      *dst[k] = (int32_t)x;
    }
  }
  if ((r = opt_prop(env, o, "hvp", &v)) < 0) return -1;
  if (r == 0) {
    if (str_eq(env, v, "analytic")) opt->hvp_mode = CS_HVP_ANALYTIC;
    else if (str_eq(env, v, "fd")) opt->hvp_mode = CS_HVP_FD;
    else { throw_type(env, "expected \"fd\" or \"analytic\"", "hvp"); return -1; }
  }
  static const char* ints[6] = {"newtonIters", "pcgMaxIters", "lineSearchIters", "flags", "device", NULL};
  int32_t* idst[5] = {&opt->newton_iters, &opt->pcg_max_iters, &opt->line_search_iters, NULL, &opt->device};
  for (int k = 0; ints[k]; k++) {
    if ((r = opt_prop(env, o, ints[k], &v)) < 0) return -1;
    if (r) continue;
    double x;
    if (num_of(env, v, ints[k], &x)) return -1;
    if (k == 3) opt->flags = (uint32_t)x;
    else *idst[k] = (int32_t)x;
  }
  if ((r = opt_prop(env, o, "gradTol", &v)) < 0) return -1;
  if (r == 0 && num_of(env, v, "gradTol", &opt->grad_tol)) return -1;
  if ((r = opt_prop(env, o, "coarse", &v)) < 0) return -1;
  if (r == 0) {
    if (str_eq(env, v, "none")) opt->coarse = CS_COARSE_NONE;
    else if (str_eq(env, v, "mas2")) opt->coarse = CS_COARSE_JACOBI;
    else if (str_eq(env, v, "c0")) opt->coarse = CS_COARSE_BLOCK;
    else if (str_eq(env, v, "c1")) opt->coarse = CS_COARSE_PCG;
    else { throw_type(env, "expected \"none\", \"mas2\", \"c0\" or \"c1\"", "coarse"); return -1; }
  }
  if ((r = opt_prop(env, o, "coarseIters", &v)) < 0) return -1;
  if (r == 0) {
    double x;
    if (num_of(env, v, "coarseIters", &x)) return -1;
    opt->coarse_iters = (int32_t)x;
  }
  if ((r = opt_prop(env, o, "coarseOmega", &v)) < 0) return -1;
  if (r == 0 && num_of(env, v, "coarseOmega", &opt->coarse_omega)) return -1;
  return 0;
}

static void finalize_solver(napi_env env, void* data, void* hint) {
  (void)env;
  (void)hint;
  cs_destroy((cs_solver*)data);
}

/* The solver behind handle argv[0]; NULL after throwing (destroyed handles
 * hold NULL, as an uninitialized CpuSolver does). */
static cs_solver* handle_of(napi_env env, napi_value v) {
  napi_valuetype t;
  void* p = NULL;
  if (napi_typeof(env, v, &t) != napi_ok || t != napi_object) {
    throw_type(env, "expected a solver handle", "handle");
    return NULL;
  }
  if (napi_unwrap(env, v, &p) != napi_ok || !p) {
    napi_throw_error(env, NULL, "solver handle not initialized or already destroyed");
    return NULL;
  }
  return (cs_solver*)p;
}

static napi_value args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  napi_value self;
  if (napi_get_cb_info(env, info, &argc, argv, &self, NULL) != napi_ok) return NULL;
  for (size_t k = argc; k < want; k++) napi_get_undefined(env, &argv[k]);
  return self;
}

/* create(scene, options) -> handle */
static napi_value Create(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  cs_scene_desc d;
  cs_options opt;
  if (scene_of(env, argv[0], &d) || options_of(env, argv[1], &opt)) return NULL;
  cs_solver* h = NULL;
  if (cs_create(&d, &opt, &h) != CS_OK) return throw_status(env);
  napi_value obj;
  CALL(env, napi_create_object(env, &obj));
  if (napi_wrap(env, obj, h, finalize_solver, NULL, NULL) != napi_ok) {
    cs_destroy(h);
    napi_throw_error(env, NULL, "could not wrap the solver handle");
    return NULL;
  }
  return obj;
}

/* destroy(handle): releases the device state now */
static napi_value Destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  void* p = NULL;
  if (napi_remove_wrap(env, argv[0], &p) == napi_ok && p) cs_destroy((cs_solver*)p);
  return NULL;
}

static int out3n(napi_env env, cs_solver* h, napi_value v, const char* name, double** p, size_t* n) {
  (void)h;
  void* q;
  if (typed_of(env, v, name, napi_float64_array, (size_t)-1, &q, n)) return -1;
  *p = (double*)q;
  return 0;
}

/* step(handle, dt, positionsOut, velocitiesOut) */
static napi_value Step(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  args(env, info, 4, argv);
  cs_solver* h = handle_of(env, argv[0]);
  if (!h) return NULL;
  double dt;
  if (num_of(env, argv[1], "dt", &dt)) return NULL;
  double *px, *pv;
  size_t nx, nv;
  if (out3n(env, h, argv[2], "positions", &px, &nx) || out3n(env, h, argv[3], "velocities", &pv, &nv)) return NULL;
  if (cs_step(h, dt) != CS_OK) return throw_status(env);
  if (cs_get_positions(h, px, nx) != CS_OK || cs_get_velocities(h, pv, nv) != CS_OK) return throw_status(env);
  return NULL;
}

static napi_value SetMaterial(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  cs_solver* h = handle_of(env, argv[0]);
  if (!h) return NULL;
  cs_material m;
  if (material_of(env, argv[1], &m)) return NULL;
  if (cs_set_material(h, &m) != CS_OK) return throw_status(env);
  return NULL;
}

/* setStaticMesh(handle, positions: Float32Array, indices: Uint32Array):
 * ContactSystem.setStaticMesh between steps (a moving rigid collider) */
static napi_value SetStaticMesh(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  args(env, info, 3, argv);
  cs_solver* h = handle_of(env, argv[0]);
  if (!h) return NULL;
  void *pp, *pi;
  size_t np, ni;
  if (typed_of(env, argv[1], "staticPos", napi_float32_array, (size_t)-1, &pp, &np) ||
      typed_of(env, argv[2], "staticIdx", napi_uint32_array, (size_t)-1, &pi, &ni))
    return NULL;
  if (np % 3 || ni % 3) return throw_type(env, "static mesh arrays must hold whole vertices and triangles", "staticPos");
  if (cs_set_static_mesh(h, (const float*)pp, (uint32_t)(np / 3), (const uint32_t*)pi, (uint32_t)(ni / 3)) != CS_OK)
    return throw_status(env);
  return NULL;
}

/* pinVertex(handle, id, [x, y, z] | null, positionsOut) */
static napi_value PinVertex(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  args(env, info, 4, argv);
  cs_solver* h = handle_of(env, argv[0]);
  if (!h) return NULL;
  double id;
  if (num_of(env, argv[1], "vertexId", &id)) return NULL;
  if (!(id >= 0) || id != (double)(uint32_t)id) return throw_type(env, "expected a non-negative integer", "vertexId");
  double pos[3];
  const double* pp = NULL;
  if (!is_nullish(env, argv[2])) {
    for (uint32_t k = 0; k < 3; k++) {
      napi_value e;
      if (napi_get_element(env, argv[2], k, &e) != napi_ok || num_of(env, e, "position", &pos[k])) return NULL;
    }
    pp = pos;
  }
  if (cs_pin_vertex(h, (uint32_t)id, pp) != CS_OK) return throw_status(env);
  if (!is_nullish(env, argv[3])) {
    double* px;
    size_t nx;
    if (out3n(env, h, argv[3], "positions", &px, &nx)) return NULL;
    if (cs_get_positions(h, px, nx) != CS_OK) return throw_status(env);
  }
  return NULL;
}

static napi_value UnpinVertex(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  cs_solver* h = handle_of(env, argv[0]);
  if (!h) return NULL;
  double id;
  if (num_of(env, argv[1], "vertexId", &id)) return NULL;
  if (cs_unpin_vertex(h, (uint32_t)id) != CS_OK) return throw_status(env);
  return NULL;
}

/* getPositions / getVelocities (handle, out): copy-out into `out` */
static napi_value CopyOut(napi_env env, napi_callback_info info, int vel) {
  napi_value argv[2];
  args(env, info, 2, argv);
  cs_solver* h = handle_of(env, argv[0]);
  if (!h) return NULL;
  double* p;
  size_t n;
  if (out3n(env, h, argv[1], vel ? "velocities" : "positions", &p, &n)) return NULL;
  if ((vel ? cs_get_velocities(h, p, n) : cs_get_positions(h, p, n)) != CS_OK) return throw_status(env);
  return NULL;
}
static napi_value GetPositions(napi_env env, napi_callback_info info) { return CopyOut(env, info, 0); }
static napi_value GetVelocities(napi_env env, napi_callback_info info) { return CopyOut(env, info, 1); }

/* setState(handle, positions, velocities) */
static napi_value SetState(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  args(env, info, 3, argv);
  cs_solver* h = handle_of(env, argv[0]);
  if (!h) return NULL;
  double *px, *pv;
  size_t nx, nv;
  if (out3n(env, h, argv[1], "positions", &px, &nx) || out3n(env, h, argv[2], "velocities", &pv, &nv)) return NULL;
  if (nx != nv) return throw_type(env, "positions and velocities differ in length", "velocities");
  if (cs_set_state(h, px, pv, nx) != CS_OK) return throw_status(env);
  return NULL;
}

static int set_num(napi_env env, napi_value o, const char* k, double x) {
  napi_value v;
  return napi_create_double(env, x, &v) != napi_ok || napi_set_named_property(env, o, k, v) != napi_ok;
}
static int set_bool(napi_env env, napi_value o, const char* k, int x) {
  napi_value v;
  return napi_get_boolean(env, x != 0, &v) != napi_ok || napi_set_named_property(env, o, k, v) != napi_ok;
}

/* lastStats(handle) -> StepStats (newton.ts:10-18) + newtonUsed,
 * lineSearchTrials; null before the first step */
static napi_value LastStats(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  cs_solver* h = handle_of(env, argv[0]);
  if (!h) return NULL;
  cs_step_stats s;
  if (cs_last_stats(h, &s) != CS_OK) return throw_status(env);
  napi_value o, c;
  if (s.newton_iters == 0 && s.newton_used == 0) {
    CALL(env, napi_get_null(env, &o));
    return o;
  }
  CALL(env, napi_create_object(env, &o));
  if (set_num(env, o, "newtonIters", s.newton_iters) || set_num(env, o, "pcgIters", s.pcg_iters) ||
      set_num(env, o, "gradNorm", s.grad_norm) || set_num(env, o, "energy", s.energy) ||
      set_num(env, o, "maxStrain", s.max_strain) || set_bool(env, o, "converged", s.converged) ||
      set_num(env, o, "newtonUsed", s.newton_used) || set_num(env, o, "lineSearchTrials", s.line_search_trials))
    return NULL;
  if (!s.has_contact) {
    CALL(env, napi_get_null(env, &c));
  } else {
    CALL(env, napi_create_object(env, &c));
    const cs_contact_diag* d = &s.contact;
    if (set_num(env, c, "candidatePairs", (double)d->candidate_pairs) ||
        set_num(env, c, "activePairs", (double)d->active_pairs) || set_num(env, c, "vtPairs", (double)d->vt_pairs) ||
        set_num(env, c, "eePairs", (double)d->ee_pairs) || set_num(env, c, "floorPairs", (double)d->floor_pairs) ||
        set_num(env, c, "minDistance", d->min_distance) || set_num(env, c, "minTOI", d->min_toi) ||
        set_num(env, c, "barrierEnergy", d->barrier_energy) || set_num(env, c, "frictionWork", d->friction_work) ||
        set_num(env, c, "rejectedLineSearchSteps", (double)d->rejected_line_search_steps) ||
        set_num(env, c, "ccdFailures", (double)d->ccd_failures))
      return NULL;
  }
  CALL(env, napi_set_named_property(env, o, "contact", c));
  return o;
}

static napi_value u32_array(napi_env env, const uint32_t* src, size_t n) {
  napi_value ab, ta;
  void* p;
  if (napi_create_arraybuffer(env, n * 4, &p, &ab) != napi_ok) return NULL;
  if (n) memcpy(p, src, n * 4);
  if (napi_create_typedarray(env, napi_uint32_array, n, ab, 0, &ta) != napi_ok) return NULL;
  return ta;
}

/* schwarzPartition(scene, target, min, max) -> {domainOf, localOf, count}
 * (buildSchwarzDomains, src/solver/schwarz.ts:44-140; no device needed) */
static napi_value SchwarzPartition(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  args(env, info, 4, argv);
  cs_scene_desc d;
  if (scene_of(env, argv[0], &d)) return NULL;
  double t, mn, mx;
  if (num_of(env, argv[1], "target", &t) || num_of(env, argv[2], "min", &mn) || num_of(env, argv[3], "max", &mx))
    return NULL;
  uint32_t* dom = (uint32_t*)malloc((size_t)(d.n_verts ? d.n_verts : 1) * 4);
  uint32_t* loc = (uint32_t*)malloc((size_t)(d.n_verts ? d.n_verts : 1) * 4);
  uint32_t cnt = 0;
  if (!dom || !loc) {
    free(dom);
    free(loc);
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  const int rc = cs_schwarz_partition(&d, (int)t, (int)mn, (int)mx, dom, loc, &cnt);
  napi_value o = NULL, a, b, c;
  if (rc == CS_OK && napi_create_object(env, &o) == napi_ok && (a = u32_array(env, dom, d.n_verts)) &&
      (b = u32_array(env, loc, d.n_verts)) && napi_create_uint32(env, cnt, &c) == napi_ok) {
    napi_set_named_property(env, o, "domainOf", a);
    napi_set_named_property(env, o, "localOf", b);
    napi_set_named_property(env, o, "count", c);
  }
  free(dom);
  free(loc);
  if (rc != CS_OK) return throw_status(env);
  return o;
}

/* hvpTiles(scene) -> {tiles, hdr, loc, lidx, off, ent} (the mixed HVP's
 * vertex-owner tiles, cs_hvp_tiles; no device needed; tiles 0 when the mesh
 * is not tiled) */
static napi_value HvpTiles(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  cs_scene_desc d;
  if (scene_of(env, argv[0], &d)) return NULL;
  uint64_t sz[5] = {0, 0, 0, 0, 0};
  if (cs_hvp_tiles(&d, sz, NULL, NULL, NULL, NULL, NULL) != CS_OK) return throw_status(env);
  uint32_t* hdr = (uint32_t*)malloc((size_t)(sz[0] ? sz[0] : 1) * 32);
  uint32_t* loc = (uint32_t*)malloc((size_t)(sz[1] ? sz[1] : 1) * 4);
  uint32_t* lidx = (uint32_t*)malloc((size_t)(sz[2] ? sz[2] : 1) * 8);
  uint16_t* off = (uint16_t*)malloc((size_t)(sz[3] ? sz[3] : 1) * 2);
  uint16_t* ent = (uint16_t*)malloc((size_t)(sz[4] ? sz[4] : 1) * 2);
  int rc = CS_OK;
  if (!hdr || !loc || !lidx || !off || !ent) {
    free(hdr); free(loc); free(lidx); free(off); free(ent);
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  if (sz[0]) rc = cs_hvp_tiles(&d, sz, hdr, loc, lidx, off, ent);
  napi_value o = NULL, t, a, b, c, e, f;
  uint32_t* off32 = (uint32_t*)malloc((size_t)(sz[3] ? sz[3] : 1) * 4);
  uint32_t* ent32 = (uint32_t*)malloc((size_t)(sz[4] ? sz[4] : 1) * 4);
  if (off32 && ent32 && rc == CS_OK) {
    for (uint64_t k = 0; k < sz[3]; k++) off32[k] = off[k];
    for (uint64_t k = 0; k < sz[4]; k++) ent32[k] = ent[k];
    if (napi_create_object(env, &o) == napi_ok && napi_create_double(env, (double)sz[0], &t) == napi_ok &&
        (a = u32_array(env, hdr, (size_t)sz[0] * 8)) && (b = u32_array(env, loc, (size_t)sz[1])) &&
        (c = u32_array(env, lidx, (size_t)sz[2] * 2)) && (e = u32_array(env, off32, (size_t)sz[3])) &&
        (f = u32_array(env, ent32, (size_t)sz[4]))) {
      napi_set_named_property(env, o, "tiles", t);
      napi_set_named_property(env, o, "hdr", a);
      napi_set_named_property(env, o, "loc", b);
      napi_set_named_property(env, o, "lidx", c);
      napi_set_named_property(env, o, "off", e);
      napi_set_named_property(env, o, "ent", f);
    }
  }
  free(hdr); free(loc); free(lidx); free(off); free(ent); free(off32); free(ent32);
  if (rc != CS_OK) return throw_status(env);
  if (!o) {
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  return o;
}

static napi_value DeviceInfo(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  double dev = 0;
  if (!is_nullish(env, argv[0]) && num_of(env, argv[0], "device", &dev)) return NULL;
  char buf[256];
  if (cs_device_info((int)dev, buf, sizeof buf) != CS_OK) return throw_status(env);
  napi_value s;
  CALL(env, napi_create_string_utf8(env, buf, NAPI_AUTO_LENGTH, &s));
  return s;
}

static napi_value AbiVersion(napi_env env, napi_callback_info info) {
  (void)info;
  napi_value v;
  CALL(env, napi_create_int32(env, cs_abi_version(), &v));
  return v;
}

static napi_value Init(napi_env env, napi_value exports) {
  const napi_property_descriptor fns[] = {
      {"create", NULL, Create, NULL, NULL, NULL, napi_enumerable, NULL},
      {"destroy", NULL, Destroy, NULL, NULL, NULL, napi_enumerable, NULL},
      {"step", NULL, Step, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setMaterial", NULL, SetMaterial, NULL, NULL, NULL, napi_enumerable, NULL},
      {"pinVertex", NULL, PinVertex, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setStaticMesh", NULL, SetStaticMesh, NULL, NULL, NULL, napi_enumerable, NULL},
      {"unpinVertex", NULL, UnpinVertex, NULL, NULL, NULL, napi_enumerable, NULL},
      {"getPositions", NULL, GetPositions, NULL, NULL, NULL, napi_enumerable, NULL},
      {"getVelocities", NULL, GetVelocities, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setState", NULL, SetState, NULL, NULL, NULL, napi_enumerable, NULL},
      {"lastStats", NULL, LastStats, NULL, NULL, NULL, napi_enumerable, NULL},
      {"schwarzPartition", NULL, SchwarzPartition, NULL, NULL, NULL, napi_enumerable, NULL},
      {"hvpTiles", NULL, HvpTiles, NULL, NULL, NULL, napi_enumerable, NULL},
      {"deviceInfo", NULL, DeviceInfo, NULL, NULL, NULL, napi_enumerable, NULL},
      {"abiVersion", NULL, AbiVersion, NULL, NULL, NULL, napi_enumerable, NULL},
  };
  if (napi_define_properties(env, exports, sizeof fns / sizeof fns[0], fns) != napi_ok) return NULL;
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
