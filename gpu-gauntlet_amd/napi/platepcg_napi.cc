// Thin Node N-API addon over the platepcg C ABI.
//
// Exposes the context / plan / solve entry points to JavaScript.  TypedArray
// arguments are read in place (napi_get_typedarray_info); solve() runs the C
// call on the libuv pool and resolves a Promise, which gives solveGPU its
// async shape (gpu/solver.ts:80-104).  Every C error becomes a rejected
// Promise or a thrown Error carrying pp_last_error().
#include <node_api.h>
#include <stdint.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/platepcg.h"

namespace {

#define NAPI_CALL(env, call)                                      \
  do {                                                            \
    if ((call) != napi_ok) {                                      \
      napi_throw_error((env), nullptr, "N-API call failed: " #call); \
      return nullptr;                                             \
    }                                                             \
  } while (0)

napi_value throw_pp(napi_env env, int rc) {
  std::string msg = "platepcg error " + std::to_string(rc) + ": " + pp_last_error();
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

template <typename T>
bool typed(napi_env env, napi_value v, napi_typedarray_type want, T** data, size_t* len) {
  bool is = false;
  if (napi_is_typedarray(env, v, &is) != napi_ok || !is) return false;
  napi_typedarray_type type;
  void* d = nullptr;
  napi_value ab;
  size_t off = 0;
  if (napi_get_typedarray_info(env, v, &type, len, &d, &ab, &off) != napi_ok) return false;
  if (type != want) return false;
  *data = (T*)d;
  return true;
}

bool is_null(napi_env env, napi_value v) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  return t == napi_null || t == napi_undefined;
}

double num(napi_env env, napi_value v) {
  double d = 0;
  napi_get_value_double(env, v, &d);
  return d;
}

double prop(napi_env env, napi_value obj, const char* name, double dflt) {
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return dflt;
  napi_value v;
  napi_get_named_property(env, obj, name, &v);
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t != napi_number) return dflt;
  return num(env, v);
}

void* external_ptr(napi_env env, napi_value v) {
  void* p = nullptr;
  napi_get_value_external(env, v, &p);
  return p;
}

napi_value make_int(napi_env env, int64_t v) {
  napi_value out;
  napi_create_int64(env, v, &out);
  return out;
}

napi_value make_num(napi_env env, double v) {
  napi_value out;
  napi_create_double(env, v, &out);
  return out;
}

// A plan as JavaScript holds it.  Solves run on the libuv pool; destroyPlan
// during one (or more) of them only marks the handle, and the last solve to
// complete frees the plan, so no worker ever uses a freed plan.  The C ABI
// itself serialises calls per context, so concurrent solves on one plan or
// one context are safe and each returns its own answer.
struct PlanHandle {
  pp_plan* plan = nullptr;
  int inflight = 0;        // solves queued or running (touched on the JS thread only)
  bool destroyed = false;  // destroyPlan was called
};

PlanHandle* plan_handle(napi_env env, napi_value v) {
  PlanHandle* h = (PlanHandle*)external_ptr(env, v);
  return h && !h->destroyed ? h : nullptr;
}

pp_plan* plan_of(napi_env env, napi_value v) {
  PlanHandle* h = plan_handle(env, v);
  return h ? h->plan : nullptr;
}

void handle_release_if_done(PlanHandle* h) {
  if (h->destroyed && h->inflight == 0) {
    pp_plan_destroy(h->plan);
    delete h;
  }
}

// ----------------------------------------------------------- functions

napi_value DeviceCount(napi_env env, napi_callback_info) { return make_int(env, pp_device_count()); }

napi_value CreateContext(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  pp_ctx* ctx = nullptr;
  int rc = pp_ctx_create(argc > 0 ? (int)num(env, argv[0]) : 0, &ctx);
  if (rc) return throw_pp(env, rc);
  napi_value out;
  NAPI_CALL(env, napi_create_external(env, ctx, nullptr, nullptr, &out));
  return out;
}

napi_value DestroyContext(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  pp_ctx_destroy((pp_ctx*)external_ptr(env, argv[0]));
  return nullptr;
}

// createPlan(ctx, mesh, material, colouring|null, mask Uint8Array)
//   mesh: {nodes: Float32Array, elements: Uint32Array, nodeCount, elementCount,
//          nodesPerElement, gridNx, gridNy}
//   material: {E, nu, t, elementT?: Float32Array}
//   colouring: {elementsByColour: Uint32Array, colourOffsets: Uint32Array}
napi_value CreatePlan(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value argv[5];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 5) {
    napi_throw_type_error(env, nullptr, "createPlan(ctx, mesh, material, colouring, mask)");
    return nullptr;
  }
  pp_ctx* ctx = (pp_ctx*)external_ptr(env, argv[0]);
  napi_value mesh = argv[1], mat = argv[2], col = argv[3];
  napi_value vn, ve;
  NAPI_CALL(env, napi_get_named_property(env, mesh, "nodes", &vn));
  NAPI_CALL(env, napi_get_named_property(env, mesh, "elements", &ve));
  float* nodes = nullptr;
  uint32_t* elems = nullptr;
  uint8_t* mask = nullptr;
  size_t ln = 0, le = 0, lm = 0;
  if (!typed(env, vn, napi_float32_array, &nodes, &ln) ||
      !typed(env, ve, napi_uint32_array, &elems, &le) ||
      !typed(env, argv[4], napi_uint8_array, &mask, &lm)) {
    napi_throw_type_error(env, nullptr, "nodes Float32Array, elements Uint32Array, mask Uint8Array");
    return nullptr;
  }
  pp_mesh m;
  m.nodes = nodes;
  m.elements = elems;
  m.node_count = (uint32_t)prop(env, mesh, "nodeCount", (double)(ln / 2));
  m.element_count = (uint32_t)prop(env, mesh, "elementCount", 0);
  m.nodes_per_element = (uint32_t)prop(env, mesh, "nodesPerElement", 4);
  m.grid_nx = (uint32_t)prop(env, mesh, "gridNx", 0);
  m.grid_ny = (uint32_t)prop(env, mesh, "gridNy", 0);
  if (ln < 2 * (size_t)m.node_count || le < (size_t)m.element_count * m.nodes_per_element ||
      lm < m.node_count) {
    napi_throw_range_error(env, nullptr, "mesh arrays shorter than the counts");
    return nullptr;
  }
  pp_material mt;
  mt.E = prop(env, mat, "E", 0);
  mt.nu = prop(env, mat, "nu", 0);
  mt.t = prop(env, mat, "t", 0);
  mt.element_t = nullptr;
  bool has_et = false;
  napi_has_named_property(env, mat, "elementT", &has_et);
  if (has_et) {
    napi_value vt;
    napi_get_named_property(env, mat, "elementT", &vt);
    if (!is_null(env, vt)) {
      float* et = nullptr;
      size_t lt = 0;
      if (!typed(env, vt, napi_float32_array, &et, &lt) || lt < m.element_count) {
        napi_throw_type_error(env, nullptr, "elementT must be a Float32Array per element");
        return nullptr;
      }
      mt.element_t = et;
    }
  }
  pp_colouring c;
  pp_colouring* cp = nullptr;
  if (!is_null(env, col)) {
    napi_value vo, vf;
    NAPI_CALL(env, napi_get_named_property(env, col, "elementsByColour", &vo));
    NAPI_CALL(env, napi_get_named_property(env, col, "colourOffsets", &vf));
    uint32_t *order = nullptr, *offs = nullptr;
    size_t lo = 0, lf = 0;
    if (!typed(env, vo, napi_uint32_array, &order, &lo) ||
        !typed(env, vf, napi_uint32_array, &offs, &lf) || lf < 2 || lo < m.element_count) {
      napi_throw_type_error(env, nullptr, "colouring arrays must be Uint32Array");
      return nullptr;
    }
    c.elements_by_colour = order;
    c.colour_offsets = offs;
    c.colour_count = (uint32_t)(lf - 1);
    cp = &c;
  }
  pp_plan* plan = nullptr;
  int rc = pp_plan_create(ctx, &m, &mt, cp, mask, &plan);
  if (rc) return throw_pp(env, rc);
  PlanHandle* h = new PlanHandle();
  h->plan = plan;
  napi_value out;
  NAPI_CALL(env, napi_create_external(env, h, nullptr, nullptr, &out));
  return out;
}

napi_value DestroyPlan(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  PlanHandle* h = plan_handle(env, argv[0]);
  if (!h) return nullptr;  // already destroyed
  h->destroyed = true;
  handle_release_if_done(h);
  return nullptr;
}

napi_value IsStructured(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  napi_value out;
  napi_get_boolean(env, pp_plan_is_structured(plan_of(env, argv[0])) != 0, &out);
  return out;
}

// ---------------------------------------------------------- async solve

struct SolveWork {
  napi_async_work work = nullptr;
  napi_deferred deferred = nullptr;
  napi_ref fref = nullptr;
  PlanHandle* handle = nullptr;  // counted in handle->inflight until complete
  pp_plan* plan = nullptr;
  const float* F = nullptr;
  size_t n = 0;
  pp_solve_opts opts;
  std::vector<float> x32;
  std::vector<double> x64;
  pp_solve_stats st;
  int rc = 0;
  std::string err;
};

void solve_execute(napi_env, void* data) {
  SolveWork* w = (SolveWork*)data;
  w->rc = pp_solve(w->plan, w->F, &w->opts, w->x32.data(), w->x64.data(), &w->st);
  if (w->rc) w->err = pp_last_error();
}

void solve_complete(napi_env env, napi_status, void* data) {
  SolveWork* w = (SolveWork*)data;
  if (w->rc) {
    napi_value msg, err;
    std::string s = "platepcg error " + std::to_string(w->rc) + ": " + w->err;
    napi_create_string_utf8(env, s.c_str(), s.size(), &msg);
    napi_create_error(env, nullptr, msg, &err);
    napi_reject_deferred(env, w->deferred, err);
  } else {
    napi_value res, ab32, ab64, a32, a64;
    void* d32 = nullptr;
    void* d64 = nullptr;
    napi_create_object(env, &res);
    napi_create_arraybuffer(env, w->n * 4, &d32, &ab32);
    napi_create_arraybuffer(env, w->n * 8, &d64, &ab64);
    memcpy(d32, w->x32.data(), w->n * 4);
    memcpy(d64, w->x64.data(), w->n * 8);
    napi_create_typedarray(env, napi_float32_array, w->n, ab32, 0, &a32);
    napi_create_typedarray(env, napi_float64_array, w->n, ab64, 0, &a64);
    napi_set_named_property(env, res, "solution", a32);
    napi_set_named_property(env, res, "solution64", a64);
    napi_set_named_property(env, res, "iterations", make_int(env, w->st.iterations));
    napi_set_named_property(env, res, "finalResidual", make_num(env, w->st.final_residual));
    napi_value conv, brk;
    napi_get_boolean(env, w->st.converged != 0, &conv);
    napi_get_boolean(env, w->st.breakdown != 0, &brk);
    napi_set_named_property(env, res, "converged", conv);
    napi_set_named_property(env, res, "breakdown", brk);
    napi_set_named_property(env, res, "trueRelResidual", make_num(env, w->st.true_rel_residual));
    napi_set_named_property(env, res, "deviceMs", make_num(env, w->st.device_ms));
    napi_set_named_property(env, res, "sweeps", make_int(env, w->st.sweeps));
    napi_resolve_deferred(env, w->deferred, res);
  }
  napi_delete_reference(env, w->fref);
  napi_delete_async_work(env, w->work);
  w->handle->inflight--;
  handle_release_if_done(w->handle);
  delete w;
}

// solve(plan, F Float32Array, opts) -> Promise
napi_value Solve(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  SolveWork* w = new SolveWork();
  w->handle = argc >= 3 ? plan_handle(env, argv[0]) : nullptr;
  w->plan = w->handle ? w->handle->plan : nullptr;
  float* F = nullptr;
  if (!w->plan || !typed(env, argv[1], napi_float32_array, &F, &w->n)) {
    delete w;
    napi_throw_type_error(env, nullptr, "solve(plan, F: Float32Array, options) on a live plan");
    return nullptr;
  }
  w->F = F;
  napi_value o = argv[2];
  w->opts.tolerance = prop(env, o, "tolerance", 1e-6);
  w->opts.max_iterations = (int32_t)prop(env, o, "maxIterations", 1000);
  w->opts.precision = (int32_t)prop(env, o, "precision", PP_PRECISION_FP64);
  w->opts.precond = (int32_t)prop(env, o, "precond", PP_PRECOND_BLOCK_JACOBI);
  w->opts.path = (int32_t)prop(env, o, "path", PP_PATH_AUTO);
  w->opts.inner_tolerance = prop(env, o, "innerTolerance", 0.0);
  w->opts.max_refinements = (int32_t)prop(env, o, "maxRefinements", 0);
  w->opts.flags = (int32_t)prop(env, o, "flags", 0);
  w->opts.poly_degree = (int32_t)prop(env, o, "polyDegree", 0);
  w->x32.resize(w->n);
  w->x64.resize(w->n);
  NAPI_CALL(env, napi_create_reference(env, argv[1], 1, &w->fref));  // keep F alive
  napi_value promise, name;
  NAPI_CALL(env, napi_create_promise(env, &w->deferred, &promise));
  napi_create_string_utf8(env, "platepcg.solve", NAPI_AUTO_LENGTH, &name);
  NAPI_CALL(env, napi_create_async_work(env, nullptr, name, solve_execute, solve_complete, w,
                                        &w->work));
  NAPI_CALL(env, napi_queue_async_work(env, w->work));
  w->handle->inflight++;
  return promise;
}

// applyK(plan, x Float64Array, y Float64Array, path)
napi_value ApplyK(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  double *x = nullptr, *y = nullptr;
  size_t lx = 0, ly = 0;
  if (!typed(env, argv[1], napi_float64_array, &x, &lx) ||
      !typed(env, argv[2], napi_float64_array, &y, &ly) || lx != ly) {
    napi_throw_type_error(env, nullptr, "applyK(plan, x: Float64Array, y: Float64Array, path)");
    return nullptr;
  }
  int rc = pp_apply_k(plan_of(env, argv[0]), x, y,
                      argc > 3 ? (int)num(env, argv[3]) : PP_PATH_AUTO);
  if (rc) return throw_pp(env, rc);
  return nullptr;
}

// pressureLoad(plan, q0, patches Float64Array (4 per patch), F Float32Array)
napi_value PressureLoad(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  double* pt = nullptr;
  float* F = nullptr;
  size_t lp = 0, lf = 0;
  if (!typed(env, argv[2], napi_float64_array, &pt, &lp) ||
      !typed(env, argv[3], napi_float32_array, &F, &lf)) {
    napi_throw_type_error(env, nullptr, "pressureLoad(plan, q0, patches, F)");
    return nullptr;
  }
  int rc = pp_pressure_load(plan_of(env, argv[0]), num(env, argv[1]),
                            lp ? pt : nullptr, (int)(lp / 4), F);
  if (rc) return throw_pp(env, rc);
  return nullptr;
}

// moments(plan, x64 | null, Mx, My, Mxy): Float32Arrays of node count; null
// x = the plan's solution on the device (computeMoments, postprocess.ts:40-270)
napi_value Moments(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value argv[5];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  double* x = nullptr;
  float* M[3] = {nullptr, nullptr, nullptr};
  size_t lx = 0, lm[3] = {0, 0, 0};
  bool ok = argc == 5 && (is_null(env, argv[1]) || typed(env, argv[1], napi_float64_array, &x, &lx));
  for (int k = 0; k < 3 && ok; k++) ok = typed(env, argv[2 + k], napi_float32_array, &M[k], &lm[k]);
  if (!ok || lm[0] != lm[1] || lm[1] != lm[2] || (x && lx != 3 * lm[0])) {
    napi_throw_type_error(env, nullptr, "moments(plan, x64|null, Mx, My, Mxy)");
    return nullptr;
  }
  int rc = pp_moments(plan_of(env, argv[0]), x, M[0], M[1], M[2]);
  if (rc) return throw_pp(env, rc);
  return nullptr;
}

// maxDisplacement(plan, x64 | null) -> {maxW, nodeIndex, x, y}
// (findMaxDisplacement, postprocess.ts:296-319)
napi_value MaxDisplacement(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  double* x = nullptr;
  size_t lx = 0;
  if (argc < 2 || !(is_null(env, argv[1]) || typed(env, argv[1], napi_float64_array, &x, &lx))) {
    napi_throw_type_error(env, nullptr, "maxDisplacement(plan, x64|null)");
    return nullptr;
  }
  double w = 0;
  int64_t node = 0;
  float cx = 0, cy = 0;
  int rc = pp_max_displacement(plan_of(env, argv[0]), x, &w, &node, &cx, &cy);
  if (rc) return throw_pp(env, rc);
  napi_value out, v;
  NAPI_CALL(env, napi_create_object(env, &out));
  napi_create_double(env, w, &v);
  napi_set_named_property(env, out, "maxW", v);
  napi_set_named_property(env, out, "nodeIndex", make_int(env, node));
  napi_create_double(env, cx, &v);
  napi_set_named_property(env, out, "x", v);
  napi_create_double(env, cy, &v);
  napi_set_named_property(env, out, "y", v);
  return out;
}

napi_value Init(napi_env env, napi_value exports) {
  const napi_property_descriptor d[] = {
      {"deviceCount", nullptr, DeviceCount, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"createContext", nullptr, CreateContext, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"destroyContext", nullptr, DestroyContext, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"createPlan", nullptr, CreatePlan, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"destroyPlan", nullptr, DestroyPlan, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"isStructured", nullptr, IsStructured, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"solve", nullptr, Solve, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"applyK", nullptr, ApplyK, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"pressureLoad", nullptr, PressureLoad, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"moments", nullptr, Moments, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"maxDisplacement", nullptr, MaxDisplacement, nullptr, nullptr, nullptr, napi_default,
       nullptr},
  };
  napi_define_properties(env, exports, sizeof(d) / sizeof(d[0]), d);
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
