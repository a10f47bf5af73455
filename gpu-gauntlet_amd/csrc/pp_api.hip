// The C ABI of platepcg: context, plan, solve, unit entry points.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>

#include "pp_cnn.h"

namespace pp {

namespace {
thread_local std::string g_last_error;

__global__ __launch_bounds__(256) void k_resid_parts(int64_t n, const double* F, const double* Kx,
                                                     double* part) {
  __shared__ double red[16];
  double v[2] = {0, 0};
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double d = F[i] - Kx[i];
    v[0] += d * d;
    v[1] += F[i] * F[i];
  }
  block_sum<2, 256>(v, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = v[0];
    part[kNormBlocks + blockIdx.x] = v[1];
  }
}

__global__ void k_f32_to_f64(int64_t n, const float* a, double* b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}

__global__ void k_f64_to_f32(int64_t n, const double* a, float* b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

__global__ void k_residual(int64_t n, const double* F, const double* Kx, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = F[i] - Kx[i];
}

int nblocks(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

bool detect_structured(const pp_mesh* m) {
  const uint32_t nx = m->grid_nx, ny = m->grid_ny;
  if (m->nodes_per_element != 4 || nx == 0 || ny == 0) return false;
  if ((uint64_t)m->node_count != (uint64_t)(nx + 1) * (ny + 1)) return false;
  if ((uint64_t)m->element_count != (uint64_t)nx * ny) return false;
  const uint32_t W = nx + 1;
  for (uint32_t j = 0; j < ny; j++)
    for (uint32_t i = 0; i < nx; i++) {
      const uint32_t* e = m->elements + 4 * ((size_t)j * nx + i);
      if (e[0] != j * W + i || e[1] != j * W + i + 1 || e[2] != (j + 1) * W + i + 1 ||
          e[3] != (j + 1) * W + i)
        return false;
    }
  for (uint32_t j = 0; j <= ny; j++)
    for (uint32_t i = 0; i <= nx; i++) {
      const float* p = m->nodes + 2 * ((size_t)j * W + i);
      if (p[0] != m->nodes[2 * i] || p[1] != m->nodes[2 * ((size_t)j * W) + 1]) return false;
      if (i > 0 && !(p[0] > p[-2])) return false;
      if (j > 0 && !(p[1] > m->nodes[2 * ((size_t)(j - 1) * W) + 1])) return false;
    }
  return true;
}

int ensure_solve_buffers(pp_plan* pl) {
  const size_t nd = (size_t)pl->n_dof;
  if (pl->x.n == nd) return PP_OK;
  int rc;
  if ((rc = pl->x.alloc(nd)) || (rc = pl->r.alloc(nd)) || (rc = pl->z.alloc(nd)) ||
      (rc = pl->p0.alloc(nd)) || (rc = pl->p1.alloc(nd)) || (rc = pl->Ap.alloc(nd)) ||
      (rc = pl->tmp.alloc(nd)))
    return rc;
  return PP_OK;
}

}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int launch_residual(pp_plan* pl, const double* dF, const double* dKx, double* out) {
  hipLaunchKernelGGL(k_residual, dim3(nblocks(pl->n_dof, 256)), dim3(256), 0, pl->ctx->stream,
                     (int64_t)pl->n_dof, dF, dKx, out);
  PP_HIP(hipGetLastError());
  return PP_OK;
}

int launch_norm_ratio(pp_plan* pl, const double* dF, const double* dKx, double* rel) {
  double a = 0, b = 0;
  int rc = launch_norms(pl, dF, dKx, &a, &b);
  if (rc) return rc;
  *rel = b > 0 ? a / b : 0.0;
  return PP_OK;
}

int launch_norms(pp_plan* pl, const double* dF, const double* dKx, double* res, double* fnorm) {
  int rc;
  if (pl->normp.n < kNormParts && (rc = pl->normp.alloc(kNormParts))) return rc;
  return launch_norms_into(pl, pl->normp.p, dF, dKx, res, fnorm);
}

int launch_norms_into(pp_plan* pl, double* part, const double* dF, const double* dKx, double* res,
                      double* fnorm) {
  hipStream_t s = pl->ctx->stream;
  hipLaunchKernelGGL(k_resid_parts, dim3(kNormBlocks), dim3(256), 0, s, (int64_t)pl->n_dof, dF, dKx,
                     part);
  PP_HIP(hipGetLastError());
  double h[kNormParts];
  PP_HIP(hipMemcpyAsync(h, part, sizeof(h), hipMemcpyDeviceToHost, s));
  PP_HIP(hipStreamSynchronize(s));
  double a = 0, b = 0;
  for (int i = 0; i < kNormBlocks; i++) {
    a += h[i];
    b += h[kNormBlocks + i];
  }
  *res = sqrt(a);
  *fnorm = sqrt(b);
  return PP_OK;
}

// A context is freed when its owner and every plan made from it have let go.
void ctx_release(pp_ctx* c) {
  if (--c->refs > 0) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

// A plan is freed when its owner and every network made on it have let go.
void plan_release(pp_plan* pl) {
  pp_ctx* c = pl->ctx;
  {
    PP_LOCK(c);
    if (--pl->refs > 0) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    mg_destroy(pl->mg_pc);
    if (pl->mgf_pinned) (void)hipHostFree(pl->mgf_pinned);
    delete pl;
  }
  ctx_release(c);
}

}  // namespace pp

using namespace pp;

extern "C" {

int pp_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char* pp_last_error(void) { return g_last_error.c_str(); }

int pp_ctx_create(int device, pp_ctx** out) {
  if (!out) return fail(PP_ERR_INVALID, "out is NULL");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(PP_ERR_NODEVICE, "no HIP device");
  if (device < 0 || device >= n) return fail(PP_ERR_INVALID, "device index out of range");
  PP_HIP(hipSetDevice(device));
  pp_ctx* c = new pp_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e != hipSuccess) {
    delete c;
    return fail(PP_ERR_HIP, std::string("context: ") + hipGetErrorString(e));
  }
  c->cu_count = prop.multiProcessorCount;
  if (int rc = noalloc_scratch_init(device)) {
    (void)hipEventDestroy(c->ev0);
    (void)hipEventDestroy(c->ev1);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  *out = c;
  return PP_OK;
}

int pp_ctx_destroy(pp_ctx* c) {
  if (!c) return PP_OK;
  ctx_release(c);
  return PP_OK;
}

int pp_plan_create(pp_ctx* ctx, const pp_mesh* m, const pp_material* mat,
                   const pp_colouring* col, const uint8_t* cmask, pp_plan** out) {
  if (!ctx || !m || !mat || !cmask || !out) return fail(PP_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (m->nodes_per_element != 3 && m->nodes_per_element != 4)
    return fail(PP_ERR_INVALID, "nodes_per_element must be 3 or 4");
  if (!m->nodes || !m->elements || m->node_count == 0 || m->element_count == 0)
    return fail(PP_ERR_INVALID, "empty mesh");
  if (!(mat->E > 0) || !(mat->t > 0) || !(mat->nu > -1.0 && mat->nu < 0.5))
    return fail(PP_ERR_INVALID, "material out of range");
  const size_t ncon = (size_t)m->element_count * m->nodes_per_element;
  for (size_t k = 0; k < ncon; k++)
    if (m->elements[k] >= m->node_count) return fail(PP_ERR_INVALID, "element node index out of range");
  PP_LOCK(ctx);
  PP_HIP(hipSetDevice(ctx->device));
  pp_plan* pl = new pp_plan();
  pl->ctx = ctx;
  pl->n_nodes = m->node_count;
  pl->n_elems = m->element_count;
  pl->n_dof = 3 * (int64_t)m->node_count;
  pl->npe = (int)m->nodes_per_element;
  pl->E = mat->E;
  pl->nu = mat->nu;
  pl->t = mat->t;
  pl->var_t = mat->element_t != nullptr;
  pl->h_nodes.assign(m->nodes, m->nodes + 2 * (size_t)m->node_count);
  pl->h_elems.assign(m->elements, m->elements + ncon);
  pl->h_cmask.assign(cmask, cmask + m->node_count);
  if (pl->var_t) pl->h_elem_t.assign(mat->element_t, mat->element_t + m->element_count);
  pl->structured = detect_structured(m);
  pl->nx = (int)m->grid_nx;
  pl->ny = (int)m->grid_ny;
  int rc = PP_OK;
  hipStream_t s = ctx->stream;
  if ((rc = pl->cmask.alloc(pl->n_nodes)) || (rc = pl->gsync.alloc(1)) ||
      (rc = pl->dstats.alloc(1)) || (rc = pl->F.alloc(pl->n_dof)) ||
      (rc = pl->F32.alloc(pl->n_dof)) || (rc = ensure_solve_buffers(pl))) {
    delete pl;
    return rc;
  }
  if (pl->var_t) {
    if ((rc = pl->elem_t.alloc(pl->n_elems))) { delete pl; return rc; }
    hipError_t e = hipMemcpyAsync(pl->elem_t.p, pl->h_elem_t.data(), pl->n_elems * 4,
                                  hipMemcpyHostToDevice, s);
    if (e != hipSuccess) { delete pl; return fail(PP_ERR_HIP, hipGetErrorString(e)); }
  }
  hipError_t e = hipMemcpyAsync(pl->cmask.p, cmask, pl->n_nodes, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) { delete pl; return fail(PP_ERR_HIP, hipGetErrorString(e)); }
  // colour offsets (the coloured path is always available for parity tests)
  if (col && col->colour_count > 0 && col->colour_offsets && col->elements_by_colour) {
    pl->h_colour_off.assign(col->colour_offsets, col->colour_offsets + col->colour_count + 1);
    // offsets must run from 0 up to element_count without decreasing before
    // any loop below indexes elements_by_colour with them
    bool ordered = pl->h_colour_off.front() == 0 && pl->h_colour_off.back() == m->element_count;
    for (uint32_t c = 0; ordered && c < col->colour_count; c++)
      ordered = pl->h_colour_off[c] <= pl->h_colour_off[c + 1];
    if (!ordered) {
      delete pl;
      return fail(PP_ERR_INVALID,
                  "colour offsets must start at 0, never decrease and end at element_count");
    }
    std::vector<uint8_t> seen(m->element_count, 0);
    for (uint32_t k = 0; k < m->element_count; k++) {
      const uint32_t el = col->elements_by_colour[k];
      if (el >= m->element_count || seen[el]) {
        delete pl;
        return fail(PP_ERR_INVALID, "colouring does not cover every element exactly once");
      }
      seen[el] = 1;
    }
    // race freedom: no node twice within a colour
    std::vector<uint32_t> stamp(m->node_count, 0xffffffffu);
    for (uint32_t c = 0; c < col->colour_count; c++)
      for (uint32_t k = col->colour_offsets[c]; k < col->colour_offsets[c + 1]; k++) {
        const uint32_t el = col->elements_by_colour[k];
        for (uint32_t i = 0; i < m->nodes_per_element; i++) {
          const uint32_t nd = m->elements[(size_t)el * m->nodes_per_element + i];
          if (stamp[nd] == c) {
            delete pl;
            return fail(PP_ERR_INVALID, "colouring has two elements sharing a node in one colour");
          }
          stamp[nd] = c;
        }
      }
    if ((rc = general_setup(pl, col))) { delete pl; return rc; }
  } else if (!pl->structured) {
    delete pl;
    return fail(PP_ERR_INVALID, "a colouring is required for a mesh that is not a tensor grid");
  }
  if (pl->structured) {
    // the structured path owns the block inverses when it exists
    if ((rc = struct_setup(pl))) { delete pl; return rc; }
  }
  ctx->refs++;
  *out = pl;
  return PP_OK;
}

int pp_plan_destroy(pp_plan* pl) {
  if (!pl) return PP_OK;
  plan_release(pl);
  return PP_OK;
}

int pp_plan_is_structured(const pp_plan* pl) { return pl && pl->structured ? 1 : 0; }

int pp_set_rhs(pp_plan* pl, const float* F) {
  if (!pl || !F) return fail(PP_ERR_INVALID, "NULL argument");
  PP_LOCK(pl->ctx);
  hipStream_t s = pl->ctx->stream;
  PP_HIP(hipMemcpyAsync(pl->F32.p, F, pl->n_dof * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_f32_to_f64, dim3(nblocks(pl->n_dof, 256)), dim3(256), 0, s,
                     (int64_t)pl->n_dof, (const float*)pl->F32.p, pl->F.p);
  PP_HIP(hipGetLastError());
  PP_HIP(hipStreamSynchronize(s));
  pl->rhs_set = true;
  return PP_OK;
}

static int pick_path(pp_plan* pl, int path, bool* structured) {
  if (path == PP_PATH_STRUCTURED) {
    if (!pl->structured) return fail(PP_ERR_UNSUPPORTED, "mesh is not a tensor grid");
    *structured = true;
  } else if (path == PP_PATH_COLOURED) {
    if (pl->h_colour_off.empty()) return fail(PP_ERR_UNSUPPORTED, "plan has no colouring");
    *structured = false;
  } else if (path == PP_PATH_AUTO) {
    *structured = pl->structured;
  } else {
    return fail(PP_ERR_INVALID, "unknown path");
  }
  return PP_OK;
}

int pp_solve_resident(pp_plan* pl, const pp_solve_opts* o, pp_solve_stats* st) {
  if (!pl || !o || !st) return fail(PP_ERR_INVALID, "NULL argument");
  PP_LOCK(pl->ctx);
  if (!pl->rhs_set) return fail(PP_ERR_INVALID, "no right-hand side set");
  if (!(o->tolerance > 0) || o->max_iterations < 0) return fail(PP_ERR_INVALID, "bad tolerance/max_iterations");
  if (o->precond != PP_PRECOND_BLOCK_JACOBI && o->precond != PP_PRECOND_LEARNED &&
      o->precond != PP_PRECOND_CHEBYSHEV && o->precond != PP_PRECOND_MULTIGRID)
    return fail(PP_ERR_INVALID, "unknown preconditioner");
  memset(st, 0, sizeof(*st));
  PP_HIP(hipSetDevice(pl->ctx->device));
  bool structured = false;
  int rc = pick_path(pl, o->path, &structured);
  if (rc) return rc;
  if (o->precond == PP_PRECOND_LEARNED) {
    if (!structured) return fail(PP_ERR_UNSUPPORTED, "the learned preconditioner needs the structured path");
    if (!pl->learned) return fail(PP_ERR_INVALID, "no network attached (pp_plan_set_learned)");
    if (o->precision != PP_PRECISION_FP64)
      return fail(PP_ERR_UNSUPPORTED, "the learned preconditioner runs with FP64 flexible PCG");
    rc = learned_solve(pl, o, st);
  } else if (o->precond == PP_PRECOND_MULTIGRID) {
    if (!structured || o->precision != PP_PRECISION_FP64)
      return fail(PP_ERR_UNSUPPORTED, "the multigrid preconditioner runs on the structured FP64 path");
    rc = mg_solve(pl, o, st);
  } else if (o->precond == PP_PRECOND_CHEBYSHEV) {
    if (!structured || o->precision != PP_PRECISION_FP64)
      return fail(PP_ERR_UNSUPPORTED, "the Chebyshev preconditioner runs on the structured FP64 path");
    rc = struct_solve_fp64(pl, o, st);
  } else if (o->precision == PP_PRECISION_FP64) {
    rc = structured ? struct_solve_fp64(pl, o, st) : general_solve_fp64(pl, o, st);
  } else if (o->precision == PP_PRECISION_MIXED) {
    if (!structured) return fail(PP_ERR_UNSUPPORTED, "mixed precision needs the structured path");
    rc = struct_solve_mixed(pl, o, st);
  } else {
    return fail(PP_ERR_INVALID, "unknown precision");
  }
  // every solver measures the FP64 true residual itself
  return rc;
}

int pp_get_solution(pp_plan* pl, double* x64, float* x32) {
  if (!pl) return fail(PP_ERR_INVALID, "NULL plan");
  PP_LOCK(pl->ctx);
  hipStream_t s = pl->ctx->stream;
  if (x64) PP_HIP(hipMemcpyAsync(x64, pl->x.p, pl->n_dof * 8, hipMemcpyDeviceToHost, s));
  if (x32) {
    DevBuf<float> f;
    int rc = f.alloc(pl->n_dof);
    if (rc) return rc;
    hipLaunchKernelGGL(k_f64_to_f32, dim3(nblocks(pl->n_dof, 256)), dim3(256), 0, s,
                       (int64_t)pl->n_dof, (const double*)pl->x.p, f.p);
    PP_HIP(hipGetLastError());
    PP_HIP(hipMemcpyAsync(x32, f.p, pl->n_dof * 4, hipMemcpyDeviceToHost, s));
    PP_HIP(hipStreamSynchronize(s));
  }
  PP_HIP(hipStreamSynchronize(s));
  return PP_OK;
}

int pp_solve(pp_plan* pl, const float* F, const pp_solve_opts* o, float* x32, double* x64,
             pp_solve_stats* st) {
  if (!pl) return fail(PP_ERR_INVALID, "NULL plan");
  PP_LOCK(pl->ctx);
  int rc = pp_set_rhs(pl, F);
  if (rc) return rc;
  rc = pp_solve_resident(pl, o, st);
  if (rc) return rc;
  return pp_get_solution(pl, x64, x32);
}

static int with_device_vectors(pp_plan* pl, const double* in, double* out,
                               int (*fn)(pp_plan*, const double*, double*, void*), void* u) {
  DevBuf<double> a, b;
  int rc;
  if ((rc = a.alloc(pl->n_dof)) || (rc = b.alloc(pl->n_dof))) return rc;
  hipStream_t s = pl->ctx->stream;
  PP_HIP(hipMemcpyAsync(a.p, in, pl->n_dof * 8, hipMemcpyHostToDevice, s));
  if ((rc = fn(pl, a.p, b.p, u))) return rc;
  PP_HIP(hipMemcpyAsync(out, b.p, pl->n_dof * 8, hipMemcpyDeviceToHost, s));
  PP_HIP(hipStreamSynchronize(s));
  return PP_OK;
}

int pp_apply_k(pp_plan* pl, const double* x, double* y, int path) {
  if (!pl || !x || !y) return fail(PP_ERR_INVALID, "NULL argument");
  PP_LOCK(pl->ctx);
  bool structured = false;
  int rc = pick_path(pl, path, &structured);
  if (rc) return rc;
  return with_device_vectors(
      pl, x, y,
      [](pp_plan* p, const double* a, double* b, void* u) {
        return *(bool*)u ? struct_apply_k(p, a, b) : general_apply_k(p, a, b, -1, true);
      },
      &structured);
}

int pp_apply_k_colour(pp_plan* pl, int colour, const double* x, double* y) {
  if (!pl || !x || !y) return fail(PP_ERR_INVALID, "NULL argument");
  PP_LOCK(pl->ctx);
  if (pl->h_colour_off.empty()) return fail(PP_ERR_UNSUPPORTED, "plan has no colouring");
  if (colour < 0 || colour + 1 >= (int)pl->h_colour_off.size())
    return fail(PP_ERR_INVALID, "colour out of range");
  return with_device_vectors(
      pl, x, y,
      [](pp_plan* p, const double* a, double* b, void* u) {
        return general_apply_k(p, a, b, *(int*)u, false);
      },
      &colour);
}

int pp_apply_precond(pp_plan* pl, const double* r, double* z) {
  if (!pl || !r || !z) return fail(PP_ERR_INVALID, "NULL argument");
  PP_LOCK(pl->ctx);
  return with_device_vectors(
      pl, r, z,
      [](pp_plan* p, const double* a, double* b, void*) {
        return p->structured ? struct_apply_precond(p, a, b) : general_apply_precond(p, a, b);
      },
      nullptr);
}

int pp_get_block_inverse(pp_plan* pl, double* blocks) {
  if (!pl || !blocks) return fail(PP_ERR_INVALID, "NULL argument");
  PP_LOCK(pl->ctx);
  std::vector<double> m(6 * (size_t)pl->n_nodes);
  PP_HIP(hipMemcpyAsync(m.data(), pl->minv.p, m.size() * 8, hipMemcpyDeviceToHost, pl->ctx->stream));
  PP_HIP(hipStreamSynchronize(pl->ctx->stream));
  for (int64_t n = 0; n < pl->n_nodes; n++) {
    const double* a = &m[6 * n];
    double* b = blocks + 9 * n;
    b[0] = a[0]; b[1] = a[1]; b[2] = a[2];
    b[3] = a[1]; b[4] = a[3]; b[5] = a[4];
    b[6] = a[2]; b[7] = a[4]; b[8] = a[5];
  }
  return PP_OK;
}

int pp_pressure_load(pp_plan* pl, double q0, const double* patches, int n_patches, float* F) {
  if (!pl || !F || n_patches < 0 || (n_patches > 0 && !patches))
    return fail(PP_ERR_INVALID, "bad argument");
  PP_LOCK(pl->ctx);
  DevBuf<double> pt;
  DevBuf<float> f;
  int rc;
  if ((rc = pt.alloc(std::max(1, 4 * n_patches))) || (rc = f.alloc(pl->n_dof))) return rc;
  hipStream_t s = pl->ctx->stream;
  if (n_patches)
    PP_HIP(hipMemcpyAsync(pt.p, patches, 4 * n_patches * 8, hipMemcpyHostToDevice, s));
  rc = pl->structured ? struct_pressure(pl, q0, pt.p, n_patches, f.p)
                      : general_pressure(pl, q0, pt.p, n_patches, f.p);
  if (rc) return rc;
  PP_HIP(hipMemcpyAsync(F, f.p, pl->n_dof * 4, hipMemcpyDeviceToHost, s));
  PP_HIP(hipStreamSynchronize(s));
  return PP_OK;
}

}  // extern "C"
