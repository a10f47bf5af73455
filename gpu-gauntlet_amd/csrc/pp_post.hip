// Post-processing on the device: nodal bending moments and the largest
// deflection of a solution (postprocess.ts:40-319), for every plan (tensor
// grids and general Q4 / triangle meshes alike).
//
// Moments: one thread per element evaluates the curvatures at the element
// centroid (Q4: isoparametric derivatives at xi = eta = 0, element.ts:84-184;
// triangles: linear shape functions, postprocess.ts:213-236) and
// M = D [k_x + nu k_y, k_y + nu k_x, (1 - nu)/2 k_xy] with
// k_x = d thy/dx, k_y = -d thx/dy, k_xy = d thy/dy - d thx/dx; then one
// thread per node averages its elements' values over a node -> element list
// in ascending element order (the order of the reference's element loop), so
// the result is deterministic and needs no atomics.  Degenerate triangles
// (|2A| < 1e-15) are skipped as in postprocess.ts:202.  Arithmetic is FP64
// throughout; the outputs are float32 like the reference's Float32Arrays.
//
// Layout: connectivity u32 [npe per element], coordinates f32 [x, y] per
// node (both as the caller passed them), the element values 4 doubles per
// element (Mx, My, Mxy, 1 = counted / 0 = skipped), the node list as CSR.

#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "pp_device.h"

namespace pp {

namespace {

constexpr int kBlock = 256;

int nblk(int64_t n) { return (int)std::max<int64_t>(1, (n + kBlock - 1) / kBlock); }

template <int NPE>
__global__ __launch_bounds__(kBlock) void k_elem_moments(int64_t ne, const uint32_t* elems,
                                                         const float* nodes, const double* x,
                                                         double Dfac, double t0, const float* elem_t,
                                                         double nu, double* em) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= ne) return;
  uint32_t nd[NPE];
  double cx[NPE], cy[NPE], thx[NPE], thy[NPE];
#pragma unroll
  for (int i = 0; i < NPE; i++) {
    nd[i] = elems[e * NPE + i];
    cx[i] = (double)nodes[2 * (size_t)nd[i]];
    cy[i] = (double)nodes[2 * (size_t)nd[i] + 1];
    thx[i] = x[3 * (size_t)nd[i] + 1];
    thy[i] = x[3 * (size_t)nd[i] + 2];
  }
  double dNdx[NPE], dNdy[NPE];
  bool ok = true;
  if constexpr (NPE == 4) {
    // dN/dxi and dN/deta of the bilinear shape functions at the centroid
    const double dxi[4] = {-0.25, 0.25, 0.25, -0.25};
    const double deta[4] = {-0.25, -0.25, 0.25, 0.25};
    double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      J11 += dxi[i] * cx[i];
      J12 += dxi[i] * cy[i];
      J21 += deta[i] * cx[i];
      J22 += deta[i] * cy[i];
    }
    const double inv = 1.0 / (J11 * J22 - J12 * J21);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      dNdx[i] = inv * (J22 * dxi[i] - J12 * deta[i]);
      dNdy[i] = inv * (-J21 * dxi[i] + J11 * deta[i]);
    }
  } else {
    const double twoA = (cx[1] - cx[0]) * (cy[2] - cy[0]) - (cx[2] - cx[0]) * (cy[1] - cy[0]);
    ok = fabs(twoA) >= 1e-15;
    const double inv = 1.0 / twoA;
    dNdx[0] = (cy[1] - cy[2]) * inv;
    dNdy[0] = (cx[2] - cx[1]) * inv;
    dNdx[1] = (cy[2] - cy[0]) * inv;
    dNdy[1] = (cx[0] - cx[2]) * inv;
    dNdx[2] = (cy[0] - cy[1]) * inv;
    dNdy[2] = (cx[1] - cx[0]) * inv;
  }
  double kx = 0, ky = 0, kxy = 0;
#pragma unroll
  for (int i = 0; i < NPE; i++) {
    kx += dNdx[i] * thy[i];
    ky += -dNdy[i] * thx[i];
    kxy += dNdy[i] * thy[i] - dNdx[i] * thx[i];
  }
  const double t = elem_t ? (double)elem_t[e] : t0;
  const double D = Dfac * t * t * t;
  double* o = em + 4 * e;
  if (ok) {
    o[0] = D * (kx + nu * ky);
    o[1] = D * (ky + nu * kx);
    o[2] = D * (1.0 - nu) * kxy / 2.0;
    o[3] = 1.0;
  } else {
    o[0] = o[1] = o[2] = o[3] = 0.0;
  }
}

__global__ __launch_bounds__(kBlock) void k_node_moments(int64_t nn, const int64_t* off,
                                                         const uint32_t* adj, const double* em,
                                                         float* out) {
  const int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (n >= nn) return;
  double s0 = 0, s1 = 0, s2 = 0, cnt = 0;
  for (int64_t k = off[n]; k < off[n + 1]; k++) {
    const double* o = em + 4 * (size_t)adj[k];
    if (o[3] != 0.0) {
      s0 += o[0];
      s1 += o[1];
      s2 += o[2];
      cnt += 1.0;
    }
  }
  out[n] = cnt > 0 ? (float)(s0 / cnt) : 0.0f;
  out[nn + n] = cnt > 0 ? (float)(s1 / cnt) : 0.0f;
  out[2 * nn + n] = cnt > 0 ? (float)(s2 / cnt) : 0.0f;
}

// Largest |w| with the lowest node index among equals (the strict '>' scan of
// postprocess.ts:304-311): per-block candidates, then one block.
__device__ __forceinline__ void better(double& v, int64_t& i, double v2, int64_t i2) {
  if (v2 > v || (v2 == v && i2 < i)) {
    v = v2;
    i = i2;
  }
}

__device__ __forceinline__ void block_argmax(double& v, int64_t& i, double* sv, int64_t* si) {
  for (int off = 32; off > 0; off >>= 1) {
    const double v2 = __shfl_xor(v, off, 64);
    const int64_t i2 = __shfl_xor(i, off, 64);
    better(v, i, v2, i2);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    sv[wid] = v;
    si[wid] = i;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < kBlock / 64; w++) better(v, i, sv[w], si[w]);
}

__global__ __launch_bounds__(kBlock) void k_absmax_w(int64_t nn, const double* x, double* pv,
                                                     int64_t* pi) {
  __shared__ double sv[kBlock / 64];
  __shared__ int64_t si[kBlock / 64];
  double v = -1.0;
  int64_t i = nn;
  for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < nn;
       n += (int64_t)gridDim.x * kBlock)
    better(v, i, fabs(x[3 * n]), n);
  block_argmax(v, i, sv, si);
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = v;
    pi[blockIdx.x] = i;
  }
}

__global__ __launch_bounds__(kBlock) void k_absmax_final(int nparts, const double* pv,
                                                         const int64_t* pi, int64_t nn,
                                                         const double* x, double* out) {
  __shared__ double sv[kBlock / 64];
  __shared__ int64_t si[kBlock / 64];
  double v = -1.0;
  int64_t i = nn;
  for (int k = threadIdx.x; k < nparts; k += kBlock) better(v, i, pv[k], pi[k]);
  block_argmax(v, i, sv, si);
  if (threadIdx.x == 0) {
    out[0] = i < nn ? x[3 * i] : 0.0;  // the signed value
    out[1] = (double)(i < nn ? i : 0);
  }
}

// Device copies of the mesh for post-processing and the node -> element
// list, built on first use.
int post_setup(pp_plan* pl) {
  if (pl->post_adj.p) return PP_OK;
  hipStream_t s = pl->ctx->stream;
  const int64_t nn = pl->n_nodes, ne = pl->n_elems;
  const int npe = pl->npe;
  std::vector<int64_t> off(nn + 1, 0);
  for (int64_t e = 0; e < ne; e++)
    for (int i = 0; i < npe; i++) off[pl->h_elems[e * npe + i] + 1]++;
  for (int64_t n = 0; n < nn; n++) off[n + 1] += off[n];
  std::vector<uint32_t> adj(off[nn]);
  std::vector<int64_t> fill(off.begin(), off.end() - 1);
  for (int64_t e = 0; e < ne; e++)  // ascending element order per node
    for (int i = 0; i < npe; i++) adj[fill[pl->h_elems[e * npe + i]]++] = (uint32_t)e;
  int rc;
  if ((rc = pl->post_nodes.alloc(2 * nn)) || (rc = pl->post_elems.alloc((size_t)ne * npe)) ||
      (rc = pl->post_off.alloc(nn + 1)) || (rc = pl->post_adj.alloc(std::max<size_t>(1, adj.size()))) ||
      (rc = pl->post_em.alloc(4 * (size_t)std::max<int64_t>(1, ne))) ||
      (rc = pl->post_out.alloc(3 * nn)) || (rc = pl->post_red.alloc(2 * 1024 + 2)))
    return rc;
  PP_HIP(hipMemcpyAsync(pl->post_nodes.p, pl->h_nodes.data(), 2 * nn * 4, hipMemcpyHostToDevice, s));
  PP_HIP(hipMemcpyAsync(pl->post_elems.p, pl->h_elems.data(), (size_t)ne * npe * 4,
                        hipMemcpyHostToDevice, s));
  PP_HIP(hipMemcpyAsync(pl->post_off.p, off.data(), (nn + 1) * 8, hipMemcpyHostToDevice, s));
  if (!adj.empty())
    PP_HIP(hipMemcpyAsync(pl->post_adj.p, adj.data(), adj.size() * 4, hipMemcpyHostToDevice, s));
  PP_HIP(hipStreamSynchronize(s));
  return PP_OK;
}

// The DOF vector to post-process: the caller's (copied to the device) or
// the plan's solution.
int post_source(pp_plan* pl, const double* x64, const double** dx) {
  if (!x64) {
    *dx = pl->x.p;
    return PP_OK;
  }
  PP_HIP(hipMemcpyAsync(pl->tmp.p, x64, pl->n_dof * 8, hipMemcpyHostToDevice, pl->ctx->stream));
  *dx = pl->tmp.p;
  return PP_OK;
}

int post_moments_device(pp_plan* pl, const double* dx) {
  int rc = post_setup(pl);
  if (rc) return rc;
  hipStream_t s = pl->ctx->stream;
  const double Dfac = pl->E / (12.0 * (1.0 - pl->nu * pl->nu));
  const float* et = pl->var_t ? pl->elem_t.p : nullptr;
  if (pl->npe == 4)
    hipLaunchKernelGGL(k_elem_moments<4>, dim3(nblk(pl->n_elems)), dim3(kBlock), 0, s,
                       (int64_t)pl->n_elems, (const uint32_t*)pl->post_elems.p,
                       (const float*)pl->post_nodes.p, dx, Dfac, pl->t, et, pl->nu, pl->post_em.p);
  else
    hipLaunchKernelGGL(k_elem_moments<3>, dim3(nblk(pl->n_elems)), dim3(kBlock), 0, s,
                       (int64_t)pl->n_elems, (const uint32_t*)pl->post_elems.p,
                       (const float*)pl->post_nodes.p, dx, Dfac, pl->t, et, pl->nu, pl->post_em.p);
  hipLaunchKernelGGL(k_node_moments, dim3(nblk(pl->n_nodes)), dim3(kBlock), 0, s,
                     (int64_t)pl->n_nodes, (const int64_t*)pl->post_off.p,
                     (const uint32_t*)pl->post_adj.p, (const double*)pl->post_em.p,
                     pl->post_out.p);
  PP_HIP(hipGetLastError());
  return PP_OK;
}

}  // namespace
}  // namespace pp

using namespace pp;

extern "C" {

int pp_moments(pp_plan* pl, const double* x64, float* Mx, float* My, float* Mxy) {
  if (!pl) return fail(PP_ERR_INVALID, "NULL plan");
  PP_LOCK(pl->ctx);
  PP_HIP(hipSetDevice(pl->ctx->device));
  const double* dx = nullptr;
  int rc;
  if ((rc = post_setup(pl)) || (rc = post_source(pl, x64, &dx)) ||
      (rc = post_moments_device(pl, dx)))
    return rc;
  hipStream_t s = pl->ctx->stream;
  const int64_t nn = pl->n_nodes;
  float* out[3] = {Mx, My, Mxy};
  for (int k = 0; k < 3; k++)
    if (out[k])
      PP_HIP(hipMemcpyAsync(out[k], pl->post_out.p + k * nn, nn * 4, hipMemcpyDeviceToHost, s));
  PP_HIP(hipStreamSynchronize(s));
  return PP_OK;
}

int pp_max_displacement(pp_plan* pl, const double* x64, double* w, int64_t* node, float* x,
                        float* y) {
  if (!pl) return fail(PP_ERR_INVALID, "NULL plan");
  PP_LOCK(pl->ctx);
  if (pl->n_nodes <= 0) return fail(PP_ERR_INVALID, "plan has no nodes");
  PP_HIP(hipSetDevice(pl->ctx->device));
  const double* dx = nullptr;
  int rc;
  if ((rc = post_setup(pl)) || (rc = post_source(pl, x64, &dx))) return rc;
  hipStream_t s = pl->ctx->stream;
  const int64_t nn = pl->n_nodes;
  const int parts = (int)std::min<int64_t>(1024, (nn + kBlock - 1) / kBlock);
  double* pv = pl->post_red.p;
  int64_t* pi = (int64_t*)(pl->post_red.p + 1024);
  double* res = pl->post_red.p + 2048;
  hipLaunchKernelGGL(k_absmax_w, dim3(parts), dim3(kBlock), 0, s, nn, dx, pv, pi);
  hipLaunchKernelGGL(k_absmax_final, dim3(1), dim3(kBlock), 0, s, parts, (const double*)pv,
                     (const int64_t*)pi, nn, dx, res);
  PP_HIP(hipGetLastError());
  double h[2];
  PP_HIP(hipMemcpyAsync(h, res, 16, hipMemcpyDeviceToHost, s));
  PP_HIP(hipStreamSynchronize(s));
  const int64_t i = (int64_t)h[1];
  if (w) *w = h[0];
  if (node) *node = i;
  if (x) *x = pl->h_nodes[2 * i];
  if (y) *y = pl->h_nodes[2 * i + 1];
  return PP_OK;
}

}  // extern "C"
