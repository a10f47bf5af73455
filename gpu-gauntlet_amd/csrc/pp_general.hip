// Coloured (general-mesh) path: per-element FP64 stiffness stored in HBM and
// applied one colour at a time, thread = element, race-free scatter because
// no two elements of a colour share a node (the scheme of
// gpu/solver.ts:413-461 + apply_k_q4.wgsl:133-350).  Used for meshes that are
// not tensor grids (triangles, general quads) and as the parity path that
// mirrors the reference's own scatter order.
//
// The FP64 PCG runs as one persistent kernel (k_gen_pcg): K x gathered per
// node from the node's 3x3 blocks of K, assembled at set-up from the element
// matrices and stored sliced-ELL (64 nodes per slice, coalesced), with a
// grid barrier between the preconditioner and the gather and one all-reduce
// per iteration.  The launch-per-operation loop (PP_FLAG_COLOURED_HOST_LOOP)
// keeps all scalars (alpha, beta, |r|, the iteration counter and the stop
// flag) on the device, every kernel returning early once the stop flag is
// set, and the host polls the flag once per batch of iterations.

#include <math.h>

#include <algorithm>

#include "pp_device.h"
#include "pp_ke.h"

namespace pp {

namespace {

constexpr int kRedBlocks = 512;  // fixed partial count: deterministic sums
constexpr int kBatch = 64;       // iterations enqueued between flag polls

// device scalar slots
enum { S_RZ = 0, S_PAP, S_ALPHA, S_BETA, S_RR, S_BNORM, S_ITER, S_DONE, S_CONV, S_BREAK, S_N };

// ----------------------------------------------------- element stiffness
// (evaluated per element on the device at plan set-up; also host-callable;
// the Q4 element is in pp_ke.h)

// The reference's 9x9 triangle (element.ts:642-943) exactly as written,
// including its signed 1/(2A) and its B columns.
__host__ __device__ void ke_dkt(const double* xy, double E, double nu, double t, double* Ke) {
  for (int k = 0; k < 81; k++) Ke[k] = 0.0;
  const double x12 = xy[0] - xy[2], y12 = xy[1] - xy[3];
  const double x23 = xy[2] - xy[4], y23 = xy[3] - xy[5];
  const double x31 = xy[4] - xy[0], y31 = xy[5] - xy[1];
  const double twoA = x12 * y31 - x31 * y12;
  const double area = fabs(twoA) / 2;
  if (area < 1e-15) return;
  const double D = E * t * t * t / (12.0 * (1.0 - nu * nu));
  const double D11 = D, D12 = D * nu, D33 = D * (1.0 - nu) / 2.0;
  // edges k = 4 (2-3), 5 (3-1), 6 (1-2)
  const double ex[3] = {x23, x31, x12}, ey[3] = {y23, y31, y12};
  double a[3], b[3], c[3], d[3], e[3];
  for (int k = 0; k < 3; k++) {
    const double L = ex[k] * ex[k] + ey[k] * ey[k];
    a[k] = -ex[k] / L;
    b[k] = 3 * ex[k] * ey[k] / (4 * L);
    c[k] = (ex[k] * ex[k] - 2 * ey[k] * ey[k]) / (4 * L);
    d[k] = -ey[k] / L;
    e[k] = (ey[k] * ey[k] - 2 * ex[k] * ex[k]) / (4 * L);
  }
  const double P4 = -6 * a[0], P5 = -6 * a[1], P6 = -6 * a[2];
  const double t4 = -6 * d[0], t5 = -6 * d[1], t6 = -6 * d[2];
  const double q4 = 3 * a[0], q5 = 3 * a[1], q6 = 3 * a[2];
  const double r4 = 3 * d[0], r5 = 3 * d[1], r6 = 3 * d[2];
  const double b4 = b[0], b5 = b[1], b6 = b[2], c4 = c[0], c5 = c[1], c6 = c[2];
  const double e4 = e[0], e5 = e[1], e6 = e[2];
  const double gL[3][2] = {{1.0 / 6, 1.0 / 6}, {2.0 / 3, 1.0 / 6}, {1.0 / 6, 2.0 / 3}};
  for (int gq = 0; gq < 3; gq++) {
    const double L1 = gL[gq][0], L2 = gL[gq][1];
    const double m1 = 1 - 2 * L1, m2 = 1 - 2 * L2;
    const double hx1[9] = {P6 * m1 + (P5 - P6) * L2, q6 * m1 - (q5 + q6) * L2,
                           -4 + 6 * (L1 + L2) + (c5 - c6) * L2 + c6 * m1,
                           P4 * L2 + P6 * m1, q4 * L2 - q6 * m1, c4 * L2 - c6 * m1,
                           -P5 * L2 - P6 * m1, q5 * L2 + q6 * m1, -c5 * L2 + c6 * m1};
    const double hx2[9] = {-P6 * m2 + (P4 - P5) * L1 - P5 * m2,
                           q6 * m2 - (q4 + q5) * L1 + q5 * m2,
                           -4 + 6 * (L1 + L2) + (c4 - c5) * L1 + c5 * m2,
                           P4 * m2 + P5 * L1, q4 * m2 - q5 * L1, c4 * m2 - c5 * L1,
                           -P4 * L1 - P5 * m2, q4 * L1 + q5 * m2, -c4 * L1 + c5 * m2};
    const double hy1[9] = {t6 * m1 + (t5 - t6) * L2, 1 + r6 * m1 - (r5 + r6) * L2,
                           -e6 * m1 + (b5 - b6) * L2,
                           t4 * L2 + t6 * m1, -1 + r4 * L2 - r6 * m1, -e4 * L2 - b6 * m1,
                           -t5 * L2 - t6 * m1, r5 * L2 + r6 * m1, e5 * L2 + b6 * m1};
    const double hy2[9] = {-t6 * m2 + (t4 - t5) * L1 - t5 * m2,
                           1 + r6 * m2 - (r4 + r5) * L1 + r5 * m2,
                           e6 * m2 + (b4 - b5) * L1 - b5 * m2,
                           t4 * m2 + t5 * L1, -1 + r4 * m2 - r5 * L1, -e4 * m2 + b5 * L1,
                           -t4 * L1 - t5 * m2, r4 * L1 + r5 * m2, e4 * L1 - b5 * m2};
    const double inv2A = 1 / twoA;
    double B[3][9], DB[3][9];
    for (int j = 0; j < 9; j++) {
      B[0][j] = inv2A * (-y23 * hx1[j] - y31 * hx2[j]);
      B[1][j] = inv2A * (x23 * hy1[j] + x31 * hy2[j]);
      B[2][j] = inv2A * (x23 * hx1[j] + x31 * hx2[j] - y23 * hy1[j] - y31 * hy2[j]);
    }
    for (int j = 0; j < 9; j++) {
      DB[0][j] = D11 * B[0][j] + D12 * B[1][j];
      DB[1][j] = D12 * B[0][j] + D11 * B[1][j];
      DB[2][j] = D33 * B[2][j];
    }
    const double f = area / 3.0;
    for (int i = 0; i < 9; i++)
      for (int j = 0; j < 9; j++)
        Ke[i * 9 + j] += f * (B[0][i] * DB[0][j] + B[1][i] * DB[1][j] + B[2][i] * DB[2][j]);
  }
}

// ---------------------------------------------------------------- kernels

template <int NPE>
__global__ void k_colour(const uint32_t* list, int count, const uint32_t* elems, const double* Ke,
                         const uint8_t* cdof, const double* x, double* y, int skip_fixed,
                         const double* sc) {
  if (sc && sc[S_DONE] != 0.0) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  constexpr int ND = 3 * NPE;
  const uint32_t e = list[k];
  int64_t dof[ND];
  double u[ND];
#pragma unroll
  for (int i = 0; i < NPE; i++)
#pragma unroll
    for (int d = 0; d < 3; d++) dof[3 * i + d] = (int64_t)elems[(size_t)e * NPE + i] * 3 + d;
#pragma unroll
  for (int j = 0; j < ND; j++) u[j] = x[dof[j]];
  const double* K = Ke + (size_t)e * ND * ND;
#pragma unroll
  for (int r = 0; r < ND; r++) {
    double s = 0;
#pragma unroll
    for (int c = 0; c < ND; c++) s += K[r * ND + c] * u[c];
    if (!(skip_fixed && cdof[dof[r]])) y[dof[r]] += s;
  }
}

__global__ void k_zero(int64_t n, double* y, const double* sc) {
  if (sc && sc[S_DONE] != 0.0) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = 0.0;
}

__global__ void k_bc_rows(int64_t n, const uint8_t* cdof, const double* x, double* y,
                          const double* sc) {
  if (sc && sc[S_DONE] != 0.0) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && cdof[i]) y[i] = x[i];
}

__global__ void k_precond6(int64_t nn, const double* minv, const double* r, double* z,
                           const double* sc) {
  if (sc && sc[S_DONE] != 0.0) return;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  const double rv[3] = {r[3 * n], r[3 * n + 1], r[3 * n + 2]};
  double zv[3];
  sym3_apply(minv + 6 * n, rv, zv);
  for (int d = 0; d < 3; d++) z[3 * n + d] = zv[d];
}

// partial sums of a.b (and c.d when c != nullptr) over a fixed grid
__global__ __launch_bounds__(256) void k_dot2(int64_t n, const double* a, const double* b,
                                              const double* c, const double* d, double* part,
                                              const double* sc) {
  if (sc && sc[S_DONE] != 0.0) return;
  __shared__ double red[16];
  double v[2] = {0, 0};
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    v[0] += a[i] * b[i];
    if (c) v[1] += c[i] * d[i];
  }
  block_sum<2, 256>(v, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = v[0];
    part[kRedBlocks + blockIdx.x] = v[1];
  }
}

__device__ double sum_parts(const double* part, double* red) {
  double v[1] = {0};
  for (int i = threadIdx.x; i < kRedBlocks; i += 256) v[0] += part[i];
  block_sum<1, 256>(v, red);
  return v[0];
}

// init: rz = r.z, bnorm = |F|
__global__ __launch_bounds__(256) void k_fin_init(const double* part, double* sc) {
  __shared__ double red[16];
  const double rz = sum_parts(part, red);
  const double ff = sum_parts(part + kRedBlocks, red);
  if (threadIdx.x == 0) {
    sc[S_RZ] = rz;
    sc[S_RR] = ff;
    sc[S_BNORM] = sqrt(ff);
    sc[S_ITER] = 0;
    sc[S_CONV] = 0;
    sc[S_BREAK] = 0;
    sc[S_DONE] = sqrt(ff) < 1e-30 ? 1.0 : 0.0;
    if (sqrt(ff) < 1e-30) { sc[S_CONV] = 1; sc[S_RR] = 0; }
  }
}

// alpha = rz / pAp, breakdown test (pcg.ts:151-164)
__global__ __launch_bounds__(256) void k_fin_pap(const double* part, double* sc, double maxit) {
  if (sc[S_DONE] != 0.0) return;
  __shared__ double red[16];
  const double pap = sum_parts(part, red);
  if (threadIdx.x == 0) {
    sc[S_PAP] = pap;
    if (fabs(pap) < 1e-30) {
      sc[S_ITER] += 1;
      sc[S_BREAK] = 1;
      sc[S_DONE] = 1;
    } else {
      sc[S_ALPHA] = sc[S_RZ] / pap;
    }
  }
}

// x += alpha p; r -= alpha Ap; z = M^-1 r; partial r.z, r.r
__global__ __launch_bounds__(256) void k_update_xrz(int64_t nn, double* x, double* r, double* z,
                                                    const double* p, const double* Ap,
                                                    const double* minv, double* part,
                                                    const double* sc) {
  if (sc[S_DONE] != 0.0) return;
  __shared__ double red[16];
  const double alpha = sc[S_ALPHA];
  double v[2] = {0, 0};
  for (int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < nn; n += (int64_t)gridDim.x * 256) {
    double rv[3], zv[3];
    for (int d = 0; d < 3; d++) {
      x[3 * n + d] += alpha * p[3 * n + d];
      rv[d] = r[3 * n + d] - alpha * Ap[3 * n + d];
      r[3 * n + d] = rv[d];
    }
    sym3_apply(minv + 6 * n, rv, zv);
    for (int d = 0; d < 3; d++) {
      z[3 * n + d] = zv[d];
      v[0] += rv[d] * zv[d];
      v[1] += rv[d] * rv[d];
    }
  }
  block_sum<2, 256>(v, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = v[0];
    part[kRedBlocks + blockIdx.x] = v[1];
  }
}

// convergence (pcg.ts:173-181) and beta (pcg.ts:194-198)
__global__ __launch_bounds__(256) void k_fin_rz(const double* part, double* sc, double tol,
                                                double maxit) {
  if (sc[S_DONE] != 0.0) return;
  __shared__ double red[16];
  const double rz = sum_parts(part, red);
  const double rr = sum_parts(part + kRedBlocks, red);
  if (threadIdx.x == 0) {
    sc[S_ITER] += 1;
    sc[S_RR] = rr;
    const double rn = sqrt(rr);
    if (rn < tol || rn / sc[S_BNORM] < tol) {
      sc[S_CONV] = 1;
      sc[S_DONE] = 1;
    } else if (sc[S_ITER] >= maxit) {
      sc[S_DONE] = 1;
    } else {
      sc[S_BETA] = rz / sc[S_RZ];
      sc[S_RZ] = rz;
    }
  }
}

__global__ void k_update_p(int64_t n, double* p, const double* z, const double* sc) {
  if (sc[S_DONE] != 0.0) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] + sc[S_BETA] * p[i];
}

__global__ void k_copy(int64_t n, const double* s, double* d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = s[i];
}

int nblk(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

int colour_pass(pp_plan* pl, int c, const double* dx, double* dy, bool skip_fixed,
                const double* sc) {
  const uint32_t off = pl->h_colour_off[c], cnt = pl->h_colour_off[c + 1] - off;
  if (cnt == 0) return PP_OK;
  hipStream_t s = pl->ctx->stream;
  if (pl->npe == 4)
    hipLaunchKernelGGL(k_colour<4>, dim3(nblk(cnt, 64)), dim3(64), 0, s, pl->by_colour.p + off,
                       (int)cnt, pl->elems.p, pl->ke.p, pl->cmask_dof.p, dx, dy, (int)skip_fixed, sc);
  else
    hipLaunchKernelGGL(k_colour<3>, dim3(nblk(cnt, 64)), dim3(64), 0, s, pl->by_colour.p + off,
                       (int)cnt, pl->elems.p, pl->ke.p, pl->cmask_dof.p, dx, dy, (int)skip_fixed, sc);
  PP_HIP(hipGetLastError());
  return PP_OK;
}

int apply_all(pp_plan* pl, const double* dx, double* dy, const double* sc) {
  hipStream_t s = pl->ctx->stream;
  hipLaunchKernelGGL(k_zero, dim3(nblk(pl->n_dof, 256)), dim3(256), 0, s, (int64_t)pl->n_dof, dy, sc);
  for (size_t c = 0; c + 1 < pl->h_colour_off.size(); c++) {
    int rc = colour_pass(pl, (int)c, dx, dy, false, sc);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_bc_rows, dim3(nblk(pl->n_dof, 256)), dim3(256), 0, s, (int64_t)pl->n_dof,
                     (const uint8_t*)pl->cmask_dof.p, dx, dy, sc);
  PP_HIP(hipGetLastError());
  return PP_OK;
}

// Ke of every element (one thread each), from the float32 coordinates as
// the reference's element routines read them.
template <int NPE>
__global__ __launch_bounds__(256) void k_elem_ke(int64_t ne, const uint32_t* elems,
                                                 const float* nodes, double E, double nu, double t,
                                                 const float* elem_t, double* Ke) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= ne) return;
  double xy[2 * NPE];
#pragma unroll
  for (int i = 0; i < NPE; i++) {
    const uint32_t n = elems[e * NPE + i];
    xy[2 * i] = (double)nodes[2 * (size_t)n];
    xy[2 * i + 1] = (double)nodes[2 * (size_t)n + 1];
  }
  const double te = elem_t ? (double)elem_t[e] : t;
  constexpr int nd = 3 * NPE;
  if constexpr (NPE == 4) ke_q4(xy, E, nu, te, Ke + e * nd * nd);
  else ke_dkt(xy, E, nu, te, Ke + e * nd * nd);
}

// 3x3 nodal diagonal blocks (solver.ts:188-225) summed over each node's
// (element, corner) list in ascending element order -- the order of the
// reference's element loop -- then the constrained-DOF rules and inverse of
// solver.ts:231-304 (one thread per node).
__global__ __launch_bounds__(256) void k_block_csr(int64_t nn, int npe, const int64_t* off,
                                                   const uint32_t* adj, const double* Ke,
                                                   const uint8_t* cmask, double* minv) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= nn) return;
  const int nd = 3 * npe;
  double A[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t k = off[n]; k < off[n + 1]; k++) {
    const uint32_t ec = adj[k];  // element * npe + corner
    const double* K = Ke + (size_t)(ec / npe) * nd * nd;
    const int c = (int)(ec % npe);
    for (int r = 0; r < 3; r++)
      for (int q = 0; q < 3; q++) A[r * 3 + q] += K[(3 * c + r) * nd + 3 * c + q];
  }
  block_inverse_sym(A, cmask[n], minv + 6 * n);
}

// Consistent nodal pressure load, one thread per node: the node's share of
// every adjacent element's quadrature (Q4 2x2 Gauss with |J|, triangles the
// 3-point rule with linear N and area/3), summed in element order and then
// quadrature order (the order of the oracle's element loop).  Rotation
// loads and constrained w are zero.
__device__ double pressure_value(double x, double y, double q0, const double* pt, int np) {
  double p = q0;
  for (int k = 0; k < np; k++) {
    const double dx = x - pt[4 * k + 1], dy = y - pt[4 * k + 2], s = pt[4 * k + 3];
    p += pt[4 * k] * exp(-(dx * dx + dy * dy) / (2 * s * s));
  }
  return p;
}

__global__ __launch_bounds__(256) void k_pressure_csr(int64_t nn, int npe, const int64_t* off,
                                                      const uint32_t* adj, const uint32_t* elems,
                                                      const float* nodes, const uint8_t* cmask,
                                                      double q0, const double* pt, int np,
                                                      float* F) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= nn) return;
  double f = 0;
  for (int64_t k = off[n]; k < off[n + 1]; k++) {
    const uint32_t ec = adj[k];
    const int64_t e = ec / npe;
    const int c = (int)(ec % npe);
    double xy[8];
    for (int i = 0; i < npe; i++) {
      const uint32_t m = elems[e * npe + i];
      xy[2 * i] = (double)nodes[2 * (size_t)m];
      xy[2 * i + 1] = (double)nodes[2 * (size_t)m + 1];
    }
    if (npe == 4) {
      const double g = 0.5773502691896257;
      const double gx[4] = {-g, g, g, -g}, gy[4] = {-g, -g, g, g};
      for (int q = 0; q < 4; q++) {
        const double xi = gx[q], eta = gy[q];
        const double N[4] = {(1 - xi) * (1 - eta) / 4, (1 + xi) * (1 - eta) / 4,
                             (1 + xi) * (1 + eta) / 4, (1 - xi) * (1 + eta) / 4};
        const double dxi[4] = {-(1 - eta) / 4, (1 - eta) / 4, (1 + eta) / 4, -(1 + eta) / 4};
        const double deta[4] = {-(1 - xi) / 4, -(1 + xi) / 4, (1 + xi) / 4, (1 - xi) / 4};
        double J11 = 0, J12 = 0, J21 = 0, J22 = 0, px = 0, py = 0;
        for (int i = 0; i < 4; i++) {
          J11 += dxi[i] * xy[2 * i]; J12 += dxi[i] * xy[2 * i + 1];
          J21 += deta[i] * xy[2 * i]; J22 += deta[i] * xy[2 * i + 1];
          px += N[i] * xy[2 * i]; py += N[i] * xy[2 * i + 1];
        }
        const double w = pressure_value(px, py, q0, pt, np) * fabs(J11 * J22 - J12 * J21);
        f += N[c] * w;
      }
    } else {
      const double gL[3][3] = {{1.0 / 6, 1.0 / 6, 2.0 / 3}, {2.0 / 3, 1.0 / 6, 1.0 / 6},
                               {1.0 / 6, 2.0 / 3, 1.0 / 6}};
      const double twoA = (xy[0] - xy[2]) * (xy[5] - xy[1]) - (xy[4] - xy[0]) * (xy[1] - xy[3]);
      const double area = fabs(twoA) / 2;
      for (int q = 0; q < 3; q++) {
        const double px = gL[q][0] * xy[0] + gL[q][1] * xy[2] + gL[q][2] * xy[4];
        const double py = gL[q][0] * xy[1] + gL[q][1] * xy[3] + gL[q][2] * xy[5];
        const double w = pressure_value(px, py, q0, pt, np) * area / 3.0;
        f += gL[q][c] * w;
      }
    }
  }
  F[3 * n] = (cmask[n] & PP_FIX_W) ? 0.0f : (float)f;
  F[3 * n + 1] = 0.0f;
  F[3 * n + 2] = 0.0f;
}

// ------------------------------------------------- persistent FP64 PCG
//
// The single-reduction (Chronopoulos-Gear) recurrence of pcg.ts:98-215 on
// the coloured path, one persistent launch:
//   A (per node):  p = u + b p, s = w + b s;  x += a p;  r -= a s;  u = M r
//   -- grid barrier: every node's u visible --
//   B (per node):  w = K u (gathered, y[c] = u[c] on constrained DOFs);
//                  partials r.u, w.u, r.r
//   -- all-reduce --  b = g'/g, pAp = d' - b g'/a, a = g'/pAp
// s = K p is carried, not recomputed.  u is the only vector other
// workgroups read (device-coherent stores and loads); x, r, p, s, w stay
// with the thread that owns the node.  Iteration count, stopping rule and
// breakdown rule as the launch-per-operation loop (pcg.ts:118-181).
constexpr int kGenNT = 512;

struct GenArgs {
  int64_t nn;
  const int64_t* slc;     // per 64-node slice: first slot, width (slots per node)
  const uint32_t* col;    // slot -> neighbour node; slot (slice, k, lane) at first + 64 k + lane
  const double* blk;      // 9 doubles per slot, entry j of slot (slice, k, lane) at
                          //   9 first + 64 (9 k + j) + lane
  const uint8_t* cmask;   // PP_FIX bits per node
  const double* minv;
  const double* F;
  const double* r0;       // restart: residual F - K x of the x held in x (else null)
  double bnorm_in;        // restart: |F|
  double *x, *r, *p, *s, *w, *u;
  double* part;           // 8 * grid tagged all-reduce slots, zeroed
  GridSync* gs;
  DevStats* st;
  double tol;
  int maxit;
};

// (K u) of node n's three rows: the node's blocks times the neighbours' u.
// The 64 lanes of a wave hold the 64 nodes of one slice (node blocks start
// at multiples of 64), so the slice's width and offset are wave-uniform and
// every block and column load is coalesced.  Padding slots carry a zero
// block and the node itself.
template <bool COH = true>
__device__ __forceinline__ void ell_kx(const int64_t* slc, const uint32_t* col, const double* blk,
                                       const double* u, int64_t n, double (&v)[3]) {
  const int sl = __builtin_amdgcn_readfirstlane((int)(n >> 6));
  const int l = (int)(n & 63);
  const int64_t first = slc[2 * (int64_t)sl];
  const int width = (int)slc[2 * (int64_t)sl + 1];
  const uint32_t* cp = col + first + l;
  const double* bp = blk + 9 * first + l;
  v[0] = v[1] = v[2] = 0.0;
#pragma unroll 4
  for (int k = 0; k < width; k++) {
    const uint32_t m = cp[64 * k];
    const double* b = bp + (size_t)576 * k;
    double bk[9];
#pragma unroll
    for (int j = 0; j < 9; j++) bk[j] = b[64 * j];
    const double* um = u + 3 * (size_t)m;
    const double u0 = COH ? ld_coh(um) : um[0], u1 = COH ? ld_coh(um + 1) : um[1],
                 u2 = COH ? ld_coh(um + 2) : um[2];
#pragma unroll
    for (int d = 0; d < 3; d++) v[d] += bk[3 * d] * u0 + bk[3 * d + 1] * u1 + bk[3 * d + 2] * u2;
  }
}
__device__ __forceinline__ void gen_kx(const GenArgs& A, int64_t n, double (&v)[3]) {
  ell_kx<true>(A.slc, A.col, A.blk, A.u, n, v);
}

// y = K x from the node blocks, y[c] = x[c] on constrained DOFs (one thread
// per node): the persistent solve's residual check in one launch.
__global__ __launch_bounds__(256) void k_ell_apply(int64_t nn, const int64_t* slc,
                                                   const uint32_t* col, const double* blk,
                                                   const uint8_t* cmask, const double* x,
                                                   double* y) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= nn) return;
  double v[3];
  ell_kx<false>(slc, col, blk, x, n, v);
  const unsigned fix = cmask[n];
#pragma unroll
  for (int d = 0; d < 3; d++) y[3 * n + d] = ((fix >> d) & 1u) ? x[3 * n + d] : v[d];
}

// Slots of a node's row whose columns a thread keeps in registers.
constexpr int kGenCols = 9;

__global__ __launch_bounds__(kGenNT) void k_gen_pcg(GenArgs A) {
  __shared__ double red[80];
  const unsigned me = blockIdx.x, nwg = gridDim.x;
  const int64_t n0 = (int64_t)me * kGenNT + threadIdx.x, step = (int64_t)nwg * kGenNT;
  unsigned epoch = 0;
  const bool timing = me == 0 && threadIdx.x == 0;
  unsigned long long tk[4] = {0, 0, 0, 0};
  auto allreduce = [&](double (&v)[3]) -> bool {
    block_sum_published<3, kGenNT>(v, red + 16);  // stores drain during the sum
    return grid_allreduce<3, 4>(v, A.part, A.gs, epoch, red, nwg, me);
  };
  auto barrier = [&]() -> bool {
    publish();
    __syncthreads();
    double z[1] = {0.0};
    return grid_allreduce<1, 4>(z, A.part, A.gs, epoch, red, nwg, me);
  };

  // The thread's first node n0 keeps its state (x, r, p, s, w, u), its
  // block inverse, constraint bits and the columns of its first kGenCols
  // slots in registers for the whole solve; further nodes (meshes larger
  // than one pass of the grid) go through memory.  Every u load of the
  // resident row is then independent of any other load.
  const bool res = n0 < A.nn;
  double X[3] = {0, 0, 0}, Rr[3] = {0, 0, 0}, Pp[3] = {0, 0, 0}, Ss[3] = {0, 0, 0};
  double W[3] = {0, 0, 0}, U[3] = {0, 0, 0}, Mi[6] = {0, 0, 0, 0, 0, 0};
  unsigned fix0 = 0;
  uint32_t rc[kGenCols];
  int64_t rfirst = 0;
  int rwidth = 0;
#pragma unroll
  for (int k = 0; k < kGenCols; k++) rc[k] = 0;
  if (res) {
    const int sl = __builtin_amdgcn_readfirstlane((int)(n0 >> 6));
    rfirst = A.slc[2 * (int64_t)sl];
    rwidth = (int)A.slc[2 * (int64_t)sl + 1];
#pragma unroll
    for (int k = 0; k < kGenCols; k++)
      rc[k] = k < rwidth ? A.col[rfirst + 64 * k + (n0 & 63)] : (uint32_t)n0;
#pragma unroll
    for (int k = 0; k < 6; k++) Mi[k] = A.minv[6 * n0 + k];
    fix0 = A.cmask[n0];
  }
  // K u of the resident node: the register columns' u and blocks need no
  // other load first, then the slots beyond kGenCols from memory; slots
  // past the width add nothing, so the sums equal gen_kx's
  auto kx_res = [&](double (&v)[3]) {
    const double* bp = A.blk + 9 * rfirst + (n0 & 63);
    v[0] = v[1] = v[2] = 0.0;
    // three slots at a time; slots past the row's width load slot 0's
    // block (clamped, branch-free) and use zero
#pragma unroll
    for (int k0 = 0; k0 < kGenCols; k0 += 3) {
      double um[3][3], bk[3][9];
#pragma unroll
      for (int q = 0; q < 3; q++) {
        const int k = k0 + q;
#pragma unroll
        for (int d = 0; d < 3; d++) um[q][d] = ld_coh(&A.u[3 * (size_t)rc[k] + d]);
        const double* b = bp + (size_t)576 * (k < rwidth ? k : 0);
#pragma unroll
        for (int j = 0; j < 9; j++) bk[q][j] = b[64 * j];
      }
#pragma unroll
      for (int q = 0; q < 3; q++) {
        const bool live = k0 + q < rwidth;
#pragma unroll
        for (int d = 0; d < 3; d++) {
          const double t = bk[q][3 * d] * um[q][0] + bk[q][3 * d + 1] * um[q][1] +
                           bk[q][3 * d + 2] * um[q][2];
          v[d] += live ? t : 0.0;
        }
      }
    }
    for (int k = kGenCols; k < rwidth; k++) {
      const uint32_t m = A.col[rfirst + 64 * k + (n0 & 63)];
      const double* b = bp + (size_t)576 * k;
      double bk[9];
#pragma unroll
      for (int j = 0; j < 9; j++) bk[j] = b[64 * j];
      const double u0 = ld_coh(&A.u[3 * (size_t)m]), u1 = ld_coh(&A.u[3 * (size_t)m + 1]),
                   u2 = ld_coh(&A.u[3 * (size_t)m + 2]);
#pragma unroll
      for (int d = 0; d < 3; d++) v[d] += bk[3 * d] * u0 + bk[3 * d + 1] * u1 + bk[3 * d + 2] * u2;
    }
  };
  // w = K u on the own nodes, partials r.u, w.u, r.r
  auto phase_b = [&](double (&acc)[3]) {
    acc[0] = acc[1] = acc[2] = 0.0;
    if (res) {
      double v[3];
      kx_res(v);
#pragma unroll
      for (int d = 0; d < 3; d++) {
        W[d] = ((fix0 >> d) & 1u) ? U[d] : v[d];
        acc[0] += Rr[d] * U[d];
        acc[1] += W[d] * U[d];
        acc[2] += Rr[d] * Rr[d];
      }
    }
    for (int64_t n = n0 + step; n < A.nn; n += step) {
      double v[3];
      gen_kx(A, n, v);
      const unsigned fix = A.cmask[n];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const double u = ld_coh(&A.u[3 * n + d]), r = A.r[3 * n + d];
        const double w = ((fix >> d) & 1u) ? u : v[d];
        A.w[3 * n + d] = w;
        acc[0] += r * u;
        acc[1] += w * u;
        acc[2] += r * r;
      }
    }
  };

  // r = F, x = 0 (restart: r = r0, x kept), u = M r
  const bool restart = A.r0 != nullptr;
  const double* src = restart ? A.r0 : A.F;
  if (res) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
      Rr[d] = src[3 * n0 + d];
      X[d] = restart ? A.x[3 * n0 + d] : 0.0;
    }
    sym3_apply(Mi, Rr, U);
#pragma unroll
    for (int d = 0; d < 3; d++) st_coh(&A.u[3 * n0 + d], U[d]);
  }
  for (int64_t n = n0 + step; n < A.nn; n += step) {
    double r[3], u[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      r[d] = src[3 * n + d];
      A.r[3 * n + d] = r[d];
      if (!restart) A.x[3 * n + d] = 0.0;
    }
    sym3_apply(A.minv + 6 * n, r, u);
#pragma unroll
    for (int d = 0; d < 3; d++) st_coh(&A.u[3 * n + d], u[d]);
  }
  if (!barrier()) return;
  double acc[3];
  phase_b(acc);
  if (!allreduce(acc)) return;
  double g = acc[0], pap = acc[1], rho = acc[2];
  const double bnorm = restart ? A.bnorm_in : sqrt(rho);
  int it = 0, converged = 0, breakdown = 0;
  if (restart ? (sqrt(rho) < A.tol || sqrt(rho) / bnorm < A.tol) : bnorm < 1e-30) {
    converged = 1;
    if (!restart) rho = 0.0;
  } else if (A.maxit > 0 && fabs(pap) < 1e-30) {
    breakdown = 1;
    it = 1;
  } else {
    double a = g / pap, b = 0.0;
    bool first = true;
    for (; it < A.maxit;) {
      const unsigned long long t0 = timing ? wall_clock64() : 0;
      // p = u + b p, s = w + b s; x += a p; r -= a s; u = M r
      if (res) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
          Pp[d] = first ? U[d] : fma(b, Pp[d], U[d]);
          Ss[d] = first ? W[d] : fma(b, Ss[d], W[d]);
          X[d] = fma(a, Pp[d], X[d]);
          Rr[d] = fma(-a, Ss[d], Rr[d]);
        }
        sym3_apply(Mi, Rr, U);
#pragma unroll
        for (int d = 0; d < 3; d++) st_coh(&A.u[3 * n0 + d], U[d]);
      }
      for (int64_t n = n0 + step; n < A.nn; n += step) {
        double r[3], u[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
          const size_t i = 3 * n + d;
          const double uo = ld_coh(&A.u[i]), wo = A.w[i];
          const double p = first ? uo : fma(b, A.p[i], uo);
          const double s = first ? wo : fma(b, A.s[i], wo);
          A.p[i] = p;
          A.s[i] = s;
          A.x[i] = fma(a, p, A.x[i]);
          r[d] = fma(-a, s, A.r[i]);
          A.r[i] = r[d];
        }
        sym3_apply(A.minv + 6 * n, r, u);
#pragma unroll
        for (int d = 0; d < 3; d++) st_coh(&A.u[3 * n + d], u[d]);
      }
      first = false;
      const unsigned long long t1 = timing ? wall_clock64() : 0;
      if (!barrier()) return;
      const unsigned long long t2 = timing ? wall_clock64() : 0;
      phase_b(acc);
      const unsigned long long t3 = timing ? wall_clock64() : 0;
      if (!allreduce(acc)) return;
      if (timing) {
        const unsigned long long t4 = wall_clock64();
        tk[0] += t1 - t0; tk[1] += t2 - t1; tk[2] += t3 - t2; tk[3] += t4 - t3;
      }
      it++;
      const double g1 = acc[0], d1 = acc[1];
      rho = acc[2];
      const double rn = sqrt(rho);
      if (rn < A.tol || rn / bnorm < A.tol) { converged = 1; break; }
      b = g1 / g;
      pap = d1 - b * g1 / a;
      if (fabs(pap) < 1e-30) { breakdown = 1; it++; break; }
      a = g1 / pap;
      g = g1;
    }
  }
  if (res)
#pragma unroll
    for (int d = 0; d < 3; d++) A.x[3 * n0 + d] = X[d];
  if (me == 0 && threadIdx.x == 0) {
    A.st->iterations = it;
    A.st->converged = converged;
    A.st->breakdown = breakdown;
    A.st->final_residual = sqrt(rho);
    A.st->bnorm = bnorm;
    for (int k = 0; k < 4; k++) A.st->ticks[k] = tk[k];
  }
}

// Clears the barrier words, the statistics and the all-reduce slots with
// stores (no memset node where a caller captures the stream).
__global__ void k_gen_zero(GridSync* gs, DevStats* st, double* part, int64_t n_part) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_part) part[i] = 0.0;
  if (i == 0) {
    *st = DevStats{};
    gs->count = 0;
    gs->abort = 0;
  }
}

// Node blocks of K for the persistent solve (one thread per node): for
// every (element, corner) of the node, in ascending element order, each
// corner block of the element's rows added into the slot of that corner's
// node.  The diagonal block is summed in the order k_block_csr sums it.
template <int NPE>
__global__ __launch_bounds__(256) void k_ell_fill(int64_t nn, const int64_t* off,
                                                  const uint32_t* adj, const uint32_t* elems,
                                                  const double* Ke, const int64_t* slc,
                                                  const uint32_t* col, double* blk) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= nn) return;
  constexpr int ND = 3 * NPE;
  const int64_t sl = n >> 6, l = n & 63;
  const int64_t first = slc[2 * sl];
  const int width = (int)slc[2 * sl + 1];
  for (int64_t k = off[n]; k < off[n + 1]; k++) {
    const uint32_t ec = adj[k];
    const uint32_t e = ec / NPE;
    const int c = (int)(ec - e * NPE);
    const double* K = Ke + (size_t)e * ND * ND;
    for (int i = 0; i < NPE; i++) {
      const uint32_t m = elems[(size_t)e * NPE + i];
      int q = 0;
      while (q < width - 1 && col[first + 64 * q + l] != m) q++;
      double* b = blk + 9 * first + 576 * (int64_t)q + l;
      for (int d = 0; d < 3; d++)
        for (int j = 0; j < 3; j++) b[64 * (3 * d + j)] += K[(3 * c + d) * ND + 3 * i + j];
    }
  }
}

}  // namespace

// Set-up on the device: every element's FP64 Ke, then the nodal block
// inverses from a node -> (element, corner) list built on the host (an
// O(elements) integer pass, kept for the pressure loads).
int general_setup(pp_plan* pl, const pp_colouring* col) {
  const int npe = pl->npe, nd = 3 * npe;
  const int64_t ne = pl->n_elems, nn = pl->n_nodes;
  std::vector<int64_t> off(nn + 1, 0);
  for (int64_t e = 0; e < ne; e++)
    for (int i = 0; i < npe; i++) off[pl->h_elems[(size_t)e * npe + i] + 1]++;
  for (int64_t n = 0; n < nn; n++) off[n + 1] += off[n];
  std::vector<uint32_t> adj(std::max<int64_t>(1, off[nn]));
  {
    std::vector<int64_t> fill(off.begin(), off.end() - 1);
    for (int64_t e = 0; e < ne; e++)
      for (int i = 0; i < npe; i++)
        adj[fill[pl->h_elems[(size_t)e * npe + i]]++] = (uint32_t)(e * npe + i);
  }
  // node-block rows of the persistent solve: each node's neighbours (the
  // nodes of its elements, itself included) in increasing order, in slices
  // of 64 nodes padded to the slice's longest row with the node itself
  const int64_t nsl = (nn + 63) / 64;
  std::vector<int64_t> slc(2 * std::max<int64_t>(1, nsl), 0);
  std::vector<uint32_t> ecol;
  {
    std::vector<std::vector<uint32_t>> nb(64);
    int64_t slots = 0;
    for (int64_t sl = 0; sl < nsl; sl++) {
      size_t width = 0;
      for (int l = 0; l < 64; l++) {
        const int64_t n = sl * 64 + l;
        std::vector<uint32_t>& v = nb[l];
        v.clear();
        if (n >= nn) continue;
        for (int64_t k = off[n]; k < off[n + 1]; k++) {
          const uint32_t e = adj[k] / (uint32_t)npe;
          for (int i = 0; i < npe; i++) v.push_back(pl->h_elems[(size_t)e * npe + i]);
        }
        if (v.empty()) v.push_back((uint32_t)n);
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        width = std::max(width, v.size());
      }
      slc[2 * sl] = slots;
      slc[2 * sl + 1] = (int64_t)width;
      ecol.resize((size_t)(slots + 64 * (int64_t)width));
      for (size_t k = 0; k < width; k++)
        for (int l = 0; l < 64; l++) {
          const int64_t n = sl * 64 + l;
          ecol[(size_t)slots + 64 * k + l] =
              k < nb[l].size() ? nb[l][k] : (uint32_t)(n < nn ? n : 0);
        }
      slots += 64 * (int64_t)width;
    }
    if (ecol.empty()) ecol.push_back(0);
  }
  std::vector<uint8_t> cdof((size_t)nn * 3);
  for (int64_t n = 0; n < nn; n++)
    for (int d = 0; d < 3; d++) cdof[(size_t)n * 3 + d] = (pl->h_cmask[n] >> d) & 1u;
  DevBuf<float>& nodes = pl->gen_nodes;
  DevBuf<int64_t>& doff = pl->gen_off;
  DevBuf<uint32_t>& dadj = pl->gen_adj;
  int rc;
  if ((rc = pl->ke.alloc((size_t)ne * nd * nd)) || (rc = pl->minv.alloc((size_t)nn * 6)) ||
      (rc = pl->cmask_dof.alloc(cdof.size())) || (rc = pl->elems.alloc(pl->h_elems.size())) ||
      (rc = pl->by_colour.alloc(std::max<size_t>(1, col->colour_offsets[col->colour_count]))) ||
      (rc = pl->scal.alloc(S_N)) || (rc = pl->partials.alloc(2 * kRedBlocks)) ||
      (rc = nodes.alloc(2 * nn)) || (rc = doff.alloc(nn + 1)) || (rc = dadj.alloc(adj.size())) ||
      (rc = pl->ell_slc.alloc(slc.size())) || (rc = pl->ell_col.alloc(ecol.size())) ||
      (rc = pl->ell_blk.alloc(9 * ecol.size())))
    return rc;
  hipStream_t s = pl->ctx->stream;
  PP_HIP(hipMemcpyAsync(nodes.p, pl->h_nodes.data(), 2 * nn * 4, hipMemcpyHostToDevice, s));
  PP_HIP(hipMemcpyAsync(doff.p, off.data(), (nn + 1) * 8, hipMemcpyHostToDevice, s));
  PP_HIP(hipMemcpyAsync(dadj.p, adj.data(), adj.size() * 4, hipMemcpyHostToDevice, s));
  PP_HIP(hipMemcpyAsync(pl->ell_slc.p, slc.data(), slc.size() * 8, hipMemcpyHostToDevice, s));
  PP_HIP(hipMemcpyAsync(pl->ell_col.p, ecol.data(), ecol.size() * 4, hipMemcpyHostToDevice, s));
  PP_HIP(hipMemsetAsync(pl->ell_blk.p, 0, pl->ell_blk.n * 8, s));
  PP_HIP(hipMemcpyAsync(pl->cmask_dof.p, cdof.data(), cdof.size(), hipMemcpyHostToDevice, s));
  PP_HIP(hipMemcpyAsync(pl->elems.p, pl->h_elems.data(), pl->h_elems.size() * 4,
                        hipMemcpyHostToDevice, s));
  const uint32_t total = col->colour_offsets[col->colour_count];
  if (total)
    PP_HIP(hipMemcpyAsync(pl->by_colour.p, col->elements_by_colour, (size_t)total * 4,
                          hipMemcpyHostToDevice, s));
  const float* et = pl->var_t ? pl->elem_t.p : nullptr;
  const dim3 ge((unsigned)std::max<int64_t>(1, (ne + 255) / 256)), gn((unsigned)((nn + 255) / 256));
  if (ne > 0) {
    if (npe == 4)
      hipLaunchKernelGGL(k_elem_ke<4>, ge, dim3(256), 0, s, ne, (const uint32_t*)pl->elems.p,
                         (const float*)nodes.p, pl->E, pl->nu, pl->t, et, pl->ke.p);
    else
      hipLaunchKernelGGL(k_elem_ke<3>, ge, dim3(256), 0, s, ne, (const uint32_t*)pl->elems.p,
                         (const float*)nodes.p, pl->E, pl->nu, pl->t, et, pl->ke.p);
  }
  hipLaunchKernelGGL(k_block_csr, gn, dim3(256), 0, s, nn, npe, (const int64_t*)doff.p,
                     (const uint32_t*)dadj.p, (const double*)pl->ke.p,
                     (const uint8_t*)pl->cmask.p, pl->minv.p);
  if (npe == 4)
    hipLaunchKernelGGL(k_ell_fill<4>, gn, dim3(256), 0, s, nn, (const int64_t*)doff.p,
                       (const uint32_t*)dadj.p, (const uint32_t*)pl->elems.p,
                       (const double*)pl->ke.p, (const int64_t*)pl->ell_slc.p,
                       (const uint32_t*)pl->ell_col.p, pl->ell_blk.p);
  else
    hipLaunchKernelGGL(k_ell_fill<3>, gn, dim3(256), 0, s, nn, (const int64_t*)doff.p,
                       (const uint32_t*)dadj.p, (const uint32_t*)pl->elems.p,
                       (const double*)pl->ke.p, (const int64_t*)pl->ell_slc.p,
                       (const uint32_t*)pl->ell_col.p, pl->ell_blk.p);
  PP_HIP(hipGetLastError());
  PP_HIP(hipStreamSynchronize(s));
  return PP_OK;
}

int general_apply_k(pp_plan* pl, const double* dx, double* dy, int colour, bool bc_rows) {
  if (colour < 0) return apply_all(pl, dx, dy, nullptr);
  hipStream_t s = pl->ctx->stream;
  hipLaunchKernelGGL(k_zero, dim3(nblk(pl->n_dof, 256)), dim3(256), 0, s, (int64_t)pl->n_dof, dy,
                     (const double*)nullptr);
  int rc = colour_pass(pl, colour, dx, dy, !bc_rows, nullptr);
  if (rc) return rc;
  if (bc_rows)
    hipLaunchKernelGGL(k_bc_rows, dim3(nblk(pl->n_dof, 256)), dim3(256), 0, s, (int64_t)pl->n_dof,
                       (const uint8_t*)pl->cmask_dof.p, dx, dy, (const double*)nullptr);
  PP_HIP(hipGetLastError());
  return PP_OK;
}

int general_apply_precond(pp_plan* pl, const double* dr, double* dz) {
  hipLaunchKernelGGL(k_precond6, dim3(nblk(pl->n_nodes, 256)), dim3(256), 0, pl->ctx->stream,
                     (int64_t)pl->n_nodes, (const double*)pl->minv.p, dr, dz, (const double*)nullptr);
  PP_HIP(hipGetLastError());
  return PP_OK;
}

int general_true_residual(pp_plan* pl, const double* dx, double* rel) {
  int rc = apply_all(pl, dx, pl->tmp.p, nullptr);
  if (rc) return rc;
  return launch_norm_ratio(pl, pl->F.p, pl->tmp.p, rel);
}

namespace {

// One persistent launch (k_gen_pcg) sized to what is co-resident: one
// 512-thread workgroup per CU at most, fewer on small meshes (cold start
// when r0 is null).  If the workgroups cannot all run at once (another
// process holds CUs) the barrier spin budget runs out and the call fails
// with PP_ERR_TIMEOUT.
int launch_gen_pcg(pp_plan* pl, const double* r0, double bnorm, double tol, int maxit,
                   DevStats* ds) {
  hipStream_t s = pl->ctx->stream;
  const int64_t nn = pl->n_nodes;
  int per_cu = 0;
  PP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gen_pcg, kGenNT, 0));
  const int64_t cap = (int64_t)std::min(per_cu, 1) * pl->ctx->cu_count;
  const int grid = (int)std::min<int64_t>((nn + kGenNT - 1) / kGenNT, cap);
  if (grid <= 0) return fail(PP_ERR_HIP, "persistent PCG kernel cannot be made resident");
  const size_t npart = (size_t)8 * grid;
  int rc;
  if (pl->partials.n < npart && (rc = pl->partials.alloc(npart))) return rc;
  GenArgs A;
  A.nn = nn;
  A.slc = (const int64_t*)pl->ell_slc.p;
  A.col = (const uint32_t*)pl->ell_col.p;
  A.blk = pl->ell_blk.p;
  A.cmask = (const uint8_t*)pl->cmask.p;
  A.minv = pl->minv.p;
  A.F = pl->F.p;
  A.r0 = r0;
  A.bnorm_in = bnorm;
  A.x = pl->x.p; A.r = pl->r.p; A.p = pl->p0.p; A.s = pl->Ap.p; A.w = pl->p1.p; A.u = pl->z.p;
  A.part = pl->partials.p;
  A.gs = pl->gsync.p;
  A.st = pl->dstats.p;
  A.tol = tol;
  A.maxit = maxit;
  hipLaunchKernelGGL(k_gen_zero, dim3((unsigned)((npart + 255) / 256)), dim3(256), 0, s, pl->gsync.p,
                     pl->dstats.p, pl->partials.p, (int64_t)npart);
  hipLaunchKernelGGL(k_gen_pcg, dim3(grid), dim3(kGenNT), 0, s, A);
  PP_HIP(hipGetLastError());
  GridSync gsh;
  PP_HIP(hipMemcpyAsync(ds, pl->dstats.p, sizeof(*ds), hipMemcpyDeviceToHost, s));
  PP_HIP(hipMemcpyAsync(&gsh, pl->gsync.p, sizeof(gsh), hipMemcpyDeviceToHost, s));
  PP_HIP(hipStreamSynchronize(s));
  if (gsh.abort) return fail(PP_ERR_TIMEOUT, "grid barrier spin limit reached");
  return PP_OK;
}

// The persistent solve, then |F - K x| / |F| in FP64 and, while it is above
// the tolerance, up to four restarts from r = F - K x (residual
// replacement), as the structured solver does.
int general_solve_persistent(pp_plan* pl, const pp_solve_opts* o, pp_solve_stats* st) {
  constexpr int kMaxRestarts = 4;
  hipStream_t s = pl->ctx->stream;
  PP_HIP(hipEventRecord(pl->ctx->ev0, s));
  DevStats ds;
  int rc = launch_gen_pcg(pl, nullptr, 0.0, o->tolerance, o->max_iterations, &ds);
  if (rc) return rc;
  const double bnorm = ds.bnorm;
  int total = ds.iterations;
  for (int k = 0; k < 4; k++) st->phase_us[k] = 0.01 * (double)ds.ticks[k];
  st->sweeps = 1;
  st->sweep_iterations[0] = ds.iterations;
  double rel = 0;
  if (ds.converged && bnorm > 0 && !(o->flags & PP_FLAG_NO_TRUE_RESIDUAL)) {
    for (int k = 0;; k++) {
      hipLaunchKernelGGL(k_ell_apply, dim3((unsigned)((pl->n_nodes + 255) / 256)), dim3(256), 0, s,
                         (int64_t)pl->n_nodes, (const int64_t*)pl->ell_slc.p,
                         (const uint32_t*)pl->ell_col.p, (const double*)pl->ell_blk.p,
                         (const uint8_t*)pl->cmask.p, (const double*)pl->x.p, pl->tmp.p);
      PP_HIP(hipGetLastError());
      double res = 0, fn = 0;
      if ((rc = launch_norms(pl, pl->F.p, pl->tmp.p, &res, &fn))) return rc;
      rel = fn > 0 ? res / fn : 0.0;
      st->sweep_residual[std::min(st->sweeps, PP_MAX_SWEEPS) - 1] = rel;
      if (rel <= o->tolerance || k == kMaxRestarts || total >= o->max_iterations) break;
      if ((rc = launch_residual(pl, pl->F.p, pl->tmp.p, pl->tmp.p))) return rc;
      rc = launch_gen_pcg(pl, pl->tmp.p, bnorm, o->tolerance, o->max_iterations - total, &ds);
      if (rc) return rc;
      total += ds.iterations;
      for (int k = 0; k < 4; k++) st->phase_us[k] += 0.01 * (double)ds.ticks[k];
      if (st->sweeps < PP_MAX_SWEEPS) st->sweep_iterations[st->sweeps] = ds.iterations;
      st->sweeps++;
      if (!ds.converged) break;
    }
  }
  PP_HIP(hipEventRecord(pl->ctx->ev1, s));
  PP_HIP(hipEventSynchronize(pl->ctx->ev1));
  float ms = 0;
  PP_HIP(hipEventElapsedTime(&ms, pl->ctx->ev0, pl->ctx->ev1));
  st->iterations = total;
  st->converged = ds.converged;
  st->breakdown = ds.breakdown;
  st->final_residual = ds.final_residual;
  st->true_rel_residual = rel;
  st->device_ms = ms;
  return PP_OK;
}

}  // namespace

// FP64 PCG on the coloured path (pcg.ts:98-215), no per-iteration read-back:
// the persistent kernel with residual replacement, or with
// PP_FLAG_COLOURED_HOST_LOOP the launch-per-operation loop (the recursive
// residual alone decides); both report the FP64 true residual unless
// PP_FLAG_NO_TRUE_RESIDUAL.
int general_solve_fp64(pp_plan* pl, const pp_solve_opts* o, pp_solve_stats* st) {
  if (!(o->flags & PP_FLAG_COLOURED_HOST_LOOP)) return general_solve_persistent(pl, o, st);
  hipStream_t s = pl->ctx->stream;
  const int64_t nd = pl->n_dof, nn = pl->n_nodes;
  double* sc = pl->scal.p;
  double* part = pl->partials.p;
  const int gb = nblk(nd, 256), gn = nblk(nn, 256);
  PP_HIP(hipEventRecord(pl->ctx->ev0, s));
  // x = 0, r = F, z = M^-1 r, p = z, rz = r.z, |F|
  PP_HIP(hipMemsetAsync(pl->x.p, 0, nd * 8, s));
  hipLaunchKernelGGL(k_copy, dim3(gb), dim3(256), 0, s, nd, (const double*)pl->F.p, pl->r.p);
  hipLaunchKernelGGL(k_precond6, dim3(gn), dim3(256), 0, s, nn, (const double*)pl->minv.p,
                     (const double*)pl->r.p, pl->z.p, (const double*)nullptr);
  hipLaunchKernelGGL(k_copy, dim3(gb), dim3(256), 0, s, nd, (const double*)pl->z.p, pl->p0.p);
  hipLaunchKernelGGL(k_dot2, dim3(kRedBlocks), dim3(256), 0, s, nd, (const double*)pl->r.p,
                     (const double*)pl->z.p, (const double*)pl->F.p, (const double*)pl->F.p, part,
                     (const double*)nullptr);
  hipLaunchKernelGGL(k_fin_init, dim3(1), dim3(256), 0, s, (const double*)part, sc);
  PP_HIP(hipGetLastError());
  const double maxit = o->max_iterations;
  double h_sc[S_N];
  for (int done_it = 0;;) {
    const int n_it = std::min(kBatch, o->max_iterations - done_it);
    for (int k = 0; k < n_it; k++) {
      int rc = apply_all(pl, pl->p0.p, pl->Ap.p, sc);
      if (rc) return rc;
      hipLaunchKernelGGL(k_dot2, dim3(kRedBlocks), dim3(256), 0, s, nd, (const double*)pl->p0.p,
                         (const double*)pl->Ap.p, (const double*)nullptr, (const double*)nullptr,
                         part, (const double*)sc);
      hipLaunchKernelGGL(k_fin_pap, dim3(1), dim3(256), 0, s, (const double*)part, sc, maxit);
      hipLaunchKernelGGL(k_update_xrz, dim3(kRedBlocks), dim3(256), 0, s, nn, pl->x.p, pl->r.p,
                         pl->z.p, (const double*)pl->p0.p, (const double*)pl->Ap.p,
                         (const double*)pl->minv.p, part, (const double*)sc);
      hipLaunchKernelGGL(k_fin_rz, dim3(1), dim3(256), 0, s, (const double*)part, sc,
                         o->tolerance, maxit);
      hipLaunchKernelGGL(k_update_p, dim3(gb), dim3(256), 0, s, nd, pl->p0.p,
                         (const double*)pl->z.p, (const double*)sc);
      PP_HIP(hipGetLastError());
    }
    done_it += n_it;
    PP_HIP(hipMemcpyAsync(h_sc, sc, sizeof(h_sc), hipMemcpyDeviceToHost, s));
    PP_HIP(hipStreamSynchronize(s));
    if (h_sc[S_DONE] != 0.0 || done_it >= o->max_iterations) break;
  }
  PP_HIP(hipEventRecord(pl->ctx->ev1, s));
  PP_HIP(hipEventSynchronize(pl->ctx->ev1));
  float ms = 0;
  PP_HIP(hipEventElapsedTime(&ms, pl->ctx->ev0, pl->ctx->ev1));
  st->iterations = (int)h_sc[S_ITER];
  st->converged = h_sc[S_CONV] != 0.0;
  st->breakdown = h_sc[S_BREAK] != 0.0;
  st->final_residual = sqrt(h_sc[S_RR]);
  st->device_ms = ms;
  st->sweeps = 1;
  st->sweep_iterations[0] = st->iterations;
  if (!(o->flags & PP_FLAG_NO_TRUE_RESIDUAL)) {
    double rel = 0;
    int rc = general_true_residual(pl, pl->x.p, &rel);
    if (rc) return rc;
    st->true_rel_residual = rel;
    st->sweep_residual[0] = rel;
  }
  return PP_OK;
}

int general_pressure(pp_plan* pl, double q0, const double* d_patches, int np, float* dF) {
  hipLaunchKernelGGL(k_pressure_csr, dim3((unsigned)((pl->n_nodes + 255) / 256)), dim3(256), 0,
                     pl->ctx->stream, (int64_t)pl->n_nodes, pl->npe,
                     (const int64_t*)pl->gen_off.p, (const uint32_t*)pl->gen_adj.p,
                     (const uint32_t*)pl->elems.p, (const float*)pl->gen_nodes.p,
                     (const uint8_t*)pl->cmask.p, q0, d_patches, np, dF);
  PP_HIP(hipGetLastError());
  return PP_OK;
}

// Element stiffness of one element (host, FP64), for the batched path.
void element_ke_q4(const double* xy, double E, double nu, double t, double* Ke) {
  ke_q4(xy, E, nu, t, Ke);
}
void element_ke_dkt(const double* xy, double E, double nu, double t, double* Ke) {
  ke_dkt(xy, E, nu, t, Ke);
}

}  // namespace pp
