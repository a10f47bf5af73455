// Coloured (general-mesh) path: per-element FP64 stiffness stored in HBM and
// applied one colour at a time, thread = element, race-free scatter because
// no two elements of a colour share a node (the scheme of
// gpu/solver.ts:413-461 + apply_k_q4.wgsl:133-350).  Used for meshes that are
// not tensor grids (triangles, general quads) and as the parity path that
// mirrors the reference's own scatter order.
//
// The PCG loop is driven from the host without per-iteration read-back: all
// scalars (alpha, beta, |r|, the iteration counter and the stop flag) live on
// the device, every kernel returns early once the stop flag is set, and the
// host only polls the flag once per batch of iterations.

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
      (rc = nodes.alloc(2 * nn)) || (rc = doff.alloc(nn + 1)) || (rc = dadj.alloc(adj.size())))
    return rc;
  hipStream_t s = pl->ctx->stream;
  PP_HIP(hipMemcpyAsync(nodes.p, pl->h_nodes.data(), 2 * nn * 4, hipMemcpyHostToDevice, s));
  PP_HIP(hipMemcpyAsync(doff.p, off.data(), (nn + 1) * 8, hipMemcpyHostToDevice, s));
  PP_HIP(hipMemcpyAsync(dadj.p, adj.data(), adj.size() * 4, hipMemcpyHostToDevice, s));
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

// Host-driven device-resident PCG (pcg.ts:98-215), no per-iteration read-back.
int general_solve_fp64(pp_plan* pl, const pp_solve_opts* o, pp_solve_stats* st) {
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
