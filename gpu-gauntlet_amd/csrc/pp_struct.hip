// Structured-grid (tensor mesh) kernels: element operator, block-Jacobi
// set-up and apply, pressure loads, and the persistent FP64 PCG.
//
// Layout in HBM: every DOF vector is node-major [w, thx, thy] doubles (the
// reference's numbering, so nothing is permuted at the boundary); the
// preconditioner is 6 doubles per node (upper triangle of the symmetric 3x3
// inverse); the geometry is two small arrays of column widths / row heights
// (a, 1/a) and (b, 1/b) because every element of a tensor grid built from
// float32 coordinates is an exact axis-aligned rectangle.
//
// K.p is computed per tile of TX x TY nodes owned by one workgroup: the tile
// plus a one-node halo of p is staged in LDS, every element touching the tile
// ((TX+1) x (TY+1)) is applied matrix-free (B^T D B p) into LDS, and each node
// then sums its four elements' contributions in a fixed order.  No atomics,
// no colour passes, no races: each node is written by its owning thread only.

#include <math.h>

#include "pp_device.h"

namespace pp {

namespace {

constexpr int TX = 15, TY = 15, NT = 256;  // tile of 15x15 nodes, 16x16 elements
constexpr int PW = TX + 2, PH = TY + 2, NP = PW * PH;
constexpr int EW = TX + 1, EH = TY + 1, NE = EW * EH;
static_assert(NE <= NT, "one element per thread");

struct KGrid {
  int nx, ny;
  const double* ax;    // 2 per column: a, 1/a
  const double* by;    // 2 per row: b, 1/b
  const double* xc;    // node x coordinates (nx+1)
  const double* yc;    // node y coordinates (ny+1)
  const float* elem_t; // nullptr = uniform t0
  double t0, Dfac, Gk, nu;
  const uint8_t* cmask;
};

__device__ __forceinline__ int tiles_x(const KGrid& G) { return (G.nx + 1 + TX - 1) / TX; }
__device__ __forceinline__ int tiles_total(const KGrid& G) {
  return tiles_x(G) * ((G.ny + 1 + TY - 1) / TY);
}

// Stage p = z + beta * pold (or p = src when pold == nullptr) for the tile
// plus halo into LDS (SoA: p_s[c * NP + idx]); outside the grid p = 0.
template <typename T, typename S>
__device__ __forceinline__ void stage_tile(const KGrid& G, int i0, int j0, const S* src,
                                           const S* pold, T beta, T* p_s) {
  for (int idx = threadIdx.x; idx < NP; idx += NT) {
    const int lx = idx % PW, ly = idx / PW;
    const int gi = i0 - 1 + lx, gj = j0 - 1 + ly;
    T v0 = 0, v1 = 0, v2 = 0;
    if (gi >= 0 && gi <= G.nx && gj >= 0 && gj <= G.ny) {
      const size_t n = (size_t)gj * (G.nx + 1) + gi;
      v0 = (T)src[3 * n];
      v1 = (T)src[3 * n + 1];
      v2 = (T)src[3 * n + 2];
      if (pold) {
        v0 += beta * (T)pold[3 * n];
        v1 += beta * (T)pold[3 * n + 1];
        v2 += beta * (T)pold[3 * n + 2];
      }
    }
    p_s[idx] = v0;
    p_s[NP + idx] = v1;
    p_s[2 * NP + idx] = v2;
  }
}

// Apply every element touching the tile; results to ye_s[k * NE + le].
template <typename T>
__device__ __forceinline__ void tile_elements(const KGrid& G, int i0, int j0, const T* p_s,
                                              T* ye_s) {
  const int le = threadIdx.x;
  if (le >= NE) return;
  const int ex = le % EW, ey = le / EW;
  const int gi = i0 - 1 + ex, gj = j0 - 1 + ey;
  T y[12];
  if (gi >= 0 && gi < G.nx && gj >= 0 && gj < G.ny) {
    const int c0 = ey * PW + ex, c1 = c0 + 1, c2 = c0 + PW + 1, c3 = c0 + PW;
    const T u[12] = {p_s[c0], p_s[NP + c0], p_s[2 * NP + c0], p_s[c1], p_s[NP + c1],
                     p_s[2 * NP + c1], p_s[c2], p_s[NP + c2], p_s[2 * NP + c2],
                     p_s[c3], p_s[NP + c3], p_s[2 * NP + c3]};
    const double t = G.elem_t ? (double)G.elem_t[(size_t)gj * G.nx + gi] : G.t0;
    const double D = G.Dfac * t * t * t;
    const T D11 = (T)D, D12 = (T)(D * G.nu), D33 = (T)(D * (1.0 - G.nu) / 2.0);
    const T Ds = (T)(G.Gk * t);
    q4_rect_apply<T>((T)G.ax[2 * gi], (T)G.by[2 * gj], (T)G.ax[2 * gi + 1], (T)G.by[2 * gj + 1],
                     D11, D12, D33, Ds, u, y);
  } else {
#pragma unroll
    for (int k = 0; k < 12; k++) y[k] = 0;
  }
#pragma unroll
  for (int k = 0; k < 12; k++) ye_s[k * NE + le] = y[k];
}

// Sum of the four elements around node (tx, ty) of the tile, fixed order.
template <typename T>
__device__ __forceinline__ void node_sum(int tx, int ty, const T* ye_s, T (&y)[3]) {
  const int sw = ty * EW + tx, se = sw + 1, ne = sw + EW + 1, nw = sw + EW;
#pragma unroll
  for (int d = 0; d < 3; d++)
    y[d] = ye_s[(6 + d) * NE + sw] + ye_s[(9 + d) * NE + se] + ye_s[d * NE + ne] +
           ye_s[(3 + d) * NE + nw];
}

// ------------------------------------------------------------ K . x

__global__ __launch_bounds__(NT) void k_apply(KGrid G, const double* x, double* y) {
  __shared__ double p_s[3 * NP];
  __shared__ double ye_s[12 * NE];
  const int tix = blockIdx.x % tiles_x(G), tiy = blockIdx.x / tiles_x(G);
  const int i0 = tix * TX, j0 = tiy * TY;
  stage_tile<double, double>(G, i0, j0, x, nullptr, 0.0, p_s);
  __syncthreads();
  tile_elements<double>(G, i0, j0, p_s, ye_s);
  __syncthreads();
  if (threadIdx.x < TX * TY) {
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int gi = i0 + tx, gj = j0 + ty;
    if (gi <= G.nx && gj <= G.ny) {
      double v[3];
      node_sum<double>(tx, ty, ye_s, v);
      const size_t n = (size_t)gj * (G.nx + 1) + gi;
      const unsigned fix = G.cmask[n];
      const int c = (ty + 1) * PW + tx + 1;
      for (int d = 0; d < 3; d++) y[3 * n + d] = (fix >> d) & 1u ? p_s[d * NP + c] : v[d];
    }
  }
}

// ------------------------------------------- block-Jacobi set-up/apply

// Diagonal 3x3 block of K at every node: each adjacent element contributes
// the (corner, corner) block of its Ke, obtained by applying the element
// operator to the three unit vectors of that corner (solver.ts:188-225),
// then the constrained-DOF rules and the inverse of solver.ts:231-304.
__global__ void k_block_inverse(KGrid G, double* minv) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nn = (int64_t)(G.nx + 1) * (G.ny + 1);
  if (n >= nn) return;
  const int i = (int)(n % (G.nx + 1)), j = (int)(n / (G.nx + 1));
  double A[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  // elements SW, SE, NE, NW of the node and the node's corner index in each
  const int ei[4] = {i - 1, i, i, i - 1}, ej[4] = {j - 1, j - 1, j, j};
  const int corner[4] = {2, 3, 0, 1};
  for (int k = 0; k < 4; k++) {
    const int gi = ei[k], gj = ej[k];
    if (gi < 0 || gi >= G.nx || gj < 0 || gj >= G.ny) continue;
    const double t = G.elem_t ? (double)G.elem_t[(size_t)gj * G.nx + gi] : G.t0;
    const double D = G.Dfac * t * t * t;
    for (int c = 0; c < 3; c++) {
      double u[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, y[12];
      u[3 * corner[k] + c] = 1.0;
      q4_rect_apply<double>(G.ax[2 * gi], G.by[2 * gj], G.ax[2 * gi + 1], G.by[2 * gj + 1], D,
                            D * G.nu, D * (1.0 - G.nu) / 2.0, G.Gk * t, u, y);
      for (int r = 0; r < 3; r++) A[r * 3 + c] += y[3 * corner[k] + r];
    }
  }
  // symmetrise (the matrix-free columns differ from the rows by rounding)
  const double s01 = 0.5 * (A[1] + A[3]), s02 = 0.5 * (A[2] + A[6]), s12 = 0.5 * (A[5] + A[7]);
  A[1] = A[3] = s01;
  A[2] = A[6] = s02;
  A[5] = A[7] = s12;
  block_inverse_sym(A, G.cmask[n], minv + 6 * n);
}

__global__ void k_precond(int64_t nn, const double* minv, const double* r, double* z) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  const double rv[3] = {r[3 * n], r[3 * n + 1], r[3 * n + 2]};
  double zv[3];
  sym3_apply(minv + 6 * n, rv, zv);
  z[3 * n] = zv[0];
  z[3 * n + 1] = zv[1];
  z[3 * n + 2] = zv[2];
}

// ------------------------------------------------------- pressure loads

__device__ double pressure_at(double x, double y, double q0, const double* pt, int np) {
  double p = q0;
  for (int k = 0; k < np; k++) {
    const double dx = x - pt[4 * k + 1], dy = y - pt[4 * k + 2], s = pt[4 * k + 3];
    p += pt[4 * k] * exp(-(dx * dx + dy * dy) / (2.0 * s * s));
  }
  return p;
}

// F_w(node) = sum over the adjacent elements of the 2x2 Gauss rule of
// N_corner * p * |J| (the rule of element.ts:29-41); rotation loads are zero
// and constrained DOFs get zero (solver.ts:400-407).
__global__ void k_pressure(KGrid G, double q0, const double* pt, int np, float* F) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nn = (int64_t)(G.nx + 1) * (G.ny + 1);
  if (n >= nn) return;
  const int i = (int)(n % (G.nx + 1)), j = (int)(n / (G.nx + 1));
  const int ei[4] = {i - 1, i, i, i - 1}, ej[4] = {j - 1, j - 1, j, j};
  const int corner[4] = {2, 3, 0, 1};
  const double g = 0.5773502691896257;
  const double gxi[4] = {-g, g, g, -g}, geta[4] = {-g, -g, g, g};
  const double cxi[4] = {-1, 1, 1, -1}, ceta[4] = {-1, -1, 1, 1};
  double f = 0;
  for (int k = 0; k < 4; k++) {
    const int gi = ei[k], gj = ej[k];
    if (gi < 0 || gi >= G.nx || gj < 0 || gj >= G.ny) continue;
    const double x0 = G.xc[gi], x1 = G.xc[gi + 1], y0 = G.yc[gj], y1 = G.yc[gj + 1];
    const double detJ = 0.25 * (x1 - x0) * (y1 - y0);
    for (int q = 0; q < 4; q++) {
      const double xi = gxi[q], eta = geta[q];
      const double px = 0.5 * (1 - xi) * x0 + 0.5 * (1 + xi) * x1;
      const double py = 0.5 * (1 - eta) * y0 + 0.5 * (1 + eta) * y1;
      const double N = 0.25 * (1 + cxi[corner[k]] * xi) * (1 + ceta[corner[k]] * eta);
      f += N * pressure_at(px, py, q0, pt, np) * fabs(detJ);
    }
  }
  const unsigned fix = G.cmask[n];
  F[3 * n] = (fix & PP_FIX_W) ? 0.0f : (float)f;
  F[3 * n + 1] = 0.0f;
  F[3 * n + 2] = 0.0f;
}

// -------------------------------------------------- persistent FP64 PCG

// One workgroup per tile when RES (all per-node state kept in registers
// across iterations; only p and z go through memory for the halos), or
// tiles strided over the grid when !RES (state in HBM).
//
// Iteration k (two grid all-reduces):
//   A: p = z + beta p_old staged with halo (p double-buffered), Ap = K p in
//      LDS, BC rows Ap = p, partial p.Ap           -> all-reduce -> alpha
//   B: x += alpha p, r -= alpha Ap, z = M^-1 r, partial r.z, r.r
//                                                  -> all-reduce -> beta, |r|
// which is pcg.ts:146-206 with the p update moved to the start of the next
// iteration.  Stopping rule |r| < tol or |r|/|F| < tol (pcg.ts:174).
struct PcgArgs {
  KGrid G;
  const double* F;
  const double* r0;  // restart: initial residual F - K x (x read from x); nullptr = cold start
  double bnorm_in;   // restart: |F| for the relative stopping test
  double *x, *r, *z, *pb0, *pb1, *Ap;
  const double* minv;
  double* part;
  GridSync* gs;
  DevStats* st;
  double tol;
  int maxit;
};

template <bool RES>
__global__ __launch_bounds__(NT) void k_pcg(PcgArgs A) {
  __shared__ double p_s[3 * NP];
  __shared__ double ye_s[12 * NE];
  __shared__ double red[64];
  const KGrid& G = A.G;
  const int ntx = tiles_x(G), ntiles = tiles_total(G);
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const bool node_thread = threadIdx.x < TX * TY;
  unsigned epoch = 0;

  // per-thread state of the owned node (RES only)
  double xr[3] = {0, 0, 0}, rr_[3] = {0, 0, 0}, zr[3] = {0, 0, 0}, pr[3] = {0, 0, 0};
  double apr[3] = {0, 0, 0}, mi[6] = {0, 0, 0, 0, 0, 0};
  int64_t my_n = -1;
  unsigned my_fix = 0;

  // ---------------- init: r = F (or r0 on a restart), z = M^-1 r, r.z, r.r
  const bool restart = A.r0 != nullptr;
  const double* src = restart ? A.r0 : A.F;
  double acc[2] = {0, 0};
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int i0 = (tile % ntx) * TX, j0 = (tile / ntx) * TY;
    const int gi = i0 + tx, gj = j0 + ty;
    if (node_thread && gi <= G.nx && gj <= G.ny) {
      const int64_t n = (int64_t)gj * (G.nx + 1) + gi;
      double rv[3] = {src[3 * n], src[3 * n + 1], src[3 * n + 2]}, zv[3], m[6];
      for (int k = 0; k < 6; k++) m[k] = A.minv[6 * n + k];
      sym3_apply(m, rv, zv);
      for (int d = 0; d < 3; d++) {
        A.z[3 * n + d] = zv[d];
        if (!RES) {
          A.r[3 * n + d] = rv[d];
          if (!restart) A.x[3 * n + d] = 0.0;
        }
        acc[0] += rv[d] * zv[d];
        acc[1] += rv[d] * rv[d];
      }
      if (RES) {
        my_n = n;
        my_fix = G.cmask[n];
        for (int d = 0; d < 3; d++) {
          rr_[d] = rv[d];
          zr[d] = zv[d];
          xr[d] = restart ? A.x[3 * n + d] : 0.0;
        }
        for (int k = 0; k < 6; k++) mi[k] = m[k];
      }
    }
  }
  block_sum<2, NT>(acc, red + 8);
  if (!grid_allreduce<2>(acc, A.part, A.gs, epoch, red)) return;
  double rz = acc[0], rr = acc[1];
  const double bnorm = restart ? A.bnorm_in : sqrt(rr);
  const bool done0 = restart ? (sqrt(rr) < A.tol || sqrt(rr) / bnorm < A.tol) : bnorm < 1e-30;
  if (done0) {
    if (RES && my_n >= 0)
      for (int d = 0; d < 3; d++) A.x[3 * my_n + d] = xr[d];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      A.st->iterations = 0; A.st->converged = 1; A.st->breakdown = 0;
      A.st->final_residual = restart ? sqrt(rr) : 0.0; A.st->bnorm = bnorm;
    }
    return;
  }

  double beta = 0;
  int it = 0, converged = 0, breakdown = 0;
  for (; it < A.maxit; it++) {
    const double* pold = (it & 1) ? A.pb1 : A.pb0;
    double* pnew = (it & 1) ? A.pb0 : A.pb1;
    // ---------------- phase A: p, Ap = K p, p.Ap
    double pap[1] = {0};
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
      const int i0 = (tile % ntx) * TX, j0 = (tile / ntx) * TY;
      stage_tile<double, double>(G, i0, j0, A.z, it == 0 ? nullptr : pold, beta, p_s);
      __syncthreads();
      tile_elements<double>(G, i0, j0, p_s, ye_s);
      __syncthreads();
      const int gi = i0 + tx, gj = j0 + ty;
      if (node_thread && gi <= G.nx && gj <= G.ny) {
        const int64_t n = (int64_t)gj * (G.nx + 1) + gi;
        const unsigned fix = RES ? my_fix : G.cmask[n];
        double v[3];
        node_sum<double>(tx, ty, ye_s, v);
        const int c = (ty + 1) * PW + tx + 1;
        for (int d = 0; d < 3; d++) {
          const double pd = p_s[d * NP + c];
          const double ad = ((fix >> d) & 1u) ? pd : v[d];
          pnew[3 * n + d] = pd;
          pap[0] += pd * ad;
          if (RES) { pr[d] = pd; apr[d] = ad; }
          else A.Ap[3 * n + d] = ad;
        }
      }
      __syncthreads();
    }
    block_sum<1, NT>(pap, red + 8);
    if (!grid_allreduce<1>(pap, A.part, A.gs, epoch, red)) return;
    if (fabs(pap[0]) < 1e-30) { breakdown = 1; it++; break; }
    const double alpha = rz / pap[0];

    // ---------------- phase B: x, r, z updates, r.z, r.r
    acc[0] = 0; acc[1] = 0;
    if (RES) {
      if (my_n >= 0) {
        for (int d = 0; d < 3; d++) {
          xr[d] += alpha * pr[d];
          rr_[d] -= alpha * apr[d];
        }
        sym3_apply(mi, rr_, zr);
        for (int d = 0; d < 3; d++) {
          A.z[3 * my_n + d] = zr[d];
          acc[0] += rr_[d] * zr[d];
          acc[1] += rr_[d] * rr_[d];
        }
      }
    } else {
      for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int i0 = (tile % ntx) * TX, j0 = (tile / ntx) * TY;
        const int gi = i0 + tx, gj = j0 + ty;
        if (node_thread && gi <= G.nx && gj <= G.ny) {
          const int64_t n = (int64_t)gj * (G.nx + 1) + gi;
          double rv[3], zv[3], m[6];
          for (int k = 0; k < 6; k++) m[k] = A.minv[6 * n + k];
          for (int d = 0; d < 3; d++) {
            A.x[3 * n + d] += alpha * pnew[3 * n + d];
            rv[d] = A.r[3 * n + d] - alpha * A.Ap[3 * n + d];
            A.r[3 * n + d] = rv[d];
          }
          sym3_apply(m, rv, zv);
          for (int d = 0; d < 3; d++) {
            A.z[3 * n + d] = zv[d];
            acc[0] += rv[d] * zv[d];
            acc[1] += rv[d] * rv[d];
          }
        }
      }
    }
    block_sum<2, NT>(acc, red + 8);
    if (!grid_allreduce<2>(acc, A.part, A.gs, epoch, red)) return;
    rr = acc[1];
    const double rn = sqrt(rr);
    if (rn < A.tol || rn / bnorm < A.tol) { converged = 1; it++; break; }
    beta = acc[0] / rz;
    rz = acc[0];
  }
  if (RES && my_n >= 0)
    for (int d = 0; d < 3; d++) A.x[3 * my_n + d] = xr[d];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    A.st->iterations = it;
    A.st->converged = converged;
    A.st->breakdown = breakdown;
    A.st->final_residual = sqrt(rr);
    A.st->bnorm = bnorm;
  }
}

__global__ void k_zero_sync(GridSync* gs) {
  if (threadIdx.x < 64) {
    gs->count = 0;
    gs->abort = 0;
  }
}

KGrid make_kgrid(const pp_plan* pl) {
  KGrid G;
  G.nx = pl->nx;
  G.ny = pl->ny;
  G.xc = pl->xc.p;
  G.yc = pl->yc.p;
  G.elem_t = pl->var_t ? pl->elem_t.p : nullptr;
  G.t0 = pl->t;
  G.Dfac = pl->E / (12.0 * (1.0 - pl->nu * pl->nu));
  G.Gk = (5.0 / 6.0) * pl->E / (2.0 * (1.0 + pl->nu));
  G.nu = pl->nu;
  G.cmask = pl->cmask.p;
  // column widths and row heights are stored after the coordinates
  G.ax = pl->xc.p + (pl->nx + 1);
  G.by = pl->yc.p + (pl->ny + 1);
  return G;
}

int grid_blocks(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace

int struct_setup(pp_plan* pl) {
  // coordinates + (a, 1/a) / (b, 1/b) per column / row, in double
  const int nx = pl->nx, ny = pl->ny;
  std::vector<double> hx((nx + 1) + 2 * nx), hy((ny + 1) + 2 * ny);
  for (int i = 0; i <= nx; i++) hx[i] = (double)pl->h_nodes[2 * i];
  for (int j = 0; j <= ny; j++) hy[j] = (double)pl->h_nodes[2 * ((size_t)j * (nx + 1)) + 1];
  for (int i = 0; i < nx; i++) {
    const double a = hx[i + 1] - hx[i];
    hx[nx + 1 + 2 * i] = a;
    hx[nx + 1 + 2 * i + 1] = 1.0 / a;
  }
  for (int j = 0; j < ny; j++) {
    const double b = hy[j + 1] - hy[j];
    hy[ny + 1 + 2 * j] = b;
    hy[ny + 1 + 2 * j + 1] = 1.0 / b;
  }
  int rc;
  if ((rc = pl->xc.alloc(hx.size())) || (rc = pl->yc.alloc(hy.size()))) return rc;
  hipStream_t s = pl->ctx->stream;
  PP_HIP(hipMemcpyAsync(pl->xc.p, hx.data(), hx.size() * 8, hipMemcpyHostToDevice, s));
  PP_HIP(hipMemcpyAsync(pl->yc.p, hy.data(), hy.size() * 8, hipMemcpyHostToDevice, s));
  if ((rc = pl->minv.alloc(6 * pl->n_nodes))) return rc;
  KGrid G = make_kgrid(pl);
  hipLaunchKernelGGL(k_block_inverse, dim3(grid_blocks(pl->n_nodes, 256)), dim3(256), 0, s, G,
                     pl->minv.p);
  PP_HIP(hipGetLastError());
  PP_HIP(hipStreamSynchronize(s));
  return PP_OK;
}

int struct_apply_k(pp_plan* pl, const double* dx, double* dy) {
  KGrid G = make_kgrid(pl);
  const int ntiles = ((pl->nx + 1 + TX - 1) / TX) * ((pl->ny + 1 + TY - 1) / TY);
  hipLaunchKernelGGL(k_apply, dim3(ntiles), dim3(NT), 0, pl->ctx->stream, G, dx, dy);
  PP_HIP(hipGetLastError());
  return PP_OK;
}

int struct_apply_precond(pp_plan* pl, const double* dr, double* dz) {
  hipLaunchKernelGGL(k_precond, dim3(grid_blocks(pl->n_nodes, 256)), dim3(256), 0,
                     pl->ctx->stream, (int64_t)pl->n_nodes, (const double*)pl->minv.p, dr, dz);
  PP_HIP(hipGetLastError());
  return PP_OK;
}

int struct_pressure(pp_plan* pl, double q0, const double* d_patches, int np, float* dF) {
  KGrid G = make_kgrid(pl);
  hipLaunchKernelGGL(k_pressure, dim3(grid_blocks(pl->n_nodes, 256)), dim3(256), 0,
                     pl->ctx->stream, G, q0, d_patches, np, dF);
  PP_HIP(hipGetLastError());
  return PP_OK;
}

int struct_true_residual(pp_plan* pl, const double* dx, double* rel) {
  int rc = struct_apply_k(pl, dx, pl->tmp.p);
  if (rc) return rc;
  return launch_norm_ratio(pl, pl->F.p, pl->tmp.p, rel);
}

// One cooperative launch of the persistent PCG (cold start when r0 is null).
static int launch_pcg(pp_plan* pl, const double* r0, double bnorm, double tol, int maxit,
                      DevStats* ds) {
  hipStream_t s = pl->ctx->stream;
  const int ntiles = ((pl->nx + 1 + TX - 1) / TX) * ((pl->ny + 1 + TY - 1) / TY);
  int per_cu = 0;
  PP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg<true>, NT, 0));
  const bool res = ntiles <= per_cu * pl->ctx->cu_count;
  int per_cu_s = 0;
  PP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_s, k_pcg<false>, NT, 0));
  const int nwg = res ? ntiles : std::min(ntiles, per_cu_s * pl->ctx->cu_count);
  if (nwg <= 0) return fail(PP_ERR_HIP, "persistent PCG kernel cannot be made resident");
  int rc;
  if (pl->partials.n < (size_t)4 * nwg && (rc = pl->partials.alloc(4 * (size_t)nwg))) return rc;
  PcgArgs A;
  A.G = make_kgrid(pl);
  A.F = pl->F.p;
  A.r0 = r0;
  A.bnorm_in = bnorm;
  A.x = pl->x.p; A.r = pl->r.p; A.z = pl->z.p; A.pb0 = pl->p0.p; A.pb1 = pl->p1.p;
  A.Ap = pl->Ap.p;
  A.minv = pl->minv.p;
  A.part = pl->partials.p;
  A.gs = pl->gsync.p;
  A.st = pl->dstats.p;
  A.tol = tol;
  A.maxit = maxit;
  void* args[] = {&A};
  hipLaunchKernelGGL(k_zero_sync, dim3(1), dim3(64), 0, s, pl->gsync.p);
  PP_HIP(hipGetLastError());
  if (res)
    PP_HIP(hipLaunchCooperativeKernel((const void*)k_pcg<true>, dim3(nwg), dim3(NT), args, 0, s));
  else
    PP_HIP(hipLaunchCooperativeKernel((const void*)k_pcg<false>, dim3(nwg), dim3(NT), args, 0, s));
  GridSync gsh;
  PP_HIP(hipMemcpyAsync(ds, pl->dstats.p, sizeof(*ds), hipMemcpyDeviceToHost, s));
  PP_HIP(hipMemcpyAsync(&gsh, pl->gsync.p, sizeof(gsh), hipMemcpyDeviceToHost, s));
  PP_HIP(hipStreamSynchronize(s));
  if (gsh.abort) return fail(PP_ERR_TIMEOUT, "grid barrier spin limit reached");
  return PP_OK;
}

// FP64 PCG with residual replacement: after the recurrence converges the
// true residual F - K x is recomputed in FP64; while it is above the
// tolerance (the recurrence drifts from it by rounding) PCG restarts from x
// with that residual, at most kMaxRestarts times.  All of it is timed.
int struct_solve_fp64(pp_plan* pl, const pp_solve_opts* o, pp_solve_stats* st) {
  constexpr int kMaxRestarts = 4;
  hipStream_t s = pl->ctx->stream;
  PP_HIP(hipEventRecord(pl->ctx->ev0, s));
  DevStats ds;
  int rc = launch_pcg(pl, nullptr, 0.0, o->tolerance, o->max_iterations, &ds);
  if (rc) return rc;
  const double bnorm = ds.bnorm;
  int total = ds.iterations;
  st->sweeps = 1;
  st->sweep_iterations[0] = ds.iterations;
  double rel = 0;
  if (ds.converged && bnorm > 0) {
    for (int k = 0;; k++) {
      if ((rc = struct_apply_k(pl, pl->x.p, pl->tmp.p))) return rc;
      if ((rc = launch_norm_ratio(pl, pl->F.p, pl->tmp.p, &rel))) return rc;
      st->sweep_residual[st->sweeps - 1] = rel;
      if (rel <= o->tolerance || k == kMaxRestarts || total >= o->max_iterations) break;
      if ((rc = launch_residual(pl, pl->F.p, pl->tmp.p, pl->tmp.p))) return rc;
      rc = launch_pcg(pl, pl->tmp.p, bnorm, o->tolerance, o->max_iterations - total, &ds);
      if (rc) return rc;
      total += ds.iterations;
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

}  // namespace pp
