// Structured-grid (tensor mesh) kernels: K.x, block-Jacobi set-up and
// apply, pressure loads, and the persistent FP64 PCG.
//
// Layout in HBM: every DOF vector is node-major [w, thx, thy] doubles (the
// reference's numbering, so nothing is permuted at the boundary); the
// preconditioner is 6 doubles per node (upper triangle of the symmetric 3x3
// inverse); the geometry is two small arrays of column widths / row heights.
//
// Persistent PCG kernels and their launch shapes:
//  - default while the plate is resident (the 61k-DOF plate: 64 tiles):
//    single-reduction PCG, one 384-thread workgroup per 18x18-node tile,
//    per-node state in registers;
//  - the recurrence order of pcg.ts (two reductions), resident on up to 64
//    tiles of 31x31 nodes or on 15x15-node tiles, state in registers;
//  - larger plates: 18x18-node tiles strided over one workgroup per CU,
//    state in HBM.

#include <math.h>

#include <algorithm>

#include "pp_grid.h"
#include "pp_load.h"

namespace pp {

using CfgS = Tile<15, 15, 256>;   // device-wide kernel and plain K.x
using CfgX = Tile<31, 31, 1024>;  // large-tile persistent kernel

KGrid make_kgrid(const pp_plan* pl) {
  KGrid G;
  G.nx = pl->nx;
  G.ny = pl->ny;
  G.xc = pl->xc.p;
  G.yc = pl->yc.p;
  G.ax = pl->xc.p + (pl->nx + 1);  // widths / heights are stored after the coordinates
  G.by = pl->yc.p + (pl->ny + 1);
  G.elem_t = pl->var_t ? pl->elem_t.p : nullptr;
  G.t0 = pl->t;
  G.Dfac = pl->E / (12.0 * (1.0 - pl->nu * pl->nu));
  G.Gk = (5.0 / 6.0) * pl->E / (2.0 * (1.0 + pl->nu));
  G.nu = pl->nu;
  G.cmask = pl->cmask.p;
  return G;
}

namespace {

int grid_blocks(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

// ------------------------------------------------------------ K . x

// y = K x with the reference's BC rows (solver.ts:479-482: y[c] = x[c]).
__global__ __launch_bounds__(CfgS::NT) void k_apply(KGrid G, const double* x, double* y) {
  using C = CfgS;
  __shared__ double p_s[3 * C::NP];
  __shared__ double ye_s[12 * C::NE];
  const int tile = xcd_band_tile((int)blockIdx.x, (int)gridDim.x);
  const int tix = tile % tiles_x<C>(G), tiy = tile / tiles_x<C>(G);
  const int i0 = tix * C::TX, j0 = tiy * C::TY;
  stage_tile<C, double, double>(G, i0, j0, x, nullptr, 0.0, p_s);
  __syncthreads();
  tile_elements<C, double>(G, i0, j0, p_s, ye_s);
  __syncthreads();
  if (threadIdx.x < C::TX * C::TY) {
    const int tx = threadIdx.x % C::TX, ty = threadIdx.x / C::TX;
    const int gi = i0 + tx, gj = j0 + ty;
    if (gi <= G.nx && gj <= G.ny) {
      double v[3];
      node_sum<C, double>(tx, ty, ye_s, v);
      const size_t n = (size_t)gj * (G.nx + 1) + gi;
      const unsigned fix = G.cmask[n];
      const int c = (ty + 1) * C::PW + tx + 1;
      for (int d = 0; d < 3; d++) y[3 * n + d] = (fix >> d) & 1u ? p_s[d * C::NP + c] : v[d];
    }
  }
}

// Fused multigrid sweep on the plan's operator: with (K x)_n as in k_apply,
// MODE 1: xo = x + c1 (x - xo) + c2 B (b - K x): a damped block-Jacobi sweep
// when c1 = 0 (xo is then not read), a Chebyshev step otherwise, xo holding
// the previous iterate (the increment x - x_prev is the last step's, so no
// increment vector is stored); xo must not alias x, whose halo other
// workgroups read.  MODE 2: xo = b - K x.
template <int MODE>
__global__ __launch_bounds__(CfgS::NT) void k_apply_sweep(KGrid G, const double* x,
                                                          const double* b, const float* minv,
                                                          double c1, double c2, double* xo) {
  using C = CfgS;
  __shared__ double p_s[3 * C::NP];
  __shared__ double ye_s[12 * C::NE];
  const int tile = xcd_band_tile((int)blockIdx.x, (int)gridDim.x);
  const int tix = tile % tiles_x<C>(G), tiy = tile / tiles_x<C>(G);
  const int i0 = tix * C::TX, j0 = tiy * C::TY;
  stage_tile<C, double, double>(G, i0, j0, x, nullptr, 0.0, p_s);
  __syncthreads();
  tile_elements<C, double>(G, i0, j0, p_s, ye_s);
  __syncthreads();
  if (threadIdx.x < C::TX * C::TY) {
    const int tx = threadIdx.x % C::TX, ty = threadIdx.x / C::TX;
    const int gi = i0 + tx, gj = j0 + ty;
    if (gi <= G.nx && gj <= G.ny) {
      double v[3];
      node_sum<C, double>(tx, ty, ye_s, v);
      const size_t n = (size_t)gj * (G.nx + 1) + gi;
      const unsigned fix = G.cmask[n];
      const int c = (ty + 1) * C::PW + tx + 1;
      double r[3];
      for (int d = 0; d < 3; d++) r[d] = b[3 * n + d] - ((fix >> d) & 1u ? p_s[d * C::NP + c] : v[d]);
      if (MODE == 2) {
        for (int d = 0; d < 3; d++) xo[3 * n + d] = r[d];
      } else {
        double z[3];
        sym3_apply(minv + 6 * n, r, z);
        for (int d = 0; d < 3; d++) {
          const double xc = p_s[d * C::NP + c];
          const double dn = (c1 != 0.0 ? c1 * (xc - xo[3 * n + d]) : 0.0) + c2 * z[d];
          xo[3 * n + d] = xc + dn;
        }
      }
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
    double D, Ds;
    elem_material(G, gi, gj, D, Ds);
    for (int c = 0; c < 3; c++) {
      double u[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, y[12];
      u[3 * corner[k] + c] = 1.0;
      q4_rect_apply<double>(G.ax[2 * gi], G.by[2 * gj], G.ax[2 * gi + 1], G.by[2 * gj + 1], D,
                            D * G.nu, D * (1.0 - G.nu) / 2.0, Ds, u, y);
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

// F_w(node) = the node's share of the 2x2 Gauss rule of every adjacent cell
// (grid_node_load_q4); rotation loads are zero and constrained DOFs get zero
// (solver.ts:400-407).
__global__ void k_pressure(KGrid G, double q0, const double* pt, int np, float* F) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nn = (int64_t)(G.nx + 1) * (G.ny + 1);
  if (n >= nn) return;
  const int i = (int)(n % (G.nx + 1)), j = (int)(n / (G.nx + 1));
  const double f = grid_node_load_q4(G.xc, G.yc, G.nx, G.ny, i, j, q0, pt, np);
  const unsigned fix = G.cmask[n];
  F[3 * n] = (fix & PP_FIX_W) ? 0.0f : (float)f;
  F[3 * n + 1] = 0.0f;
  F[3 * n + 2] = 0.0f;
}

// -------------------------------------------------- persistent FP64 PCG

// Iteration k (two all-reduces):
//   A: p = z + beta p_old staged with halo (p double-buffered so no
//      workgroup overwrites what a neighbour still reads), Ap = K p in LDS,
//      BC rows Ap = p, partial p.Ap                  -> all-reduce -> alpha
//   B: x += alpha p, r -= alpha Ap, z = M^-1 r, partials r.z, r.r
//                                                    -> all-reduce -> beta, |r|
// i.e. pcg.ts:146-206 with the p update moved to the start of the next
// iteration.  Stopping rule |r| < tol or |r|/|F| < tol (pcg.ts:174), pAp
// breakdown (pcg.ts:154), zero right-hand side (pcg.ts:118).  A restart
// (r0 != nullptr) starts from the x in memory with residual r0 and measures
// the relative test against the given |F|.  Only z and p are read by other
// workgroups; they go through st_coh / ld_coh.
struct PcgArgs {
  KGrid G;
  const double* F;
  const double* r0;
  double bnorm_in;
  double *x, *r, *z, *pb0, *pb1, *Ap;
  const double* minv;
  double* part;
  GridSync* gs;
  DevStats* st;
  double tol;
  int maxit;
};

// Tiles strided over a resident grid of workgroups, state in HBM; the
// two-reduction recurrence of pcg.ts:146-206.  Phase A stages p = z + b p
// (tile plus halo) and applies K; phase B updates x, r, z.  Phase A loads
// the next tile's staging values into registers while the current tile is
// applied (the current tile's element data is loaded before that prefetch,
// because loads complete in issue order; the workgroup barriers wait on LDS
// only, so the prefetch stays in flight across them): 5-6% per iteration at
// 700^2 and 1024^2.  Phase B and the init run over equal node ranges instead
// of tiles (Ap then goes through device-coherent stores and loads).
template <class C>
__global__ __launch_bounds__(C::NT) void k_pcg(PcgArgs A) {
  static_assert(C::NE <= C::NT, "one element per thread");
  static_assert(6 * C::NT <= 12 * C::NE, "phase B stages two chunk vectors in ye_s");
  constexpr int SPT = (C::NP + C::NT - 1) / C::NT;  // staged nodes per thread
  __shared__ double p_s[3 * C::NP];
  __shared__ double ye_s[12 * C::NE];
  __shared__ double red[80];
  const unsigned me = blockIdx.x, nwg = gridDim.x;
  const KGrid& G = A.G;
  const int ntx = tiles_x<C>(G), ntiles = tiles_total<C>(G);
  const int tx = threadIdx.x % C::TX, ty = threadIdx.x / C::TX;
  const bool node_thread = threadIdx.x < C::TX * C::TY;
  unsigned epoch = 0;
  const bool timing = me == 0 && threadIdx.x == 0;
  unsigned long long tk[4] = {0, 0, 0, 0}, t0 = 0, t1 = 0, t2 = 0, t3 = 0;
  auto allreduce = [&](auto& v) -> bool {
    block_sum_published<sizeof(v) / sizeof(double), C::NT>(v, red + 16);  // stores drain during the sum
    return grid_allreduce(v, A.part, A.gs, epoch, red, nwg, me);
  };

  // Node-wise work (init, phase B) goes by equal contiguous node ranges,
  // one per workgroup, so every workgroup reaches the all-reduce after the
  // same share and the accesses are whole rows; x and r are only ever
  // touched by their range's owner.
  const int64_t nn = (int64_t)(G.nx + 1) * (G.ny + 1);
  const int64_t lo = nn * me / nwg, hi = nn * (me + 1) / nwg;
  const TileBand band = xcd_tile_band(me, nwg, ntiles);

  // ---------------- init: r = F (or r0 on a restart), z = M^-1 r, r.z, r.r
  const bool restart = A.r0 != nullptr;
  const double* src = restart ? A.r0 : A.F;
  double acc[2] = {0, 0};
  for (int64_t n = lo + threadIdx.x; n < hi; n += C::NT) {
    {
      double rv[3] = {src[3 * n], src[3 * n + 1], src[3 * n + 2]}, zv[3], m[6];
      for (int k = 0; k < 6; k++) m[k] = A.minv[6 * n + k];
      sym3_apply(m, rv, zv);
      for (int d = 0; d < 3; d++) {
        st_coh(&A.z[3 * n + d], zv[d]);
        A.r[3 * n + d] = rv[d];
        if (!restart) A.x[3 * n + d] = 0.0;
        acc[0] += rv[d] * zv[d];
        acc[1] += rv[d] * rv[d];
      }
    }
  }
  if (!allreduce(acc)) return;
  double rz = acc[0], rr = acc[1];
  const double bnorm = restart ? A.bnorm_in : sqrt(rr);
  const bool done0 = restart ? (sqrt(rr) < A.tol || sqrt(rr) / bnorm < A.tol) : bnorm < 1e-30;
  if (done0) {
    if (me == 0 && threadIdx.x == 0) {
      A.st->iterations = 0; A.st->converged = 1; A.st->breakdown = 0;
      A.st->final_residual = restart ? sqrt(rr) : 0.0; A.st->bnorm = bnorm;
    }
    return;
  }

  // staging values (z, and p_old when there is one) of this thread's nodes
  // of a tile's staged block, loaded ahead of use
  double sv[SPT][6];
  auto stage_load = [&](int tile, const double* pold) {
    const int i0 = (tile % ntx) * C::TX, j0 = (tile / ntx) * C::TY;
#pragma unroll
    for (int k = 0; k < SPT; k++) {
      const int idx = threadIdx.x + k * C::NT;
#pragma unroll
      for (int q = 0; q < 6; q++) sv[k][q] = 0.0;
      if (idx < C::NP) {
        const int gi = i0 - 1 + idx % C::PW, gj = j0 - 1 + idx / C::PW;
        if (gi >= 0 && gi <= G.nx && gj >= 0 && gj <= G.ny) {
          const size_t n = (size_t)gj * (G.nx + 1) + gi;
#pragma unroll
          for (int d = 0; d < 3; d++) sv[k][d] = ld_coh(A.z + 3 * n + d);
          if (pold)
#pragma unroll
            for (int d = 0; d < 3; d++) sv[k][3 + d] = ld_coh(pold + 3 * n + d);
        }
      }
    }
  };
  auto stage_store = [&](bool with_pold, double beta) {
#pragma unroll
    for (int k = 0; k < SPT; k++) {
      const int idx = threadIdx.x + k * C::NT;
      if (idx < C::NP)
#pragma unroll
        for (int d = 0; d < 3; d++)
          p_s[d * C::NP + idx] = with_pold ? sv[k][d] + beta * sv[k][3 + d] : sv[k][d];
    }
  };

  double beta = 0;
  int it = 0, converged = 0, breakdown = 0;
  for (; it < A.maxit; it++) {
    const double* pold = it == 0 ? nullptr : (it & 1) ? A.pb1 : A.pb0;
    double* pnew = (it & 1) ? A.pb0 : A.pb1;
    // ---------------- phase A: p, Ap = K p, p.Ap
    if (timing) t0 = wall_clock64();
    double pap[1] = {0};
    if (band.first < band.last) stage_load(band.first, pold);
    for (int tile = band.first; tile < band.last; tile += band.step) {
      const int i0 = (tile % ntx) * C::TX, j0 = (tile / ntx) * C::TY;
      // this tile's element data and node mask, then its staged block, then
      // the next tile's staging loads
      const int le = threadIdx.x, ex = le % C::EW, ey = le / C::EW;
      const int ei = i0 - 1 + ex, ej = j0 - 1 + ey;
      const bool elem = le < C::NE && ei >= 0 && ei < G.nx && ej >= 0 && ej < G.ny;
      double ea = 0, eb = 0, eia = 0, eib = 0, eD = 0, eDs = 0;
      if (elem) {
        ea = G.ax[2 * ei]; eia = G.ax[2 * ei + 1];
        eb = G.by[2 * ej]; eib = G.by[2 * ej + 1];
        elem_material(G, ei, ej, eD, eDs);
      }
      const int gi = i0 + tx, gj = j0 + ty;
      const bool mine = node_thread && gi <= G.nx && gj <= G.ny;
      const int64_t n = mine ? (int64_t)gj * (G.nx + 1) + gi : 0;
      const unsigned fix = mine ? G.cmask[n] : 0u;
      stage_store(pold != nullptr, beta);
      if (tile + band.step < band.last) stage_load(tile + band.step, pold);
      __syncthreads();
      if (le < C::NE) {
        double y[12];
        if (elem) {
          const int c0 = ey * C::PW + ex, c1 = c0 + 1, c2 = c0 + C::PW + 1, c3 = c0 + C::PW;
          constexpr int NP = C::NP;
          const double u[12] = {p_s[c0], p_s[NP + c0], p_s[2 * NP + c0], p_s[c1], p_s[NP + c1],
                                p_s[2 * NP + c1], p_s[c2], p_s[NP + c2], p_s[2 * NP + c2],
                                p_s[c3], p_s[NP + c3], p_s[2 * NP + c3]};
          q4_rect_apply_fast<double>(ea, eb, eia, eib, eD, eD * G.nu, eD * (1.0 - G.nu) / 2.0,
                                     eDs, u, y);
        } else {
#pragma unroll
          for (int k = 0; k < 12; k++) y[k] = 0.0;
        }
#pragma unroll
        for (int k = 0; k < 12; k++) ye_s[k * C::NE + le] = y[k];
      }
      __syncthreads();
      double v[3] = {0, 0, 0};
      if (mine) node_sum<C, double>(tx, ty, ye_s, v);
      __syncthreads();  // every node sum has read ye_s
      if (mine) {
        const int c = (ty + 1) * C::PW + tx + 1;
        for (int d = 0; d < 3; d++) {
          const double pd = p_s[d * C::NP + c];
          const double ad = ((fix >> d) & 1u) ? pd : v[d];
          pap[0] += pd * ad;
          ye_s[3 * (ty * C::TX + tx) + d] = ad;
        }
      }
      __syncthreads();
      // p and K p of the tile's own nodes go out row by row with consecutive
      // lanes on consecutive doubles: a device-coherent store is written
      // through per instruction, so node-strided stores (lane = node,
      // one instruction per component) would write each line three times
      {
        const int rw = min(C::TX, G.nx + 1 - i0), rows = min(C::TY, G.ny + 1 - j0);
        for (int t = threadIdx.x; t < 3 * C::TX * rows; t += C::NT) {
          const int r = t / (3 * C::TX), k = t % (3 * C::TX);
          if (k < 3 * rw) {
            const int lx = k / 3, d = k % 3;
            const size_t g = 3 * ((size_t)(j0 + r) * (G.nx + 1) + i0) + k;
            st_coh(&pnew[g], p_s[d * C::NP + (r + 1) * C::PW + lx + 1]);
            st_coh(&A.Ap[g], ye_s[3 * (r * C::TX + lx) + d]);  // read by the node range's owner
          }
        }
      }
      __syncthreads();
    }
    if (timing) t1 = wall_clock64();
    if (!allreduce(pap)) return;
    if (timing) t2 = wall_clock64();
    if (fabs(pap[0]) < 1e-30) { breakdown = 1; it++; break; }
    const double alpha = rz / pap[0];

    // ---------------- phase B: x, r, z updates, r.z, r.r
    // Chunks of NT nodes: p and K p come in and z goes out through LDS with
    // consecutive lanes on consecutive doubles (whole lines per coherent
    // access); thread k of a chunk owns node nb + k as before.
    acc[0] = 0; acc[1] = 0;
    double* lb = ye_s;  // [3 NT] p then z, [3 NT] K p
    for (int64_t nb = lo; nb < hi; nb += C::NT) {
      const int cnt = (int)min((int64_t)C::NT, hi - nb);
      for (int t = threadIdx.x; t < 3 * cnt; t += C::NT) {
        lb[t] = ld_coh(&pnew[3 * nb + t]);
        lb[3 * C::NT + t] = ld_coh(&A.Ap[3 * nb + t]);
      }
      __syncthreads();
      if ((int)threadIdx.x < cnt) {
        const int64_t n = nb + threadIdx.x;
        double* own = lb + 3 * threadIdx.x;
        double rv[3], zv[3], m[6];
        for (int k = 0; k < 6; k++) m[k] = A.minv[6 * n + k];
        for (int d = 0; d < 3; d++) {
          A.x[3 * n + d] += alpha * own[d];
          rv[d] = A.r[3 * n + d] - alpha * own[3 * C::NT + d];
          A.r[3 * n + d] = rv[d];
        }
        sym3_apply(m, rv, zv);
        for (int d = 0; d < 3; d++) {
          own[d] = zv[d];  // this thread's own entries only
          acc[0] += rv[d] * zv[d];
          acc[1] += rv[d] * rv[d];
        }
      }
      __syncthreads();
      for (int t = threadIdx.x; t < 3 * cnt; t += C::NT) st_coh(&A.z[3 * nb + t], lb[t]);
      __syncthreads();  // before the next chunk overwrites lb
    }
    if (timing) t3 = wall_clock64();
    if (!allreduce(acc)) return;
    if (timing) {
      const unsigned long long t4 = wall_clock64();
      tk[0] += t1 - t0; tk[1] += t2 - t1; tk[2] += t3 - t2; tk[3] += t4 - t3;
    }
    rr = acc[1];
    const double rn = sqrt(rr);
    if (rn < A.tol || rn / bnorm < A.tol) { converged = 1; it++; break; }
    beta = acc[0] / rz;
    rz = acc[0];
  }
  if (me == 0 && threadIdx.x == 0) {
    A.st->iterations = it;
    A.st->converged = converged;
    A.st->breakdown = breakdown;
    A.st->final_residual = sqrt(rr);
    A.st->bnorm = bnorm;
    // phase A, its all-reduce, phase B, its all-reduce
    for (int k = 0; k < 4; k++) A.st->ticks[k] = tk[k];
  }
}

// Resident variant: one tile per workgroup and every per-node vector in
// registers of the node's thread.  Only tile-edge nodes are ever read by
// another workgroup (as its halo), so only they publish z and p (st_coh),
// and each phase A fetches just the one-node halo ring (ld_coh): the tile
// interior never goes through memory.  x is written once at the end.
template <class C>
__device__ __forceinline__ void ring_pos(int k, int& lx, int& ly) {
  if (k < C::PW) { lx = k; ly = 0; }
  else if (k < 2 * C::PW) { lx = k - C::PW; ly = C::PH - 1; }
  else if (k < 2 * C::PW + C::TY) { lx = 0; ly = 1 + (k - 2 * C::PW); }
  else { lx = C::PW - 1; ly = 1 + (k - 2 * C::PW - C::TY); }
}

template <class C>
__global__ __launch_bounds__(C::NT) void k_pcg_res(PcgArgs A) {
  constexpr int NRING = 2 * C::PW + 2 * C::TY;
  static_assert(NRING <= C::NT, "one ring node per thread");
  __shared__ double p_s[3 * C::NP];
  __shared__ double ye_s[12 * C::NE];
  __shared__ double red[80];
  const unsigned me = blockIdx.x, nwg = gridDim.x;
  const KGrid& G = A.G;
  const int ntx = tiles_x<C>(G);
  const int i0 = (me % ntx) * C::TX, j0 = (me / ntx) * C::TY;
  const int tx = threadIdx.x % C::TX, ty = threadIdx.x / C::TX;
  const int gi = i0 + tx, gj = j0 + ty;
  const bool mine = threadIdx.x < C::TX * C::TY && gi <= G.nx && gj <= G.ny;
  const int64_t n = mine ? (int64_t)gj * (G.nx + 1) + gi : 0;
  // tile-edge nodes are some neighbour's halo
  const bool edge = mine && (tx == 0 || ty == 0 || tx == C::TX - 1 || ty == C::TY - 1 ||
                             gi == G.nx || gj == G.ny);
  const int c = (ty + 1) * C::PW + tx + 1;  // own node in p_s
  // this thread's halo ring entry, if any
  int rn_ = -1;
  int64_t rnode = -1;
  if (threadIdx.x < NRING) {
    int lx, ly;
    ring_pos<C>(threadIdx.x, lx, ly);
    rn_ = ly * C::PW + lx;
    const int hi = i0 - 1 + lx, hj = j0 - 1 + ly;
    if (hi >= 0 && hi <= G.nx && hj >= 0 && hj <= G.ny) rnode = (int64_t)hj * (G.nx + 1) + hi;
  }
  unsigned epoch = 0;
  const bool timing = me == 0 && threadIdx.x == 0;
  unsigned long long tk[6] = {0, 0, 0, 0, 0, 0};
  auto allreduce = [&](auto& v) -> bool {
    block_sum_published<sizeof(v) / sizeof(double), C::NT>(v, red + 16);  // stores drain during the sum
    return grid_allreduce(v, A.part, A.gs, epoch, red, nwg, me);
  };

  // x, r, z in registers; p lives in the node's own p_s slot from one
  // iteration to the next and Ap in ap_s between the two phases
  __shared__ double ap_s[3 * C::NT];
  double x[3] = {0, 0, 0}, r[3] = {0, 0, 0}, z[3] = {0, 0, 0};
  const unsigned fix = mine ? G.cmask[n] : 0u;
  const bool restart = A.r0 != nullptr;
  double acc[2] = {0, 0};
  if (mine) {
    const double* src = restart ? A.r0 : A.F;
    for (int d = 0; d < 3; d++) {
      r[d] = src[3 * n + d];
      x[d] = restart ? A.x[3 * n + d] : 0.0;
    }
    sym3_apply(A.minv + 6 * n, r, z);
    for (int d = 0; d < 3; d++) {
      if (edge) st_coh(&A.z[3 * n + d], z[d]);
      acc[0] += r[d] * z[d];
      acc[1] += r[d] * r[d];
    }
  }
  if (!allreduce(acc)) return;
  double rz = acc[0], rr = acc[1];
  const double bnorm = restart ? A.bnorm_in : sqrt(rr);
  const bool done0 = restart ? (sqrt(rr) < A.tol || sqrt(rr) / bnorm < A.tol) : bnorm < 1e-30;
  if (done0) {
    if (mine)
      for (int d = 0; d < 3; d++) A.x[3 * n + d] = x[d];
    if (me == 0 && threadIdx.x == 0) {
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
    unsigned long long t0 = timing ? wall_clock64() : 0;
    // ---------------- phase A: p = z + beta p, halo ring, Ap = K p, p.Ap
    // every staged position is written once: own nodes, the ring (zero
    // outside the grid), and zero for tile slots past the grid's edge
    if (mine) {
      for (int d = 0; d < 3; d++) {
        const double pd = it == 0 ? z[d] : z[d] + beta * p_s[d * C::NP + c];
        p_s[d * C::NP + c] = pd;
        if (edge) st_coh(&pnew[3 * n + d], pd);
      }
    } else if (threadIdx.x < C::TX * C::TY) {
      for (int d = 0; d < 3; d++) p_s[d * C::NP + c] = 0.0;
    }
    if (rn_ >= 0) {
      for (int d = 0; d < 3; d++) {
        double v = 0.0;
        if (rnode >= 0) {
          v = ld_coh(&A.z[3 * rnode + d]);
          if (it > 0) v += beta * ld_coh(&pold[3 * rnode + d]);
        }
        p_s[d * C::NP + rn_] = v;
      }
    }
    __syncthreads();
    unsigned long long t1 = timing ? wall_clock64() : 0;
    tile_elements<C, double>(G, i0, j0, p_s, ye_s);
    __syncthreads();
    unsigned long long t2 = timing ? wall_clock64() : 0;
    double pap[1] = {0};
    if (mine) {
      double v[3];
      node_sum<C, double>(tx, ty, ye_s, v);
      for (int d = 0; d < 3; d++) {
        const double pd = p_s[d * C::NP + c];
        const double ad = ((fix >> d) & 1u) ? pd : v[d];
        ap_s[d * C::NT + threadIdx.x] = ad;
        pap[0] += pd * ad;
      }
    }
    unsigned long long t3 = timing ? wall_clock64() : 0;
    if (!allreduce(pap)) return;
    unsigned long long t4 = timing ? wall_clock64() : 0;
    if (fabs(pap[0]) < 1e-30) { breakdown = 1; it++; break; }
    const double alpha = rz / pap[0];
    // ---------------- phase B: x, r, z, r.z, r.r
    acc[0] = 0; acc[1] = 0;
    if (mine) {
      for (int d = 0; d < 3; d++) {
        x[d] += alpha * p_s[d * C::NP + c];
        r[d] -= alpha * ap_s[d * C::NT + threadIdx.x];
      }
      sym3_apply(A.minv + 6 * n, r, z);
      for (int d = 0; d < 3; d++) {
        if (edge) st_coh(&A.z[3 * n + d], z[d]);
        acc[0] += r[d] * z[d];
        acc[1] += r[d] * r[d];
      }
    }
    unsigned long long t5 = timing ? wall_clock64() : 0;
    if (!allreduce(acc)) return;
    if (timing) {
      const unsigned long long t6 = wall_clock64();
      tk[0] += t1 - t0; tk[1] += t2 - t1; tk[2] += t3 - t2;
      tk[3] += t4 - t3; tk[4] += t5 - t4; tk[5] += t6 - t5;
    }
    rr = acc[1];
    const double rnorm = sqrt(rr);
    if (rnorm < A.tol || rnorm / bnorm < A.tol) { converged = 1; it++; break; }
    beta = acc[0] / rz;
    rz = acc[0];
  }
  if (mine)
    for (int d = 0; d < 3; d++) A.x[3 * n + d] = x[d];
  if (me == 0 && threadIdx.x == 0) {
    A.st->iterations = it;
    A.st->converged = converged;
    A.st->breakdown = breakdown;
    A.st->final_residual = sqrt(rr);
    A.st->bnorm = bnorm;
    for (int k = 0; k < 6; k++) A.st->ticks[k] = tk[k];
  }
}

// ------------------------------ single-reduction (Chronopoulos-Gear) PCG

// Same iterates as pcg.ts:98-215 in exact arithmetic, one all-reduce per
// iteration (Chronopoulos & Gear 1989):
//   x += a p;  r -= a s;  u = M^-1 r;  w = K u
//   g' = r.u, d' = w.u, rho' = r.r                  (one all-reduce)
//   b = g'/g;  pAp = d' - b g'/a;  a' = g'/pAp;  p = u + b p;  s = w + b s
// (s = K p is carried instead of recomputed.)  K u needs u on the one-node
// halo ring: every workgroup keeps r and s of its ring nodes itself, updated
// with the owners' arithmetic (cg_step) from the w values the owners publish
// before the all-reduce, so the ring copies stay bit-identical and no second
// barrier is needed for the halo.  Stopping rule, breakdown and zero
// right-hand side as pcg.ts:118-181; |r_{k+1}| arrives with the all-reduce
// that follows the matvec of u_{k+1}.
__device__ __forceinline__ void cg_rs(double a, const double (&s)[3], double (&r)[3]) {
#pragma unroll
  for (int d = 0; d < 3; d++) r[d] = fma(-a, s[d], r[d]);
}
__device__ __forceinline__ void cg_sw(double b, const double (&w)[3], double (&s)[3]) {
#pragma unroll
  for (int d = 0; d < 3; d++) s[d] = fma(b, s[d], w[d]);
}
__device__ __forceinline__ void precond_fma(const double* m, const double (&r)[3], double (&u)[3]) {
  const double m00 = m[0], m01 = m[1], m02 = m[2], m11 = m[3], m12 = m[4], m22 = m[5];
  u[0] = fma(m00, r[0], fma(m01, r[1], m02 * r[2]));
  u[1] = fma(m01, r[0], fma(m11, r[1], m12 * r[2]));
  u[2] = fma(m02, r[0], fma(m12, r[1], m22 * r[2]));
}

template <class C>
__global__ __launch_bounds__(C::NT) void k_pcg_cg(PcgArgs A) {
  constexpr int NRING = 2 * C::PW + 2 * C::TY;
  static_assert(NRING <= C::NT, "one ring node per thread");
  __shared__ double u_s[3 * C::NP];
  __shared__ double ye_s[12 * C::NE];
  __shared__ double red[80];
  const unsigned me = blockIdx.x, nwg = gridDim.x;
  const KGrid& G = A.G;
  const int ntx = tiles_x<C>(G);
  const int i0 = (me % ntx) * C::TX, j0 = (me / ntx) * C::TY;
  const int tx = threadIdx.x % C::TX, ty = threadIdx.x / C::TX;
  const int gi = i0 + tx, gj = j0 + ty;
  const bool mine = threadIdx.x < C::TX * C::TY && gi <= G.nx && gj <= G.ny;
  const int64_t n = mine ? (int64_t)gj * (G.nx + 1) + gi : 0;
  const bool edge = mine && (tx == 0 || ty == 0 || tx == C::TX - 1 || ty == C::TY - 1 ||
                             gi == G.nx || gj == G.ny);
  const int c = (ty + 1) * C::PW + tx + 1;
  int rpos = -1;
  int64_t rnode = -1;
  if (threadIdx.x < NRING) {
    int lx, ly;
    ring_pos<C>(threadIdx.x, lx, ly);
    rpos = ly * C::PW + lx;
    const int hi = i0 - 1 + lx, hj = j0 - 1 + ly;
    if (hi >= 0 && hi <= G.nx && hj >= 0 && hj <= G.ny) rnode = (int64_t)hj * (G.nx + 1) + hi;
  }
  unsigned epoch = 0;
  const bool timing = me == 0 && threadIdx.x == 0;
  unsigned long long tk[6] = {0, 0, 0, 0, 0, 0};

  // own state
  double x[3] = {0, 0, 0}, r[3] = {0, 0, 0}, u[3] = {0, 0, 0}, w[3] = {0, 0, 0};
  double p[3] = {0, 0, 0}, s[3] = {0, 0, 0};
  // ring copies (r, s) and the ring node's M^-1
  double rr_[3] = {0, 0, 0}, rs_[3] = {0, 0, 0}, rm[6] = {0, 0, 0, 0, 0, 0};
  const unsigned fix = mine ? G.cmask[n] : 0u;
  const bool restart = A.r0 != nullptr;
  const double* src = restart ? A.r0 : A.F;
  const double* mown = A.minv + 6 * n;
  if (mine)
    for (int d = 0; d < 3; d++) {
      r[d] = src[3 * n + d];
      x[d] = restart ? A.x[3 * n + d] : 0.0;
    }
  if (rnode >= 0) {
    for (int d = 0; d < 3; d++) rr_[d] = src[3 * rnode + d];
    for (int k = 0; k < 6; k++) rm[k] = A.minv[6 * rnode + k];
  }

  // u = M^-1 r on own nodes and ring, staged; w = K u; edge w published;
  // partials r.u, w.u, r.r
  auto matvec = [&](double* wpub, double (&acc)[3]) {
    if (mine) {
      precond_fma(mown, r, u);
      for (int d = 0; d < 3; d++) u_s[d * C::NP + c] = u[d];
    } else if (threadIdx.x < C::TX * C::TY) {
      for (int d = 0; d < 3; d++) u_s[d * C::NP + c] = 0.0;
    }
    if (rpos >= 0) {
      double ru[3] = {0, 0, 0};
      if (rnode >= 0) precond_fma(rm, rr_, ru);
      for (int d = 0; d < 3; d++) u_s[d * C::NP + rpos] = ru[d];
    }
    __syncthreads();
    tile_elements<C, double>(G, i0, j0, u_s, ye_s);
    __syncthreads();
    acc[0] = acc[1] = acc[2] = 0;
    if (mine) {
      double v[3];
      node_sum<C, double>(tx, ty, ye_s, v);
      for (int d = 0; d < 3; d++) {
        w[d] = ((fix >> d) & 1u) ? u[d] : v[d];
        if (edge) st_coh(&wpub[3 * n + d], w[d]);
        acc[0] += r[d] * u[d];
        acc[1] += w[d] * u[d];
        acc[2] += r[d] * r[d];
      }
    }
  };
  auto allreduce = [&](double (&v)[3]) -> bool {
    block_sum_published<3, C::NT>(v, red + 16);  // stores drain during the sum
    // at most a few hundred workgroups: two loads in flight per lane suffice
    return grid_allreduce<3, 2>(v, A.part, A.gs, epoch, red, nwg, me);
  };
  // ring w published by the owners into wpub before the last all-reduce
  auto ring_w = [&](const double* wpub, double (&wv)[3]) {
    for (int d = 0; d < 3; d++) wv[d] = rnode >= 0 ? ld_coh(&wpub[3 * rnode + d]) : 0.0;
  };

  // ---------------- init: u0, w0 = K u0, g0 = r.u, d0 = w.u
  double acc[3];
  matvec(A.pb0, acc);
  if (!allreduce(acc)) return;
  double g = acc[0], pap = acc[1];
  const double bnorm = restart ? A.bnorm_in : sqrt(acc[2]);
  double rho = acc[2];
  int it = 0, converged = 0, breakdown = 0;
  const bool done0 = restart ? (sqrt(rho) < A.tol || sqrt(rho) / bnorm < A.tol) : bnorm < 1e-30;
  if (done0) {
    converged = 1;
  } else if (A.maxit > 0 && fabs(pap) < 1e-30) {
    breakdown = 1;
    it = 1;
  } else {
    double a = g / pap;
    for (int d = 0; d < 3; d++) { p[d] = u[d]; s[d] = w[d]; }
    if (rpos >= 0) ring_w(A.pb0, rs_);
    for (; it < A.maxit;) {
      unsigned long long t0 = timing ? wall_clock64() : 0;
      double* wpub = (it & 1) ? A.pb0 : A.pb1;
      // x += a p, r -= a s (own); r -= a s (ring copies)
      if (mine)
        for (int d = 0; d < 3; d++) x[d] = fma(a, p[d], x[d]);
      cg_rs(a, s, r);
      if (rnode >= 0) cg_rs(a, rs_, rr_);
      unsigned long long t1 = timing ? wall_clock64() : 0;
      matvec(wpub, acc);
      unsigned long long t2 = timing ? wall_clock64() : 0;
      if (!allreduce(acc)) return;
      unsigned long long t3 = timing ? wall_clock64() : 0;
      it++;
      const double g1 = acc[0], d1 = acc[1];
      rho = acc[2];
      const double rn = sqrt(rho);
      if (rn < A.tol || rn / bnorm < A.tol) { converged = 1; break; }
      const double b = g1 / g;
      pap = d1 - b * g1 / a;
      if (fabs(pap) < 1e-30) { breakdown = 1; it++; break; }
      a = g1 / pap;
      g = g1;
      for (int d = 0; d < 3; d++) p[d] = fma(b, p[d], u[d]);
      cg_sw(b, w, s);
      if (rpos >= 0) {
        double wv[3];
        ring_w(wpub, wv);
        if (rnode >= 0) cg_sw(b, wv, rs_);
      }
      if (timing) {
        const unsigned long long t4 = wall_clock64();
        tk[0] += t1 - t0; tk[1] += t2 - t1; tk[3] += t3 - t2; tk[4] += t4 - t3;
      }
    }
  }
  if (mine)
    for (int d = 0; d < 3; d++) A.x[3 * n + d] = x[d];
  if (me == 0 && threadIdx.x == 0) {
    A.st->iterations = it;
    A.st->converged = converged;
    A.st->breakdown = breakdown;
    A.st->final_residual = sqrt(rho);
    A.st->bnorm = bnorm;
    for (int k = 0; k < 6; k++) A.st->ticks[k] = tk[k];
  }
}

// ------------------------- Chebyshev-preconditioned single-reduction PCG
//
// The recurrence of k_pcg_cg with M = q_d(B K) B: u = M r is d steps of the
// Chebyshev semi-iteration for B K z = B r from z = 0 on [lmin, lmax]
// (theta = (lmax + lmin) / 2, delta = (lmax - lmin) / 2, sigma = theta /
// delta):
//   z1 = B r / theta;  rho_1 = 1 / sigma
//   z_{k+1} = z_k + dz_k,  dz_k = rho_k rho_{k-1} dz_{k-1} + (2 rho_k / delta) B (r - K z_k),
//   rho_k = 1 / (2 sigma - rho_{k-1}),
// i.e. q_d(l) = (1 - T_d((theta - l) / delta) / T_d(sigma)) / l > 0 on (0, lmax]: M is
// symmetric positive definite, and CG needs about d times fewer iterations
// (and all-reduces) for about the same number of K applies.  Each K apply
// needs the halo of its input: the owners publish their tile edges as
// tagged words (the exchange step's low 8 bits replace the mantissa's, as in
// grid_allreduce; relative change 2^-44), double-buffered by step parity, and
// each halo-ring thread polls its node's three words until the tags match
// (bounded spin; the abort word ends every waiter).  No flag and no store
// fence: the tag travels with the data.  A neighbour can be at most one step
// ahead: to write step s + 2 into the same buffer it must first have read
// this workgroup's step s + 1, written only after this workgroup read step s.
struct ChebArgs {
  double theta, delta;
  int degree;
};

template <class C>
__global__ __launch_bounds__(C::NT) void k_pcg_cheb(PcgArgs A, ChebArgs Q) {
  constexpr int NRING = 2 * C::PW + 2 * C::TY;
  static_assert(NRING <= C::NT, "one ring node per thread");
  __shared__ double u_s[3 * C::NP];
  __shared__ double ye_s[12 * C::NE];
  __shared__ double red[80];
  const unsigned me = blockIdx.x, nwg = gridDim.x;
  const KGrid& G = A.G;
  const int ntx = tiles_x<C>(G);
  const int i0 = (int)(me % ntx) * C::TX, j0 = (int)(me / ntx) * C::TY;
  const int tx = threadIdx.x % C::TX, ty = threadIdx.x / C::TX;
  const int gi = i0 + tx, gj = j0 + ty;
  const bool mine = threadIdx.x < C::TX * C::TY && gi <= G.nx && gj <= G.ny;
  const int64_t n = mine ? (int64_t)gj * (G.nx + 1) + gi : 0;
  const bool edge = mine && (tx == 0 || ty == 0 || tx == C::TX - 1 || ty == C::TY - 1 ||
                             gi == G.nx || gj == G.ny);
  const int c = (ty + 1) * C::PW + tx + 1;
  int rpos = -1;
  int64_t rnode = -1;
  if (threadIdx.x < NRING) {
    int lx, ly;
    ring_pos<C>(threadIdx.x, lx, ly);
    rpos = ly * C::PW + lx;
    const int hi = i0 - 1 + lx, hj = j0 - 1 + ly;
    if (hi >= 0 && hi <= G.nx && hj >= 0 && hj <= G.ny) rnode = (int64_t)hj * (G.nx + 1) + hi;
  }
  unsigned epoch = 0, step = 0;

  double x[3] = {0, 0, 0}, r[3] = {0, 0, 0}, u[3] = {0, 0, 0}, w[3] = {0, 0, 0};
  double p[3] = {0, 0, 0}, s[3] = {0, 0, 0};
  const unsigned fix = mine ? G.cmask[n] : 0u;
  const bool restart = A.r0 != nullptr;
  const double* src = restart ? A.r0 : A.F;
  const double* mown = A.minv + 6 * n;
  if (mine)
    for (int d = 0; d < 3; d++) {
      r[d] = src[3 * n + d];
      x[d] = restart ? A.x[3 * n + d] : 0.0;
    }

  // stage v on the tile and its halo ring in u_s (see above); false on abort
  auto exchange = [&](const double (&v)[3]) -> bool {
    step++;
    const unsigned tag = (step - 1u) % 255u + 1u;  // never 0: the buffers start zeroed
    unsigned long long* pub = (unsigned long long*)((step & 1u) ? A.pb1 : A.pb0);
    if (edge)
      for (int d = 0; d < 3; d++)
        __hip_atomic_store(&pub[3 * n + d], tag_word(v[d], tag), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    if (mine) {
      for (int d = 0; d < 3; d++) u_s[d * C::NP + c] = v[d];
    } else if (threadIdx.x < C::TX * C::TY) {
      for (int d = 0; d < 3; d++) u_s[d * C::NP + c] = 0.0;
    }
    int bad = 0;
    if (rpos >= 0) {
      double hv[3] = {0.0, 0.0, 0.0};
      if (rnode >= 0) {
        unsigned spins = 0;
        for (;;) {
          unsigned long long q[3];
          for (int d = 0; d < 3; d++)
            q[d] = __hip_atomic_load(&pub[3 * rnode + d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if ((unsigned)(q[0] & 0xffu) == tag && (unsigned)(q[1] & 0xffu) == tag &&
              (unsigned)(q[2] & 0xffu) == tag) {
            for (int d = 0; d < 3; d++) hv[d] = untag_word(q[d]);
            break;
          }
          if (++spins > kSpinLimit ||
              __hip_atomic_load(&A.gs->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
            __hip_atomic_store(&A.gs->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            bad = 1;
            break;
          }
        }
      }
      for (int d = 0; d < 3; d++) u_s[d * C::NP + rpos] = hv[d];
    }
    return !__syncthreads_or(bad);
  };
  // K v on the own nodes from the staged u_s, with the constraint rows
  auto apply = [&](const double (&v)[3], double (&kv)[3]) {
    tile_elements<C, double>(G, i0, j0, u_s, ye_s);
    __syncthreads();
    if (mine) {
      double t[3];
      node_sum<C, double>(tx, ty, ye_s, t);
      for (int d = 0; d < 3; d++) kv[d] = ((fix >> d) & 1u) ? v[d] : t[d];
    }
    __syncthreads();
  };
  // u = q_d(B K) B r
  auto precond = [&]() -> bool {
    const double sigma = Q.theta / Q.delta;
    double rho = 1.0 / sigma;
    double z[3] = {0, 0, 0}, dz[3] = {0, 0, 0};
    if (mine) {
      double res[3];
      precond_fma(mown, r, res);
      for (int d = 0; d < 3; d++) dz[d] = res[d] / Q.theta;
    }
    for (int k = 0; k < Q.degree; k++) {
      for (int d = 0; d < 3; d++) z[d] += dz[d];
      if (k == Q.degree - 1) break;
      if (!exchange(z)) return false;
      double kz[3] = {0, 0, 0};
      apply(z, kz);
      const double rn = 1.0 / (2.0 * sigma - rho);
      if (mine) {
        double t[3], res[3];
        for (int d = 0; d < 3; d++) t[d] = r[d] - kz[d];
        precond_fma(mown, t, res);
        for (int d = 0; d < 3; d++) dz[d] = rn * rho * dz[d] + 2.0 * rn / Q.delta * res[d];
      }
      rho = rn;
    }
    for (int d = 0; d < 3; d++) u[d] = z[d];
    return true;
  };
  // u = M r, w = K u, partials r.u, w.u, r.r
  auto matvec = [&](double (&acc)[3]) -> bool {
    if (!precond() || !exchange(u)) return false;
    apply(u, w);
    acc[0] = acc[1] = acc[2] = 0;
    if (mine)
      for (int d = 0; d < 3; d++) {
        acc[0] += r[d] * u[d];
        acc[1] += w[d] * u[d];
        acc[2] += r[d] * r[d];
      }
    return true;
  };
  auto allreduce = [&](double (&v)[3]) -> bool {
    block_sum_published<3, C::NT>(v, red + 16);  // stores drain during the sum
    return grid_allreduce<3, 2>(v, A.part, A.gs, epoch, red, nwg, me);
  };

  double acc[3];
  if (!matvec(acc) || !allreduce(acc)) return;
  double g = acc[0], pap = acc[1];
  const double bnorm = restart ? A.bnorm_in : sqrt(acc[2]);
  double rho = acc[2];
  int it = 0, converged = 0, breakdown = 0;
  const bool done0 = restart ? (sqrt(rho) < A.tol || sqrt(rho) / bnorm < A.tol) : bnorm < 1e-30;
  if (done0) {
    converged = 1;
  } else if (A.maxit > 0 && fabs(pap) < 1e-30) {
    breakdown = 1;
    it = 1;
  } else {
    double a = g / pap;
    for (int d = 0; d < 3; d++) { p[d] = u[d]; s[d] = w[d]; }
    for (; it < A.maxit;) {
      if (mine)
        for (int d = 0; d < 3; d++) x[d] = fma(a, p[d], x[d]);
      cg_rs(a, s, r);
      if (!matvec(acc) || !allreduce(acc)) return;
      it++;
      const double g1 = acc[0], d1 = acc[1];
      rho = acc[2];
      const double rn = sqrt(rho);
      if (rn < A.tol || rn / bnorm < A.tol) { converged = 1; break; }
      const double b = g1 / g;
      pap = d1 - b * g1 / a;
      if (fabs(pap) < 1e-30) { breakdown = 1; it++; break; }
      a = g1 / pap;
      g = g1;
      for (int d = 0; d < 3; d++) p[d] = fma(b, p[d], u[d]);
      cg_sw(b, w, s);
    }
  }
  if (mine)
    for (int d = 0; d < 3; d++) A.x[3 * n + d] = x[d];
  if (me == 0 && threadIdx.x == 0) {
    A.st->iterations = it;
    A.st->converged = converged;
    A.st->breakdown = breakdown;
    A.st->final_residual = sqrt(rho);
    A.st->bnorm = bnorm;
    for (int k = 0; k < 6; k++) A.st->ticks[k] = 0;
  }
}

__global__ void k_scale_dof(int64_t n, double* v, double f) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] *= f;
}

__global__ void k_zero_sync(GridSync* gs) {
  if (threadIdx.x < 64) {
    gs->count = 0;
    gs->abort = 0;
  }
}

// Clears what a persistent launch needs at zero (the statistics block and
// the tagged all-reduce slots) with stores, not hipMemsetAsync, so that the
// launch sequence holds no memset node if a caller captures it.
__global__ void k_zero_launch(DevStats* st, double* part, int64_t n_part) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_part) part[i] = 0.0;
  if (i == 0) *st = DevStats{};
}

// Launch shape of the persistent PCG for this plan.
// Tile of the resident single-reduction kernel and of the strided kernel.
// Measured on MI355X (us per iteration): 143^2 nodes 5.46 on 18x18 tiles vs
// 5.63 on 15x15; 700^2 elements strided 55.8 / 57.0 / 63.1 on 18 / 23 / 15
// (before the staging prefetch, which leaves a 23x23 tile short of
// registers for one 576-thread workgroup per CU).
using CfgR = Tile<18, 18, 384>;

struct PcgShape {
  // 0 = 31x31 resident, 1 = 15x15 resident, 2 = 18x18 strided,
  // 3 = 18x18 single-reduction resident, 4 = 18x18 Chebyshev resident
  int kind;
  int grid;
};

// Large tiles (fewer workgroups, cheaper barriers) while they are few;
// smaller resident tiles next; tiles strided over a resident grid last.
constexpr int kMaxLargeTiles = 64;

int pcg_shape(pp_plan* pl, int variant, PcgShape* sh) {
  const int nx = pl->nx, ny = pl->ny, cu = pl->ctx->cu_count;
  auto tiles = [&](int T) { return ((nx + T) / T) * ((ny + T) / T); };
  const int nl = tiles(CfgX::TX), ns = tiles(CfgS::TX);
  int per_cu = 0;
  const bool small_tiles = variant == PP_FLAG_SMALL_TILES;
  if (variant == 3) {  // Chebyshev preconditioner: resident tiles only
    PP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg_cheb<CfgR>, CfgR::NT, 0));
    if (tiles(CfgR::TX) > per_cu * cu)
      return fail(PP_ERR_UNSUPPORTED,
                  "the Chebyshev preconditioner needs every 18x18 tile resident at once");
    sh->kind = 4;
    sh->grid = tiles(CfgR::TX);
    return PP_OK;
  }
  if (variant == 0) {  // default: single-reduction kernel while it is resident
    PP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg_cg<CfgR>, CfgR::NT, 0));
    if (tiles(CfgR::TX) <= per_cu * cu) {
      sh->kind = 3;
      sh->grid = tiles(CfgR::TX);
      return PP_OK;
    }
  }
  if (!small_tiles && nl <= kMaxLargeTiles) {
    PP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg_res<CfgX>, CfgX::NT, 0));
    if (nl <= per_cu * cu) {
      sh->kind = 0;
      sh->grid = nl;
      return PP_OK;
    }
  }
  PP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg_res<CfgS>, CfgS::NT, 0));
  if (ns <= per_cu * cu) {
    sh->kind = 1;
    sh->grid = ns;
    return PP_OK;
  }
  // Strided: 18x18-node tiles, one workgroup per CU (more add all-reduce
  // slots and tile-count imbalance without hiding more latency: 63, 64, 70
  // and 78 us per iteration at 1-4 per CU on 15x15 tiles at 700^2).
  PP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg<CfgR>, CfgR::NT, 0));
  sh->kind = 2;
  sh->grid = std::min(tiles(CfgR::TX), std::min(per_cu, 1) * cu);
  if (sh->grid <= 0) return fail(PP_ERR_HIP, "persistent PCG kernel cannot be made resident");
  return PP_OK;
}

// One launch of the persistent PCG (cold start when r0 is null).  A plain
// launch sized to what is co-resident: the workgroups then all run at once
// on an otherwise idle device; if they cannot (another process holds CUs)
// the barrier spin budget runs out and the call fails with PP_ERR_TIMEOUT
// instead of hanging.  (A cooperative launch would refuse to run at all
// while another process has the device open.)
int launch_pcg(pp_plan* pl, const StructWs& ws, const double* r0, double bnorm, double tol,
               int maxit, int variant, bool resident_only, DevStats* ds) {
  hipStream_t s = pl->ctx->stream;
  PcgShape sh;
  int rc = pcg_shape(pl, variant, &sh);
  if (rc) return rc;
  if (resident_only && sh.kind != 3)
    return fail(PP_ERR_UNSUPPORTED,
                "this entry point runs the single-reduction PCG with every 18x18 tile resident");
  const size_t npart = (size_t)8 * sh.grid;
  if (ws.part_n < npart) return fail(PP_ERR_INVALID, "grid-reduction partials buffer too small");
  PcgArgs A;
  A.G = make_kgrid(pl);
  A.F = ws.F;
  A.r0 = r0;
  A.bnorm_in = bnorm;
  A.x = ws.x; A.r = pl->r.p; A.z = pl->z.p; A.pb0 = ws.pb0; A.pb1 = ws.pb1;
  A.Ap = pl->Ap.p;
  A.minv = pl->minv.p;
  A.part = ws.part;
  A.gs = ws.gs;
  A.st = ws.ds;
  A.tol = tol;
  A.maxit = maxit;
  hipLaunchKernelGGL(k_zero_sync, dim3(1), dim3(64), 0, s, ws.gs);
  hipLaunchKernelGGL(k_zero_launch, dim3(grid_blocks((int64_t)npart, 256)), dim3(256), 0, s, ws.ds,
                     ws.part, (int64_t)npart);
  if (sh.kind == 0)
    hipLaunchKernelGGL((k_pcg_res<CfgX>), dim3(sh.grid), dim3(CfgX::NT), 0, s, A);
  else if (sh.kind == 1)
    hipLaunchKernelGGL((k_pcg_res<CfgS>), dim3(sh.grid), dim3(CfgS::NT), 0, s, A);
  else if (sh.kind == 2)
    hipLaunchKernelGGL((k_pcg<CfgR>), dim3(sh.grid), dim3(CfgR::NT), 0, s, A);
  else if (sh.kind == 3)
    hipLaunchKernelGGL((k_pcg_cg<CfgR>), dim3(sh.grid), dim3(CfgR::NT), 0, s, A);
  else {
    // tag 0 never matches: stale words of an earlier launch cannot pass
    PP_HIP(hipMemsetAsync(ws.pb0, 0, pl->n_dof * sizeof(double), s));
    PP_HIP(hipMemsetAsync(ws.pb1, 0, pl->n_dof * sizeof(double), s));
    // interval [lmax / kChebRatio, lmax], lmax with a 5% margin over the
    // power-iteration estimate (q_d stays positive up to lmax)
    constexpr double kChebRatio = 100.0;
    const double hi = 1.05 * pl->cheb_lmax, lo = hi / kChebRatio;
    ChebArgs Q;
    Q.theta = 0.5 * (hi + lo);
    Q.delta = 0.5 * (hi - lo);
    Q.degree = pl->cheb_degree;
    hipLaunchKernelGGL((k_pcg_cheb<CfgR>), dim3(sh.grid), dim3(CfgR::NT), 0, s, A, Q);
  }
  PP_HIP(hipGetLastError());
  GridSync gsh;
  PP_HIP(hipMemcpyAsync(ds, ws.ds, sizeof(*ds), hipMemcpyDeviceToHost, s));
  PP_HIP(hipMemcpyAsync(&gsh, ws.gs, sizeof(gsh), hipMemcpyDeviceToHost, s));
  PP_HIP(hipStreamSynchronize(s));
  if (gsh.abort) return fail(PP_ERR_TIMEOUT, "grid barrier spin limit reached");
  return PP_OK;
}

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
  const int ntiles =
      ((pl->nx + 1 + CfgS::TX - 1) / CfgS::TX) * ((pl->ny + 1 + CfgS::TY - 1) / CfgS::TY);
  hipLaunchKernelGGL(k_apply, dim3(ntiles), dim3(CfgS::NT), 0, pl->ctx->stream, G, dx, dy);
  PP_HIP(hipGetLastError());
  return PP_OK;
}

int struct_sweep(pp_plan* pl, int mode, const double* x, const double* b, const float* minv,
                 double c1, double c2, double* xo) {
  KGrid G = make_kgrid(pl);
  const int ntiles =
      ((pl->nx + 1 + CfgS::TX - 1) / CfgS::TX) * ((pl->ny + 1 + CfgS::TY - 1) / CfgS::TY);
  if (mode == 1)
    hipLaunchKernelGGL(k_apply_sweep<1>, dim3(ntiles), dim3(CfgS::NT), 0, pl->ctx->stream, G, x, b,
                       minv, c1, c2, xo);
  else
    hipLaunchKernelGGL(k_apply_sweep<2>, dim3(ntiles), dim3(CfgS::NT), 0, pl->ctx->stream, G, x, b,
                       minv, c1, c2, xo);
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

// FP64 PCG with residual replacement: after the recurrence converges the
// true residual F - K x is recomputed in FP64; while it is above the
// tolerance (the recurrence drifts from it by rounding) PCG restarts from x
// with that residual, at most kMaxRestarts times.  All of it is timed.
// lambda_max(B K) by 60 power iterations from a fixed pseudo-random start on
// the free DOFs (pl->tmp, z and Ap as scratch); kept in the plan.
int cheb_lambda_max(pp_plan* pl) {
  const int64_t nd = pl->n_dof;
  std::vector<double> h(nd);
  uint64_t sd = 0x9E3779B97F4A7C15ull;
  for (int64_t i = 0; i < nd; i++) {
    sd = sd * 6364136223846793005ull + 1442695040888963407ull;
    h[i] = pl->h_cmask[i / 3] >> (i % 3) & 1 ? 0.0 : (double)(sd >> 11) * (1.0 / 9007199254740992.0) - 0.5;
  }
  hipStream_t s = pl->ctx->stream;
  PP_HIP(hipMemcpyAsync(pl->z.p, h.data(), nd * 8, hipMemcpyHostToDevice, s));
  double lam = 0.0, nv = 0.0, unused = 0.0;
  int rc;
  if ((rc = launch_norms(pl, pl->z.p, pl->z.p, &unused, &nv))) return rc;
  for (int it = 0; it < 60 && nv > 0; it++) {
    if ((rc = struct_apply_k(pl, pl->z.p, pl->Ap.p)) || (rc = struct_apply_precond(pl, pl->Ap.p, pl->tmp.p)))
      return rc;
    double nb = 0.0;
    if ((rc = launch_norms(pl, pl->tmp.p, pl->tmp.p, &unused, &nb))) return rc;
    if (!(nb > 0)) break;
    lam = nb / nv;
    hipLaunchKernelGGL(k_scale_dof, dim3(grid_blocks(nd, 256)), dim3(256), 0, s, nd, pl->tmp.p, 1.0 / nb);
    PP_HIP(hipGetLastError());
    std::swap(pl->z.p, pl->tmp.p);
    nv = 1.0;
  }
  if (!(lam > 0)) return fail(PP_ERR_INVALID, "Chebyshev preconditioner: no lambda_max estimate");
  pl->cheb_lmax = lam;
  return PP_OK;
}

// 0: single-reduction kernel; 1: two-reduction kernel; 2: two-reduction on
// small tiles; 3: single-reduction with the Chebyshev preconditioner
static int solve_variant(const pp_solve_opts* o) {
  if (o->precond == PP_PRECOND_CHEBYSHEV) return 3;
  return (o->flags & PP_FLAG_SMALL_TILES) ? 2 : (o->flags & PP_FLAG_CLASSIC_PCG) ? 1 : 0;
}

int struct_partials_needed(pp_plan* pl, int variant, size_t* n) {
  PcgShape sh;
  int rc = pcg_shape(pl, variant, &sh);
  if (rc) return rc;
  *n = (size_t)8 * sh.grid;
  return PP_OK;
}

int struct_solve_fp64(pp_plan* pl, const pp_solve_opts* o, pp_solve_stats* st) {
  int rc;
  size_t npart = 0;
  if ((rc = struct_partials_needed(pl, solve_variant(o), &npart))) return rc;
  if (pl->partials.n < npart && (rc = pl->partials.alloc(npart))) return rc;
  if (pl->normp.n < kNormParts && (rc = pl->normp.alloc(kNormParts))) return rc;
  StructWs ws{pl->F.p, pl->x.p, pl->p0.p, pl->p1.p, pl->tmp.p, pl->partials.p, pl->partials.n,
              pl->normp.p, pl->gsync.p, pl->dstats.p};
  return struct_solve_fp64_ws(pl, o, st, ws, false);
}

int struct_solve_fp64_ws(pp_plan* pl, const pp_solve_opts* o, pp_solve_stats* st,
                         const StructWs& ws, bool resident_only) {
  constexpr int kMaxRestarts = 4;
  hipStream_t s = pl->ctx->stream;
  int rc;
  const int dw = solve_variant(o);
  if (o->precond == PP_PRECOND_CHEBYSHEV) {
    pl->cheb_degree = o->poly_degree > 0 ? o->poly_degree : 8;
    if (pl->cheb_degree > 32) return fail(PP_ERR_INVALID, "poly_degree must be at most 32");
    if (pl->cheb_lmax <= 0 && (rc = cheb_lambda_max(pl))) return rc;
  }
  PP_HIP(hipEventRecord(pl->ctx->ev0, s));
  DevStats ds;
  rc = launch_pcg(pl, ws, nullptr, 0.0, o->tolerance, o->max_iterations, dw, resident_only, &ds);
  if (rc) return rc;
  const double bnorm = ds.bnorm;
  int total = ds.iterations;
  for (int k = 0; k < 6; k++) st->phase_us[k] = 0.01 * (double)ds.ticks[k];
  st->sweeps = 1;
  st->sweep_iterations[0] = ds.iterations;
  double rel = 0;
  if (ds.converged && bnorm > 0 && !(o->flags & PP_FLAG_NO_TRUE_RESIDUAL)) {
    for (int k = 0;; k++) {
      if ((rc = struct_apply_k(pl, ws.x, ws.tmp))) return rc;
      double res = 0, fn = 0;
      if ((rc = launch_norms_into(pl, ws.normp, ws.F, ws.tmp, &res, &fn))) return rc;
      rel = fn > 0 ? res / fn : 0.0;
      st->sweep_residual[std::min(st->sweeps, PP_MAX_SWEEPS) - 1] = rel;
      if (rel <= o->tolerance || k == kMaxRestarts || total >= o->max_iterations) break;
      if ((rc = launch_residual(pl, ws.F, ws.tmp, ws.tmp))) return rc;
      rc = launch_pcg(pl, ws, ws.tmp, bnorm, o->tolerance, o->max_iterations - total, dw,
                      resident_only, &ds);
      if (rc) return rc;
      total += ds.iterations;
      for (int k = 0; k < 6; k++) st->phase_us[k] += 0.01 * (double)ds.ticks[k];
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
