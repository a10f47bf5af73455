// Batched plates: hundreds of rectangular-grid plates, Q4 Mindlin or DKT
// (each grid cell split into triangles (n0,n1,n2), (n0,n2,n3)), solved by
// ONE persistent launch.
//
// Every plate is a tensor grid, so its operator is a sum of 12x12 cell
// matrices: the Q4 element itself, or for DKT the two triangles' 9x9
// matrices assembled on the cell's four corners.  Q4 cells apply the
// closed-form rectangle operator of the structured path with the exact
// column width and row height.  DKT cells read their exact cell matrix from
// a per-plate table indexed by (row-height class, column-width class): the
// float32 node coordinates give a plate only a few distinct widths, so the
// table is small, built once on the host in FP64 and L1/L2-resident.  The
// apply is the tile gather of the structured path (stage the tile plus halo,
// apply every touching cell, each node sums its four cells in fixed order):
// race-free with no colour passes, so DKT needs no colouring.
//
// Workgroups are dealt to plates; the workgroups of one plate run its PCG
// (the two-reduction recurrence of pcg.ts:146-206, OR stopping rule, FP64)
// with a per-plate atomic-free all-reduce, and leave when their plate has
// converged.  A plate dealt one workgroup per tile may instead keep every
// per-node vector in registers and run the single-reduction recurrences
// (one pass and one all-reduce per iteration, only the tile edge through
// memory): latency-bound small plates with many iterations gain about 4x
// per iteration, but the kernel then holds one workgroup per CU less.  The
// schedule (tile shape per plate, workgroups per plate, workgroups per CU,
// resident plates or not) is chosen by pp_batch_autotune: it measures each
// plate's iteration count once, then times candidate schedules (workgroups
// dealt greedily to minimise the slowest plate's iterations x tiles per
// workgroup, at several occupancies and tile shapes) and keeps the fastest.

#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <vector>

#include "pp_grid.h"
#include "pp_load.h"

namespace pp {

namespace {

constexpr int kNT = 256;         // threads per workgroup; every tile shape has 256 cells
constexpr int kMaxNP = 33 * 9;   // staged nodes of the largest shape (31 x 7 nodes)
constexpr int kShapes = 3;
const int kShapeTX[kShapes] = {15, 31, 7};
const int kShapeTY[kShapes] = {15, 7, 31};

struct PlateDesc {
  int nx, ny, TX, TY, ntx, ntiles, wg0, nwg, ncc, nclass, dkt;
  double D11, D12, D33, Ds;  // Q4 material (bending D, nu D, (1 - nu) D / 2, shear)
  const double* ax;     // Q4: per column a, 1/a
  const double* by;     // Q4: per row b, 1/b
  const double* xc;     // column x coordinates (nx + 1), for the loads
  const double* yc;     // row y coordinates (ny + 1)
  const uint8_t* cmask;
  const uint16_t* ccls;  // column class per column
  const uint16_t* rcls;  // row class per row
  const double* kcell;   // [nclass][144], class = rcls * ncc + ccls
  const double* minv;   // 6 per node
  const double* F;
  double *x, *r, *z, *p0, *p1, *Ap;
  double* part;  // 8 * nwg
  GridSync* gs;
  DevStats* st;
};

// y (12) = K_cell u (12) for cell (gi, gj) of the plate
__device__ __forceinline__ void cell_apply(const PlateDesc& D, int gi, int gj,
                                           const double (&u)[12], double (&y)[12]) {
  if (!D.dkt) {
    q4_rect_apply_fast<double>(D.ax[2 * gi], D.by[2 * gj], D.ax[2 * gi + 1], D.by[2 * gj + 1],
                               D.D11, D.D12, D.D33, D.Ds, u, y);
    return;
  }
  const double* K = D.kcell + (size_t)(D.rcls[gj] * D.ncc + D.ccls[gi]) * 144;
  for (int a = 0; a < 12; a++) {
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < 12; b++) s += K[a * 12 + b] * u[b];
    y[a] = s;
  }
}

// Stage p = src + beta pold for the tile at (i0, j0) plus halo (runtime shape).
template <bool COH>
__device__ __forceinline__ void stage(const PlateDesc& D, int i0, int j0, const double* src,
                                      const double* pold, double beta, double* p_s) {
  const int PW = D.TX + 2, NP = PW * (D.TY + 2);
  for (int idx = threadIdx.x; idx < NP; idx += kNT) {
    const int lx = idx % PW, ly = idx / PW;
    const int gi = i0 - 1 + lx, gj = j0 - 1 + ly;
    double v0 = 0, v1 = 0, v2 = 0;
    if (gi >= 0 && gi <= D.nx && gj >= 0 && gj <= D.ny) {
      const size_t n = (size_t)gj * (D.nx + 1) + gi;
      v0 = ld_in<COH>(src + 3 * n);
      v1 = ld_in<COH>(src + 3 * n + 1);
      v2 = ld_in<COH>(src + 3 * n + 2);
      if (pold) {
        v0 += beta * ld_in<COH>(pold + 3 * n);
        v1 += beta * ld_in<COH>(pold + 3 * n + 1);
        v2 += beta * ld_in<COH>(pold + 3 * n + 2);
      }
    }
    p_s[idx] = v0;
    p_s[kMaxNP + idx] = v1;
    p_s[2 * kMaxNP + idx] = v2;
  }
}

// Apply every cell touching the tile (one cell per thread) into ye_s.
__device__ __forceinline__ void cells(const PlateDesc& D, int i0, int j0, const double* p_s,
                                      double* ye_s) {
  const int EW = D.TX + 1, PW = D.TX + 2;
  const int le = threadIdx.x;  // (TX + 1) (TY + 1) == 256 for every shape
  const int ex = le % EW, ey = le / EW;
  const int gi = i0 - 1 + ex, gj = j0 - 1 + ey;
  double y[12];
  if (gi >= 0 && gi < D.nx && gj >= 0 && gj < D.ny) {
    const int c0 = ey * PW + ex, c1 = c0 + 1, c2 = c0 + PW + 1, c3 = c0 + PW;
    const double u[12] = {p_s[c0], p_s[kMaxNP + c0], p_s[2 * kMaxNP + c0],
                          p_s[c1], p_s[kMaxNP + c1], p_s[2 * kMaxNP + c1],
                          p_s[c2], p_s[kMaxNP + c2], p_s[2 * kMaxNP + c2],
                          p_s[c3], p_s[kMaxNP + c3], p_s[2 * kMaxNP + c3]};
    cell_apply(D, gi, gj, u, y);
  } else {
#pragma unroll
    for (int k = 0; k < 12; k++) y[k] = 0.0;
  }
#pragma unroll
  for (int k = 0; k < 12; k++) ye_s[k * kNT + le] = y[k];
}

__device__ __forceinline__ void gather(const PlateDesc& D, int tx, int ty, const double* ye_s,
                                       double (&v)[3]) {
  const int EW = D.TX + 1;
  const int sw = ty * EW + tx, se = sw + 1, ne = sw + EW + 1, nw = sw + EW;
#pragma unroll
  for (int d = 0; d < 3; d++)
    v[d] = ye_s[(6 + d) * kNT + sw] + ye_s[(9 + d) * kNT + se] + ye_s[d * kNT + ne] +
           ye_s[(3 + d) * kNT + nw];
}

// Ring slot k of a TX x TY tile's one-node halo (PW = TX + 2, PH = TY + 2):
// bottom row, top row, left column, right column.
__device__ __forceinline__ void ring_slot(int k, int PW, int PH, int TY, int& lx, int& ly) {
  if (k < PW) { lx = k; ly = 0; }
  else if (k < 2 * PW) { lx = k - PW; ly = PH - 1; }
  else if (k < 2 * PW + TY) { lx = 0; ly = 1 + (k - 2 * PW); }
  else { lx = PW - 1; ly = 1 + (k - 2 * PW - TY); }
}

// Resident single-reduction PCG (Chronopoulos-Gear) for a plate that has
// one workgroup per tile: the recurrences of the structured path's
//   x += a p; r -= a s; u = M^-1 r; w = K u; (g, d, rho) = (r.u, w.u, r.r)
//   b = g'/g; pAp = d' - b g'/a; a' = g'/pAp; p = u + b p; s = w + b s
// with every per-node vector in registers.  K u needs u on the halo ring:
// the workgroup keeps r and s of its ring nodes itself, updated with the
// owners' arithmetic from the w values the owners publish (tile edge, in
// p0/p1 by iteration parity) before the all-reduce, so the copies stay
// bit-identical and one all-reduce per iteration is the only barrier.
// Same iterates as pcg.ts:146-206 in exact arithmetic; OR stopping rule,
// breakdown and zero right-hand side as the strided path.
__device__ __forceinline__ void batch_cg(const PlateDesc& D, unsigned me, unsigned nwg,
                                         double tol, int maxit, const double* src, bool restart,
                                         double bnorm_in, double* u_s, double* ye_s,
                                         double* red) {
  const int TX = D.TX, TY = D.TY, PW = TX + 2, PH = TY + 2;
  const int NRING = 2 * PW + 2 * TY;  // <= 80 for every shape
  const int i0 = (me % D.ntx) * TX, j0 = (me / D.ntx) * TY;
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int gi = i0 + tx, gj = j0 + ty;
  const bool node_thread = threadIdx.x < TX * TY;
  const bool mine = node_thread && gi <= D.nx && gj <= D.ny;
  const int64_t n = mine ? (int64_t)gj * (D.nx + 1) + gi : 0;
  const bool edge = mine && (tx == 0 || ty == 0 || tx == TX - 1 || ty == TY - 1 ||
                             gi == D.nx || gj == D.ny);
  const int c = (ty + 1) * PW + tx + 1;
  int rpos = -1;
  int64_t rnode = -1;
  if ((int)threadIdx.x < NRING) {
    int lx, ly;
    ring_slot(threadIdx.x, PW, PH, TY, lx, ly);
    rpos = ly * PW + lx;
    const int hi = i0 - 1 + lx, hj = j0 - 1 + ly;
    if (hi >= 0 && hi <= D.nx && hj >= 0 && hj <= D.ny) rnode = (int64_t)hj * (D.nx + 1) + hi;
  }
  unsigned epoch = 0;
  double x[3] = {0, 0, 0}, r[3] = {0, 0, 0}, u[3] = {0, 0, 0}, w[3] = {0, 0, 0};
  double p[3] = {0, 0, 0}, s[3] = {0, 0, 0};
  double rr_[3] = {0, 0, 0}, rs_[3] = {0, 0, 0}, rm[6] = {0, 0, 0, 0, 0, 0};
  const unsigned fix = mine ? D.cmask[n] : 0u;
  const double* mown = D.minv + 6 * n;
  if (mine)
    for (int d = 0; d < 3; d++) {
      r[d] = src[3 * n + d];
      x[d] = restart ? D.x[3 * n + d] : 0.0;
    }
  if (rnode >= 0) {
    for (int d = 0; d < 3; d++) rr_[d] = src[3 * rnode + d];
    for (int k = 0; k < 6; k++) rm[k] = D.minv[6 * rnode + k];
  }
  // u = M^-1 r (own nodes and ring), w = K u, edge w published, partials
  auto matvec = [&](double* wpub, double (&acc)[3]) {
    if (mine) {
      sym3_apply(mown, r, u);
      for (int d = 0; d < 3; d++) u_s[d * kMaxNP + c] = u[d];
    } else if (node_thread) {
      for (int d = 0; d < 3; d++) u_s[d * kMaxNP + c] = 0.0;
    }
    if (rpos >= 0) {
      double ru[3] = {0, 0, 0};
      if (rnode >= 0) sym3_apply(rm, rr_, ru);
      for (int d = 0; d < 3; d++) u_s[d * kMaxNP + rpos] = ru[d];
    }
    __syncthreads();
    cells(D, i0, j0, u_s, ye_s);
    __syncthreads();
    acc[0] = acc[1] = acc[2] = 0;
    if (mine) {
      double v[3];
      gather(D, tx, ty, ye_s, v);
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
    block_sum_published<3, kNT>(v, red + 16);  // stores drain during the sum
    return grid_allreduce<3, 2>(v, D.part, D.gs, epoch, red, nwg, me);
  };
  auto ring_w = [&](const double* wpub, double (&wv)[3]) {
    for (int d = 0; d < 3; d++) wv[d] = rnode >= 0 ? ld_coh(&wpub[3 * rnode + d]) : 0.0;
  };
  auto sw = [](double b, const double (&wv)[3], double (&sv)[3]) {
    for (int d = 0; d < 3; d++) sv[d] = fma(b, sv[d], wv[d]);
  };
  auto rs = [](double a, const double (&sv)[3], double (&rv)[3]) {
    for (int d = 0; d < 3; d++) rv[d] = fma(-a, sv[d], rv[d]);
  };

  double acc[3];
  matvec(D.p0, acc);
  if (!allreduce(acc)) return;
  double g = acc[0], pap = acc[1], rho = acc[2];
  const double bnorm = restart ? bnorm_in : sqrt(rho);
  int it = 0, converged = 0, breakdown = 0;
  const bool done0 = restart ? (bnorm < 1e-30 || sqrt(rho) < tol || sqrt(rho) / bnorm < tol)
                             : bnorm < 1e-30;
  if (done0) {
    converged = 1;
  } else if (maxit > 0 && fabs(pap) < 1e-30) {
    breakdown = 1;
    it = 1;
  } else {
    double a = g / pap;
    for (int d = 0; d < 3; d++) { p[d] = u[d]; s[d] = w[d]; }
    if (rpos >= 0) ring_w(D.p0, rs_);
    while (it < maxit) {
      double* wpub = (it & 1) ? D.p0 : D.p1;
      if (mine)
        for (int d = 0; d < 3; d++) x[d] = fma(a, p[d], x[d]);
      rs(a, s, r);
      if (rnode >= 0) rs(a, rs_, rr_);
      matvec(wpub, acc);
      if (!allreduce(acc)) return;
      it++;
      const double g1 = acc[0], d1 = acc[1];
      rho = acc[2];
      const double rn = sqrt(rho);
      if (rn < tol || rn / bnorm < tol) { converged = 1; break; }
      const double b = g1 / g;
      pap = d1 - b * g1 / a;
      if (fabs(pap) < 1e-30) { breakdown = 1; it++; break; }
      a = g1 / pap;
      g = g1;
      for (int d = 0; d < 3; d++) p[d] = fma(b, p[d], u[d]);
      sw(b, w, s);
      if (rpos >= 0) {
        double wv[3];
        ring_w(wpub, wv);
        if (rnode >= 0) sw(b, wv, rs_);
      }
    }
  }
  if (mine)
    for (int d = 0; d < 3; d++) D.x[3 * n + d] = x[d];
  if (me == 0 && threadIdx.x == 0) {
    D.st->iterations = it;
    D.st->converged = converged;
    D.st->breakdown = breakdown;
    D.st->final_residual = sqrt(rho);
    D.st->bnorm = bnorm;
  }
}

// restart: r = r0 (F - K x, precomputed), x kept, |F| from the cold run.
// RES: plates with one workgroup per tile take the resident path (batch_cg);
// without it the kernel keeps fewer registers and one more workgroup per CU.
template <bool RES>
__global__ __launch_bounds__(kNT) void k_batch_pcg(const PlateDesc* descs, const int* wg_plate,
                                                   double tol, int maxit, const double* r0pool,
                                                   const int64_t* dof_off) {
  __shared__ double p_s[3 * kMaxNP];
  __shared__ double ye_s[12 * kNT];
  __shared__ double red[80];
  const PlateDesc D = descs[wg_plate[blockIdx.x]];
  const unsigned me = blockIdx.x - D.wg0, nwg = D.nwg;
  if (RES && (int)nwg == D.ntiles) {  // one workgroup per tile: registers, one all-reduce
    // These plates are latency-bound chains of barriers sharing their CUs
    // with streaming workgroups of the strided plates: their waves issue first.
    __builtin_amdgcn_s_setprio(2);
    const bool rs = r0pool != nullptr;
    batch_cg(D, me, nwg, tol, maxit, rs ? r0pool + dof_off[wg_plate[blockIdx.x]] : D.F, rs,
             rs ? D.st->bnorm : 0.0, p_s, ye_s, red);
    return;
  }
  const int tx = threadIdx.x % D.TX, ty = threadIdx.x / D.TX;
  const bool node_thread = threadIdx.x < D.TX * D.TY;
  const int PW = D.TX + 2;
  unsigned epoch = 0;
  auto allreduce = [&](auto& v) -> bool {
    block_sum_published<sizeof(v) / sizeof(double), kNT>(v, red + 16);  // stores drain during the sum
    return grid_allreduce(v, D.part, D.gs, epoch, red, nwg, me);
  };
  // init: r = F (or r0), x = 0 (or kept), z = M^-1 r
  const bool restart = r0pool != nullptr;
  const double* src = restart ? r0pool + dof_off[wg_plate[blockIdx.x]] : D.F;
  const double bnorm_in = restart ? D.st->bnorm : 0.0;
  double acc[2] = {0, 0};
  for (int tile = me; tile < D.ntiles; tile += nwg) {
    const int gi = (tile % D.ntx) * D.TX + tx, gj = (tile / D.ntx) * D.TY + ty;
    if (node_thread && gi <= D.nx && gj <= D.ny) {
      const int64_t n = (int64_t)gj * (D.nx + 1) + gi;
      double rv[3] = {src[3 * n], src[3 * n + 1], src[3 * n + 2]}, zv[3];
      sym3_apply(D.minv + 6 * n, rv, zv);
      for (int d = 0; d < 3; d++) {
        st_coh(&D.z[3 * n + d], zv[d]);
        D.r[3 * n + d] = rv[d];
        if (!restart) D.x[3 * n + d] = 0.0;
        acc[0] += rv[d] * zv[d];
        acc[1] += rv[d] * rv[d];
      }
    }
  }
  if (!allreduce(acc)) return;
  double rz = acc[0], rr = acc[1];
  const double bnorm = restart ? bnorm_in : sqrt(rr);
  double beta = 0;
  int it = 0, converged = 0, breakdown = 0;
  const bool done0 = restart ? (bnorm < 1e-30 || sqrt(rr) < tol || sqrt(rr) / bnorm < tol)
                             : bnorm < 1e-30;
  if (done0) {
    converged = 1;
  } else {
    for (; it < maxit; it++) {
      const double* pold = (it & 1) ? D.p1 : D.p0;
      double* pnew = (it & 1) ? D.p0 : D.p1;
      double pap[1] = {0};
      for (int tile = me; tile < D.ntiles; tile += nwg) {
        const int i0 = (tile % D.ntx) * D.TX, j0 = (tile / D.ntx) * D.TY;
        stage<true>(D, i0, j0, D.z, it == 0 ? nullptr : pold, beta, p_s);
        __syncthreads();
        cells(D, i0, j0, p_s, ye_s);
        __syncthreads();
        const int gi = i0 + tx, gj = j0 + ty;
        if (node_thread && gi <= D.nx && gj <= D.ny) {
          const int64_t n = (int64_t)gj * (D.nx + 1) + gi;
          const unsigned fix = D.cmask[n];
          double v[3];
          gather(D, tx, ty, ye_s, v);
          const int c = (ty + 1) * PW + tx + 1;
          for (int d = 0; d < 3; d++) {
            const double pd = p_s[d * kMaxNP + c];
            const double ad = ((fix >> d) & 1u) ? pd : v[d];
            D.Ap[3 * n + d] = ad;
            pap[0] += pd * ad;
          }
        }
        // p of the tile's own nodes goes out row by row with consecutive
        // lanes on consecutive doubles: a device-coherent store is written
        // through per instruction, so node-strided stores (lane = node, one
        // instruction per component) would write each line three times
        {
          const int rw = min(D.TX, D.nx + 1 - i0), rows = min(D.TY, D.ny + 1 - j0);
          const int W3 = 3 * D.TX;
          for (int t = threadIdx.x; t < W3 * rows; t += kNT) {
            const int r = t / W3, k = t % W3;
            if (k < 3 * rw)
              st_coh(&pnew[3 * ((size_t)(j0 + r) * (D.nx + 1) + i0) + k],
                     p_s[(k % 3) * kMaxNP + (r + 1) * PW + k / 3 + 1]);
          }
        }
        __syncthreads();
      }
      if (!allreduce(pap)) return;
      if (fabs(pap[0]) < 1e-30) { breakdown = 1; it++; break; }
      const double alpha = rz / pap[0];
      acc[0] = 0; acc[1] = 0;
      // p comes in and z goes out through LDS row by row (whole lines per
      // coherent access); the node's thread works on its own entries
      for (int tile = me; tile < D.ntiles; tile += nwg) {
        const int i0 = (tile % D.ntx) * D.TX, j0 = (tile / D.ntx) * D.TY;
        const int rw = min(D.TX, D.nx + 1 - i0), rows = min(D.TY, D.ny + 1 - j0);
        const int W3 = 3 * D.TX;
        for (int t = threadIdx.x; t < W3 * rows; t += kNT) {
          const int r = t / W3, k = t % W3;
          if (k < 3 * rw) ye_s[t] = ld_coh(&pnew[3 * ((size_t)(j0 + r) * (D.nx + 1) + i0) + k]);
        }
        __syncthreads();
        const int gi = i0 + tx, gj = j0 + ty;
        if (node_thread && gi <= D.nx && gj <= D.ny) {
          const int64_t n = (int64_t)gj * (D.nx + 1) + gi;
          double* own = ye_s + 3 * (ty * D.TX + tx);
          double rv[3], zv[3];
          for (int d = 0; d < 3; d++) {
            D.x[3 * n + d] += alpha * own[d];
            rv[d] = D.r[3 * n + d] - alpha * D.Ap[3 * n + d];
            D.r[3 * n + d] = rv[d];
          }
          sym3_apply(D.minv + 6 * n, rv, zv);
          for (int d = 0; d < 3; d++) {
            own[d] = zv[d];
            acc[0] += rv[d] * zv[d];
            acc[1] += rv[d] * rv[d];
          }
        }
        __syncthreads();
        for (int t = threadIdx.x; t < W3 * rows; t += kNT) {
          const int r = t / W3, k = t % W3;
          if (k < 3 * rw) st_coh(&D.z[3 * ((size_t)(j0 + r) * (D.nx + 1) + i0) + k], ye_s[t]);
        }
        __syncthreads();
      }
      if (!allreduce(acc)) return;
      rr = acc[1];
      const double rn = sqrt(rr);
      if (rn < tol || rn / bnorm < tol) { converged = 1; it++; break; }
      beta = acc[0] / rz;
      rz = acc[0];
    }
  }
  if (me == 0 && threadIdx.x == 0) {
    D.st->iterations = it;
    D.st->converged = converged;
    D.st->breakdown = breakdown;
    D.st->final_residual = sqrt(rr);
    D.st->bnorm = bnorm;
  }
}

// y = K x (BC rows y[c] = x[c]) for every tile of every plate: block b works
// on tile tile_map[b].y of plate tile_map[b].x.
__global__ __launch_bounds__(kNT) void k_batch_apply(const PlateDesc* descs, const int2* tile_map,
                                                     const double* xpool, double* ypool,
                                                     const int64_t* dof_off) {
  __shared__ double p_s[3 * kMaxNP];
  __shared__ double ye_s[12 * kNT];
  const int2 tm = tile_map[blockIdx.x];
  const PlateDesc D = descs[tm.x];
  const double* x = xpool + dof_off[tm.x];
  double* y = ypool + dof_off[tm.x];
  const int i0 = (tm.y % D.ntx) * D.TX, j0 = (tm.y / D.ntx) * D.TY;
  stage<false>(D, i0, j0, x, nullptr, 0.0, p_s);
  __syncthreads();
  cells(D, i0, j0, p_s, ye_s);
  __syncthreads();
  const int tx = threadIdx.x % D.TX, ty = threadIdx.x / D.TX;
  const int gi = i0 + tx, gj = j0 + ty;
  if (threadIdx.x < D.TX * D.TY && gi <= D.nx && gj <= D.ny) {
    const int64_t n = (int64_t)gj * (D.nx + 1) + gi;
    const unsigned fix = D.cmask[n];
    double v[3];
    gather(D, tx, ty, ye_s, v);
    for (int d = 0; d < 3; d++) y[3 * n + d] = ((fix >> d) & 1u) ? x[3 * n + d] : v[d];
  }
}

// Per plate: |F - K x|^2 and |F|^2 (one workgroup per plate, fixed order).
__global__ __launch_bounds__(kNT) void k_batch_norms(const int64_t* dof_off, const double* F,
                                                     const double* Kx, double* out) {
  __shared__ double red[16];
  const int64_t a = dof_off[blockIdx.x], b = dof_off[blockIdx.x + 1];
  double v[2] = {0, 0};
  for (int64_t i = a + threadIdx.x; i < b; i += kNT) {
    const double d = F[i] - Kx[i];
    v[0] += d * d;
    v[1] += F[i] * F[i];
  }
  block_sum<2, kNT>(v, red);
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = v[0];
    out[2 * blockIdx.x + 1] = v[1];
  }
}

// out = F - Kx over the whole pool
__global__ void k_batch_residual(int64_t n, const double* F, const double* Kx, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = F[i] - Kx[i];
}

// 3x3 nodal block-Jacobi inverse from the cell matrices (solver.ts:188-304 rules).
__global__ void k_batch_block_inverse(const PlateDesc* descs, int plate, double* minv) {
  const PlateDesc D = descs[plate];
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)(D.nx + 1) * (D.ny + 1)) return;
  const int i = (int)(n % (D.nx + 1)), j = (int)(n / (D.nx + 1));
  double A[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  // (cell offset, corner of the node in that cell): SW cell -> NE corner, ...
  const int di[4] = {-1, 0, 0, -1}, dj[4] = {-1, -1, 0, 0}, corner[4] = {2, 3, 0, 1};
  for (int k = 0; k < 4; k++) {
    const int ci = i + di[k], cj = j + dj[k];
    if (ci < 0 || ci >= D.nx || cj < 0 || cj >= D.ny) continue;
    const double* K = D.kcell + (D.rcls[cj] * D.ncc + D.ccls[ci]) * 144;
    const int c = corner[k];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) A[3 * a + b] += K[(3 * c + a) * 12 + 3 * c + b];
  }
  block_inverse_sym(A, D.cmask[n], minv + 6 * n);
}

// Consistent pressure loads of plates [p0, p1), one thread per node (nodes
// node0 .. node0 + nodes - 1 of the pool): each plate's own q0 and Gaussian
// patches (q0[p - p0], patches poff[p - p0] .. poff[p - p0 + 1]), Q4 cells or
// split-quad triangles (pp_load.h).  The w load is rounded to float32 (the
// reference builds a Float32Array) and written to the pool's FP64 right-hand
// side Fb and/or to the float vector Ff (indexed from node0); rotation loads
// and constrained w are zero (solver.ts:400-407).
__global__ __launch_bounds__(256) void k_batch_pressure(const PlateDesc* descs,
                                                        const int64_t* dof_off, int p0, int p1,
                                                        int64_t node0, int64_t nodes,
                                                        const double* q0, const double* pt,
                                                        const int32_t* poff, double* Fb,
                                                        float* Ff) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= nodes) return;
  const int64_t ng = node0 + g;
  int lo = p0, hi = p1 - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (dof_off[mid] / 3 <= ng) lo = mid;
    else hi = mid - 1;
  }
  const PlateDesc& D = descs[lo];
  const int64_t n = ng - dof_off[lo] / 3;
  const int i = (int)(n % (D.nx + 1)), j = (int)(n / (D.nx + 1));
  const int k = lo - p0;
  const double* pp = pt + 4 * (int64_t)poff[k];
  const int np = poff[k + 1] - poff[k];
  const double f = D.dkt ? grid_node_load_split(D.xc, D.yc, D.nx, D.ny, i, j, q0[k], pp, np)
                         : grid_node_load_q4(D.xc, D.yc, D.nx, D.ny, i, j, q0[k], pp, np);
  const float fw = (D.cmask[n] & PP_FIX_W) ? 0.0f : (float)f;
  if (Fb) {
    Fb[3 * ng] = (double)fw;
    Fb[3 * ng + 1] = 0.0;
    Fb[3 * ng + 2] = 0.0;
  }
  if (Ff) {
    Ff[3 * g] = fw;
    Ff[3 * g + 1] = 0.0f;
    Ff[3 * g + 2] = 0.0f;
  }
}

__global__ void k_batch_zero(GridSync* gs, DevStats* st, int n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  gs[p].count = 0;
  gs[p].abort = 0;
  st[p].iterations = 0;
  st[p].converged = 0;
  st[p].breakdown = 0;
}

}  // namespace

}  // namespace pp

using namespace pp;

struct BatchPlate {
  int nx = 0, ny = 0, npe = 4;
  double E = 0, nu = 0, t = 0;
  int ncc = 0, ncr = 0;
  std::vector<uint16_t> ccls, rcls;
  std::vector<uint8_t> cmask;
  std::vector<double> ax, by;  // per column (a, 1/a), per row (b, 1/b)
  std::vector<double> kcell;
  int64_t dof_off = 0, node_off = 0, cls_off = 0, k_off = 0, geo_off = 0, coord_off = 0;
  int shape = 0, nwg = 1, last_iterations = 0, last_converged = 0;
};

struct pp_batch {
  pp_ctx* ctx = nullptr;
  std::vector<BatchPlate> P;
  int64_t total_dof = 0, total_nodes = 0;
  DevBuf<double> F, x, r, z, p0, p1, Ap, tmp, minv, kcell, geo, coord, part, norms;
  DevBuf<uint8_t> cmask;
  DevBuf<uint16_t> cls;
  DevBuf<PlateDesc> descs;
  DevBuf<int> wg_plate;
  DevBuf<int2> tile_map;
  DevBuf<int64_t> dof_off;
  DevBuf<GridSync> gs;
  DevBuf<DevStats> st;
  std::vector<PlateDesc> hdesc;
  int total_wg = 0, per_cu_max = 0, per_cu_res = 0;
  bool last_checked = false;
  pp_batch_schedule sched{};
};

namespace {

int ceil_div(int a, int b) { return (a + b - 1) / b; }

// exact classes of the widths (FP64 differences of the float32 coordinates)
int classes(const std::vector<double>& w, std::vector<uint16_t>& cls, std::vector<double>& vals) {
  std::map<double, int> m;
  cls.resize(w.size());
  for (size_t i = 0; i < w.size(); i++) {
    auto it = m.find(w[i]);
    if (it == m.end()) {
      const int k = (int)vals.size();
      m[w[i]] = k;
      vals.push_back(w[i]);
      cls[i] = (uint16_t)k;
    } else {
      cls[i] = (uint16_t)it->second;
    }
  }
  return (int)vals.size();
}

// 12x12 cell matrix of an a x b cell: Q4 directly, DKT from its two triangles
void cell_matrix(int npe, double a, double b, double E, double nu, double t, double* K) {
  std::fill(K, K + 144, 0.0);
  if (npe == 4) {
    const double xy[8] = {0, 0, a, 0, a, b, 0, b};
    element_ke_q4(xy, E, nu, t, K);
    return;
  }
  const double corners[4][2] = {{0, 0}, {a, 0}, {a, b}, {0, b}};
  const int tri[2][3] = {{0, 1, 2}, {0, 2, 3}};
  for (int q = 0; q < 2; q++) {
    double xy[6], Ke[81];
    for (int k = 0; k < 3; k++) {
      xy[2 * k] = corners[tri[q][k]][0];
      xy[2 * k + 1] = corners[tri[q][k]][1];
    }
    element_ke_dkt(xy, E, nu, t, Ke);
    for (int r = 0; r < 9; r++)
      for (int c = 0; c < 9; c++)
        K[(3 * tri[q][r / 3] + r % 3) * 12 + 3 * tri[q][c / 3] + c % 3] += Ke[r * 9 + c];
  }
}

// Is the mesh the rectangular grid (Q4) or its split-quad triangulation (DKT)?
bool grid_mesh(const pp_mesh* m) {
  const uint32_t nx = m->grid_nx, ny = m->grid_ny, W = nx + 1;
  if (nx == 0 || ny == 0 || (uint64_t)m->node_count != (uint64_t)W * (ny + 1)) return false;
  const uint64_t cells = (uint64_t)nx * ny;
  if (m->nodes_per_element == 4 ? m->element_count != cells
                                : (m->nodes_per_element != 3 || m->element_count != 2 * cells))
    return false;
  for (uint32_t j = 0; j < ny; j++)
    for (uint32_t i = 0; i < nx; i++) {
      const uint32_t n0 = j * W + i, n1 = n0 + 1, n2 = n0 + W + 1, n3 = n0 + W;
      const size_t c = (size_t)j * nx + i;
      if (m->nodes_per_element == 4) {
        const uint32_t* e = m->elements + 4 * c;
        if (e[0] != n0 || e[1] != n1 || e[2] != n2 || e[3] != n3) return false;
      } else {
        const uint32_t* e = m->elements + 6 * c;
        if (e[0] != n0 || e[1] != n1 || e[2] != n2 || e[3] != n0 || e[4] != n2 || e[5] != n3)
          return false;
      }
    }
  for (uint32_t j = 0; j <= ny; j++)
    for (uint32_t i = 0; i <= nx; i++) {
      const float* p = m->nodes + 2 * ((size_t)j * W + i);
      if (p[0] != m->nodes[2 * i] || p[1] != m->nodes[2 * ((size_t)j * W) + 1]) return false;
      if (i > 0 && !(p[0] > p[-2])) return false;
    }
  return true;
}

// Deal workgroups: every plate one, then one at a time to the plate whose
// estimated time (iterations x tiles per workgroup) is largest.
void deal(pp_batch* b, int total, bool by_iterations) {
  const int n = (int)b->P.size();
  std::vector<int> tiles(n);
  for (int p = 0; p < n; p++) {
    const BatchPlate& q = b->P[p];
    const int TX = kShapeTX[q.shape], TY = kShapeTY[q.shape];
    tiles[p] = ceil_div(q.nx + 1, TX) * ceil_div(q.ny + 1, TY);
    b->P[p].nwg = 1;
  }
  int used = n;
  auto cost = [&](int p, int w) {
    const double its = by_iterations ? std::max(1, b->P[p].last_iterations) : 1.0;
    return its * ceil_div(tiles[p], w);
  };
  std::vector<std::pair<double, int>> heap;
  for (int p = 0; p < n; p++) heap.push_back({cost(p, 1), p});
  std::make_heap(heap.begin(), heap.end());
  while (used < total && !heap.empty()) {
    std::pop_heap(heap.begin(), heap.end());
    const int p = heap.back().second;
    heap.pop_back();
    if (b->P[p].nwg >= tiles[p]) continue;  // one workgroup per tile at most
    b->P[p].nwg++;
    used++;
    heap.push_back({cost(p, b->P[p].nwg), p});
    std::push_heap(heap.begin(), heap.end());
  }
}

int apply_schedule(pp_batch* b, const pp_batch_schedule* s) {
  const int n = (int)b->P.size();
  if (s->tile_shape < 0 || s->tile_shape > kShapes) return fail(PP_ERR_INVALID, "bad tile shape");
  for (int p = 0; p < n; p++) {
    BatchPlate& q = b->P[p];
    if (s->tile_shape < kShapes) {
      q.shape = s->tile_shape;
    } else {  // per plate: the shape with the fewest tiles
      int best = 0, bt = 1 << 30;
      for (int k = 0; k < kShapes; k++) {
        const int t = ceil_div(q.nx + 1, kShapeTX[k]) * ceil_div(q.ny + 1, kShapeTY[k]);
        if (t < bt) { bt = t; best = k; }
      }
      q.shape = best;
    }
  }
  if (s->resident_plates != 0 && s->resident_plates != 1)
    return fail(PP_ERR_INVALID, "resident_plates must be 0 or 1");
  const int cap = s->resident_plates ? b->per_cu_res : b->per_cu_max;
  const int wpc = std::max(1, std::min(s->workgroups_per_cu, cap));
  const int total = std::max(n, wpc * b->ctx->cu_count);
  if (total > cap * b->ctx->cu_count)
    return fail(PP_ERR_UNSUPPORTED, "more plates than resident workgroups");
  deal(b, total, s->balance_iterations != 0);  // else in proportion to tiles only
  // descriptors, workgroup table, partial slots
  int wg = 0;
  int64_t parts = 0;
  for (auto& q : b->P) parts += 8 * q.nwg;
  int rc;
  if ((int64_t)b->part.n < parts && (rc = b->part.alloc(parts))) return rc;
  std::vector<int> wg_plate;
  std::vector<int2> tmap;
  parts = 0;
  for (int p = 0; p < n; p++) {
    BatchPlate& q = b->P[p];
    PlateDesc& D = b->hdesc[p];
    D.TX = kShapeTX[q.shape];
    D.TY = kShapeTY[q.shape];
    D.ntx = ceil_div(q.nx + 1, D.TX);
    D.ntiles = D.ntx * ceil_div(q.ny + 1, D.TY);
    D.wg0 = wg;
    D.nwg = q.nwg;
    D.part = b->part.p + parts;
    parts += 8 * q.nwg;
    for (int k = 0; k < q.nwg; k++) wg_plate.push_back(p);
    wg += q.nwg;
    for (int t = 0; t < D.ntiles; t++) tmap.push_back(make_int2(p, t));
  }
  b->total_wg = wg;
  if ((rc = b->wg_plate.alloc(wg)) || (rc = b->tile_map.alloc(tmap.size()))) return rc;
  hipStream_t st = b->ctx->stream;
  PP_HIP(hipMemcpyAsync(b->descs.p, b->hdesc.data(), n * sizeof(PlateDesc), hipMemcpyHostToDevice, st));
  PP_HIP(hipMemcpyAsync(b->wg_plate.p, wg_plate.data(), wg * sizeof(int), hipMemcpyHostToDevice, st));
  PP_HIP(hipMemcpyAsync(b->tile_map.p, tmap.data(), tmap.size() * sizeof(int2), hipMemcpyHostToDevice, st));
  PP_HIP(hipStreamSynchronize(st));
  b->sched = *s;
  return PP_OK;
}

// |F - K x| / |F| of every plate into b->norms (and rel, when given);
// returns the worst.
int true_residuals(pp_batch* b, double* worst, std::vector<double>* rel = nullptr) {
  hipStream_t s = b->ctx->stream;
  const int n = (int)b->P.size();
  hipLaunchKernelGGL(k_batch_apply, dim3(b->tile_map.n), dim3(kNT), 0, s,
                     (const PlateDesc*)b->descs.p, (const int2*)b->tile_map.p,
                     (const double*)b->x.p, b->tmp.p, (const int64_t*)b->dof_off.p);
  hipLaunchKernelGGL(k_batch_norms, dim3(n), dim3(kNT), 0, s, (const int64_t*)b->dof_off.p,
                     (const double*)b->F.p, (const double*)b->tmp.p, b->norms.p);
  PP_HIP(hipGetLastError());
  std::vector<double> h(2 * n);
  PP_HIP(hipMemcpyAsync(h.data(), b->norms.p, 2 * n * 8, hipMemcpyDeviceToHost, s));
  PP_HIP(hipStreamSynchronize(s));
  *worst = 0.0;
  if (rel) rel->assign(n, 0.0);
  for (int p = 0; p < n; p++) {
    const double v = h[2 * p + 1] > 0 ? sqrt(h[2 * p] / h[2 * p + 1]) : 0.0;
    *worst = std::max(*worst, v);
    if (rel) (*rel)[p] = v;
  }
  return PP_OK;
}

// One persistent launch (cold, or restart from x with r0 = F - K x in tmp)
// of at most maxit iterations per plate.
int launch(pp_batch* b, const pp_solve_opts* o, int maxit, bool restart,
           std::vector<DevStats>& ds) {
  hipStream_t s = b->ctx->stream;
  const int n = (int)b->P.size();
  hipLaunchKernelGGL(k_batch_zero, dim3(ceil_div(n, 256)), dim3(256), 0, s, b->gs.p, b->st.p, n);
  PP_HIP(hipMemsetAsync(b->part.p, 0, b->part.n * 8, s));
  const double* r0 = restart ? (const double*)b->tmp.p : (const double*)nullptr;
  if (b->sched.resident_plates)
    hipLaunchKernelGGL(k_batch_pcg<true>, dim3(b->total_wg), dim3(kNT), 0, s,
                       (const PlateDesc*)b->descs.p, (const int*)b->wg_plate.p, o->tolerance,
                       maxit, r0, (const int64_t*)b->dof_off.p);
  else
    hipLaunchKernelGGL(k_batch_pcg<false>, dim3(b->total_wg), dim3(kNT), 0, s,
                       (const PlateDesc*)b->descs.p, (const int*)b->wg_plate.p, o->tolerance,
                       maxit, r0, (const int64_t*)b->dof_off.p);
  PP_HIP(hipGetLastError());
  std::vector<GridSync> gs(n);
  ds.resize(n);
  PP_HIP(hipMemcpyAsync(ds.data(), b->st.p, n * sizeof(DevStats), hipMemcpyDeviceToHost, s));
  PP_HIP(hipMemcpyAsync(gs.data(), b->gs.p, n * sizeof(GridSync), hipMemcpyDeviceToHost, s));
  PP_HIP(hipStreamSynchronize(s));
  for (int p = 0; p < n; p++)
    if (gs[p].abort)
      return fail(PP_ERR_TIMEOUT, "grid barrier spin limit reached (plate " + std::to_string(p) + ")");
  return PP_OK;
}

// Solve every plate; with the true-residual check, plates whose FP64
// |F - K x| / |F| is above tol after the recurrence has converged are
// restarted from x with that residual (as the single-plate FP64 path), at
// most four times, within what is left of max_iterations (counted for the
// plate that has used the most).  A plate counts as converged when its
// recurrence converged and, with the check, its true residual is at most
// tol.  The device time covers all launches and checks.
int batch_run(pp_batch* b, const pp_solve_opts* o, pp_batch_stats* out) {
  constexpr int kMaxRestarts = 4;
  hipStream_t s = b->ctx->stream;
  const int n = (int)b->P.size();
  const bool check = !(o->flags & PP_FLAG_NO_TRUE_RESIDUAL);
  PP_HIP(hipEventRecord(b->ctx->ev0, s));
  std::vector<DevStats> ds;
  int rc;
  if ((rc = launch(b, o, o->max_iterations, false, ds))) return rc;
  std::vector<int> its(n), conv(n);
  for (int p = 0; p < n; p++) {
    its[p] = ds[p].iterations;
    conv[p] = ds[p].converged;
  }
  double worst = 0.0;
  if (check) {
    std::vector<double> rel;
    if ((rc = true_residuals(b, &worst, &rel))) return rc;
    for (int k = 0; k < kMaxRestarts && worst > o->tolerance; k++) {
      const int left = o->max_iterations - *std::max_element(its.begin(), its.end());
      if (left <= 0) break;
      hipLaunchKernelGGL(k_batch_residual, dim3(ceil_div((int)std::min<int64_t>(b->total_dof, 1 << 30), 256)),
                         dim3(256), 0, s, b->total_dof, (const double*)b->F.p,
                         (const double*)b->tmp.p, b->tmp.p);
      if ((rc = launch(b, o, left, true, ds))) return rc;
      for (int p = 0; p < n; p++) {
        its[p] += ds[p].iterations;
        conv[p] = ds[p].converged;
      }
      if ((rc = true_residuals(b, &worst, &rel))) return rc;
    }
    for (int p = 0; p < n; p++) conv[p] = conv[p] && rel[p] <= o->tolerance;
  }
  PP_HIP(hipEventRecord(b->ctx->ev1, s));
  PP_HIP(hipEventSynchronize(b->ctx->ev1));
  float ms = 0;
  PP_HIP(hipEventElapsedTime(&ms, b->ctx->ev0, b->ctx->ev1));
  int nconv = 0, maxit = 0;
  int64_t total = 0;
  for (int p = 0; p < n; p++) {
    b->P[p].last_iterations = its[p];
    b->P[p].last_converged = conv[p];
    nconv += conv[p];
    maxit = std::max(maxit, its[p]);
    total += its[p];
  }
  b->last_checked = check;
  if (out) {
    out->plates = n;
    out->converged = nconv;
    out->workgroups = b->total_wg;
    out->max_iterations = maxit;
    out->total_iterations = total;
    out->device_ms = ms;
    out->total_dof = b->total_dof;
    out->worst_true_rel_residual = worst;
  }
  return PP_OK;
}

}  // namespace

extern "C" {

int pp_batch_create(pp_ctx* ctx, int32_t n, const pp_mesh* meshes, const pp_material* mats,
                    const uint8_t* const* masks, pp_batch** out) {
  if (!ctx || !meshes || !mats || !masks || !out || n <= 0) return fail(PP_ERR_INVALID, "bad argument");
  *out = nullptr;
  PP_HIP(hipSetDevice(ctx->device));
  pp_batch* b = new pp_batch();
  b->ctx = ctx;
  ctx->refs++;
  b->P.resize(n);
  std::vector<uint8_t> hmask;
  std::vector<uint16_t> hcls;
  std::vector<double> hk, hgeo, hcoord;
  std::vector<int64_t> hoff(n + 1, 0);
  int rc = PP_OK;
  for (int p = 0; p < n && rc == PP_OK; p++) {
    const pp_mesh* m = meshes + p;
    BatchPlate& q = b->P[p];
    if (!m->nodes || !m->elements || !masks[p] || !grid_mesh(m)) {
      rc = fail(PP_ERR_UNSUPPORTED, "plate " + std::to_string(p) + " is not a rectangular grid (Q4 or split-quad DKT)");
      break;
    }
    if (!(mats[p].E > 0) || !(mats[p].t > 0) || !(mats[p].nu > -1.0 && mats[p].nu < 0.5) || mats[p].element_t) {
      rc = fail(PP_ERR_INVALID, "plate " + std::to_string(p) + ": material out of range (uniform thickness only)");
      break;
    }
    q.nx = m->grid_nx;
    q.ny = m->grid_ny;
    q.npe = m->nodes_per_element;
    q.E = mats[p].E;
    q.nu = mats[p].nu;
    q.t = mats[p].t;
    const int W = q.nx + 1;
    std::vector<double> wa(q.nx), hb(q.ny), va, vb;
    for (int i = 0; i < q.nx; i++) wa[i] = (double)m->nodes[2 * (i + 1)] - (double)m->nodes[2 * i];
    for (int j = 0; j < q.ny; j++)
      hb[j] = (double)m->nodes[2 * ((size_t)(j + 1) * W) + 1] - (double)m->nodes[2 * ((size_t)j * W) + 1];
    q.ncc = classes(wa, q.ccls, va);
    q.ncr = classes(hb, q.rcls, vb);
    if ((int64_t)q.ncc * q.ncr > 4096) {
      rc = fail(PP_ERR_UNSUPPORTED, "plate " + std::to_string(p) + ": too many distinct cell sizes (graded grid)");
      break;
    }
    for (int i = 0; i < q.nx; i++) {
      q.ax.push_back(wa[i]);
      q.ax.push_back(1.0 / wa[i]);
    }
    for (int j = 0; j < q.ny; j++) {
      q.by.push_back(hb[j]);
      q.by.push_back(1.0 / hb[j]);
    }
    q.geo_off = (int64_t)hgeo.size();
    hgeo.insert(hgeo.end(), q.ax.begin(), q.ax.end());
    hgeo.insert(hgeo.end(), q.by.begin(), q.by.end());
    q.coord_off = (int64_t)hcoord.size();
    for (int i = 0; i <= q.nx; i++) hcoord.push_back((double)m->nodes[2 * i]);
    for (int j = 0; j <= q.ny; j++) hcoord.push_back((double)m->nodes[2 * ((size_t)j * W) + 1]);
    q.kcell.resize((size_t)q.ncc * q.ncr * 144);
    for (int cr = 0; cr < q.ncr; cr++)
      for (int cc = 0; cc < q.ncc; cc++)
        cell_matrix(q.npe, va[cc], vb[cr], q.E, q.nu, q.t, &q.kcell[(size_t)(cr * q.ncc + cc) * 144]);
    q.cmask.assign(masks[p], masks[p] + m->node_count);
    q.dof_off = b->total_dof;
    q.node_off = b->total_nodes;
    q.cls_off = (int64_t)hcls.size();
    q.k_off = (int64_t)hk.size();
    hcls.insert(hcls.end(), q.ccls.begin(), q.ccls.end());
    hcls.insert(hcls.end(), q.rcls.begin(), q.rcls.end());
    hk.insert(hk.end(), q.kcell.begin(), q.kcell.end());
    hmask.insert(hmask.end(), q.cmask.begin(), q.cmask.end());
    b->total_dof += 3 * (int64_t)m->node_count;
    b->total_nodes += m->node_count;
    hoff[p + 1] = b->total_dof;
  }
  const int64_t nd = b->total_dof;
  if (rc == PP_OK) {
    (void)((rc = b->F.alloc(nd)) || (rc = b->x.alloc(nd)) || (rc = b->r.alloc(nd)) ||
           (rc = b->z.alloc(nd)) || (rc = b->p0.alloc(nd)) || (rc = b->p1.alloc(nd)) ||
           (rc = b->Ap.alloc(nd)) || (rc = b->tmp.alloc(nd)) ||
           (rc = b->minv.alloc(6 * b->total_nodes)) || (rc = b->kcell.alloc(hk.size())) ||
           (rc = b->geo.alloc(hgeo.size())) || (rc = b->coord.alloc(hcoord.size())) ||
           (rc = b->cmask.alloc(hmask.size())) || (rc = b->cls.alloc(hcls.size())) ||
           (rc = b->descs.alloc(n)) || (rc = b->gs.alloc(n)) || (rc = b->st.alloc(n)) ||
           (rc = b->norms.alloc(2 * n)) || (rc = b->dof_off.alloc(n + 1)));
  }
  if (rc) {
    pp_batch_destroy(b);
    return rc;
  }
  hipStream_t s = ctx->stream;
  auto up = [&](void* d, const void* h, size_t bytes) {
    return hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s);
  };
  if (up(b->kcell.p, hk.data(), hk.size() * 8) != hipSuccess ||
      up(b->cmask.p, hmask.data(), hmask.size()) != hipSuccess ||
      up(b->cls.p, hcls.data(), hcls.size() * 2) != hipSuccess ||
      up(b->geo.p, hgeo.data(), hgeo.size() * 8) != hipSuccess ||
      up(b->coord.p, hcoord.data(), hcoord.size() * 8) != hipSuccess ||
      up(b->dof_off.p, hoff.data(), hoff.size() * 8) != hipSuccess ||
      hipMemsetAsync(b->F.p, 0, nd * 8, s) != hipSuccess) {
    pp_batch_destroy(b);
    return fail(PP_ERR_HIP, "batch upload failed");
  }
  b->hdesc.resize(n);
  for (int p = 0; p < n; p++) {
    const BatchPlate& q = b->P[p];
    PlateDesc& D = b->hdesc[p];
    memset(&D, 0, sizeof(D));
    D.nx = q.nx;
    D.ny = q.ny;
    D.ncc = q.ncc;
    D.nclass = q.ncc * q.ncr;
    D.dkt = q.npe == 3;
    const double Dm = q.E * q.t * q.t * q.t / (12.0 * (1.0 - q.nu * q.nu));
    D.D11 = Dm;
    D.D12 = Dm * q.nu;
    D.D33 = Dm * (1.0 - q.nu) / 2.0;
    D.Ds = (5.0 / 6.0) * q.E / (2.0 * (1.0 + q.nu)) * q.t;
    D.ax = b->geo.p + q.geo_off;
    D.by = b->geo.p + q.geo_off + 2 * q.nx;
    D.xc = b->coord.p + q.coord_off;
    D.yc = b->coord.p + q.coord_off + q.nx + 1;
    D.cmask = b->cmask.p + q.node_off;
    D.ccls = b->cls.p + q.cls_off;
    D.rcls = b->cls.p + q.cls_off + q.nx;
    D.kcell = b->kcell.p + q.k_off;
    D.minv = b->minv.p + 6 * q.node_off;
    D.F = b->F.p + q.dof_off;
    D.x = b->x.p + q.dof_off;
    D.r = b->r.p + q.dof_off;
    D.z = b->z.p + q.dof_off;
    D.p0 = b->p0.p + q.dof_off;
    D.p1 = b->p1.p + q.dof_off;
    D.Ap = b->Ap.p + q.dof_off;
    D.gs = b->gs.p + p;
    D.st = b->st.p + p;
  }
  if (hipMemcpyAsync(b->descs.p, b->hdesc.data(), n * sizeof(PlateDesc), hipMemcpyHostToDevice, s) != hipSuccess) {
    pp_batch_destroy(b);
    return fail(PP_ERR_HIP, "batch upload failed");
  }
  for (int p = 0; p < n; p++) {
    hipLaunchKernelGGL(k_batch_block_inverse, dim3(ceil_div((b->P[p].nx + 1) * (b->P[p].ny + 1), 256)),
                       dim3(256), 0, s, (const PlateDesc*)b->descs.p, p, b->minv.p + 6 * b->P[p].node_off);
  }
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
    pp_batch_destroy(b);
    return fail(PP_ERR_HIP, "batch set-up kernels failed");
  }
  int per_cu = 0, per_cu_res = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_pcg<false>, kNT, 0) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_res, k_batch_pcg<true>, kNT, 0) != hipSuccess ||
      per_cu <= 0 || per_cu_res <= 0) {
    pp_batch_destroy(b);
    return fail(PP_ERR_HIP, "batched PCG kernel cannot be made resident");
  }
  b->per_cu_max = per_cu;
  b->per_cu_res = per_cu_res;
  pp_batch_schedule def;
  def.workgroups_per_cu = std::min(2, per_cu_res);
  def.tile_shape = 0;
  def.balance_iterations = 0;
  def.resident_plates = 1;
  if ((rc = apply_schedule(b, &def))) {
    pp_batch_destroy(b);
    return rc;
  }
  *out = b;
  return PP_OK;
}

int pp_batch_destroy(pp_batch* b) {
  if (!b) return PP_OK;
  pp_ctx* c = b->ctx;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  delete b;
  ctx_release(c);
  return PP_OK;
}

int32_t pp_batch_plate_count(const pp_batch* b) { return b ? (int32_t)b->P.size() : 0; }

int pp_batch_set_rhs(pp_batch* b, int32_t plate, const float* F) {
  if (!b || !F || plate < 0 || plate >= (int)b->P.size()) return fail(PP_ERR_INVALID, "bad argument");
  PP_LOCK(b->ctx);
  const BatchPlate& q = b->P[plate];
  const int64_t n = 3 * (int64_t)(q.nx + 1) * (q.ny + 1);
  std::vector<double> h(F, F + n);
  PP_HIP(hipMemcpyAsync(b->F.p + q.dof_off, h.data(), n * 8, hipMemcpyHostToDevice, b->ctx->stream));
  PP_HIP(hipStreamSynchronize(b->ctx->stream));
  return PP_OK;
}

int pp_batch_pressure_load(pp_batch* b, int32_t plate, double q0, const double* patches,
                           int32_t n_patches, float* F) {
  if (!b || !F || plate < 0 || plate >= (int)b->P.size() || n_patches < 0 ||
      (n_patches > 0 && !patches))
    return fail(PP_ERR_INVALID, "bad argument");
  PP_LOCK(b->ctx);
  PP_HIP(hipSetDevice(b->ctx->device));
  const BatchPlate& q = b->P[plate];
  const int64_t nn = (int64_t)(q.nx + 1) * (q.ny + 1);
  std::vector<double> hv(1 + 4 * (size_t)n_patches);
  hv[0] = q0;
  if (n_patches) memcpy(hv.data() + 1, patches, 4 * (size_t)n_patches * sizeof(double));
  const int32_t hpo[2] = {0, n_patches};
  DevBuf<double> v;
  DevBuf<int32_t> po;
  DevBuf<float> f;
  int rc;
  if ((rc = v.alloc(hv.size())) || (rc = po.alloc(2)) || (rc = f.alloc(3 * nn))) return rc;
  hipStream_t s = b->ctx->stream;
  PP_HIP(hipMemcpyAsync(v.p, hv.data(), hv.size() * 8, hipMemcpyHostToDevice, s));
  PP_HIP(hipMemcpyAsync(po.p, hpo, sizeof(hpo), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_batch_pressure, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s,
                     (const PlateDesc*)b->descs.p, (const int64_t*)b->dof_off.p, plate, plate + 1,
                     q.node_off, nn, (const double*)v.p, (const double*)v.p + 1,
                     (const int32_t*)po.p, (double*)nullptr, f.p);
  PP_HIP(hipGetLastError());
  PP_HIP(hipMemcpyAsync(F, f.p, 3 * nn * sizeof(float), hipMemcpyDeviceToHost, s));
  PP_HIP(hipStreamSynchronize(s));
  return PP_OK;
}

int pp_batch_pressure_loads(pp_batch* b, const double* q0, const double* patches,
                            const int32_t* patch_offsets) {
  if (!b || !q0 || !patch_offsets) return fail(PP_ERR_INVALID, "NULL argument");
  const int n = (int)b->P.size();
  if (patch_offsets[0] != 0) return fail(PP_ERR_INVALID, "patch_offsets[0] must be 0");
  for (int p = 0; p < n; p++)
    if (patch_offsets[p + 1] < patch_offsets[p])
      return fail(PP_ERR_INVALID, "patch_offsets must be nondecreasing");
  const int32_t total = patch_offsets[n];
  if (total > 0 && !patches) return fail(PP_ERR_INVALID, "patches is NULL");
  PP_LOCK(b->ctx);
  PP_HIP(hipSetDevice(b->ctx->device));
  std::vector<double> hv((size_t)n + 4 * (size_t)total);
  memcpy(hv.data(), q0, (size_t)n * sizeof(double));
  if (total) memcpy(hv.data() + n, patches, 4 * (size_t)total * sizeof(double));
  DevBuf<double> v;
  DevBuf<int32_t> po;
  int rc;
  if ((rc = v.alloc(hv.size())) || (rc = po.alloc((size_t)n + 1))) return rc;
  hipStream_t s = b->ctx->stream;
  PP_HIP(hipMemcpyAsync(v.p, hv.data(), hv.size() * 8, hipMemcpyHostToDevice, s));
  PP_HIP(hipMemcpyAsync(po.p, patch_offsets, ((size_t)n + 1) * sizeof(int32_t),
                        hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_batch_pressure, dim3((unsigned)((b->total_nodes + 255) / 256)), dim3(256),
                     0, s, (const PlateDesc*)b->descs.p, (const int64_t*)b->dof_off.p, 0, n,
                     (int64_t)0, b->total_nodes, (const double*)v.p, (const double*)v.p + n,
                     (const int32_t*)po.p, b->F.p, (float*)nullptr);
  PP_HIP(hipGetLastError());
  PP_HIP(hipStreamSynchronize(s));
  return PP_OK;
}

int pp_batch_set_schedule(pp_batch* b, const pp_batch_schedule* s) {
  if (!b || !s) return fail(PP_ERR_INVALID, "NULL argument");
  PP_LOCK(b->ctx);
  PP_HIP(hipSetDevice(b->ctx->device));
  return apply_schedule(b, s);
}

int pp_batch_get_schedule(const pp_batch* b, pp_batch_schedule* s, int32_t* wg_per_plate) {
  if (!b || !s) return fail(PP_ERR_INVALID, "NULL argument");
  PP_LOCK(b->ctx);
  *s = b->sched;
  if (wg_per_plate)
    for (size_t p = 0; p < b->P.size(); p++) wg_per_plate[p] = b->P[p].nwg;
  return PP_OK;
}

int pp_batch_solve(pp_batch* b, const pp_solve_opts* o, pp_batch_stats* st) {
  if (!b || !o || !st) return fail(PP_ERR_INVALID, "NULL argument");
  PP_LOCK(b->ctx);
  if (!(o->tolerance > 0) || o->max_iterations < 0) return fail(PP_ERR_INVALID, "bad tolerance/max_iterations");
  memset(st, 0, sizeof(*st));
  PP_HIP(hipSetDevice(b->ctx->device));
  return batch_run(b, o, st);
}

int pp_batch_autotune(pp_batch* b, const pp_solve_opts* o, pp_batch_schedule* best,
                      pp_batch_stats* best_stats, int32_t* trials) {
  if (!b || !o || !best) return fail(PP_ERR_INVALID, "NULL argument");
  PP_LOCK(b->ctx);
  PP_HIP(hipSetDevice(b->ctx->device));
  pp_solve_opts q = *o;
  q.flags |= PP_FLAG_NO_TRUE_RESIDUAL;
  pp_batch_stats s;
  int rc;
  // measure the iteration counts once (they do not depend on the schedule)
  pp_batch_schedule base;
  base.workgroups_per_cu = std::min(2, b->per_cu_max);
  base.tile_shape = 0;
  base.balance_iterations = 0;
  base.resident_plates = 0;
  if ((rc = apply_schedule(b, &base)) || (rc = batch_run(b, &q, &s))) return rc;
  std::vector<int> its(b->P.size());
  for (size_t p = 0; p < b->P.size(); p++) its[p] = b->P[p].last_iterations;
  double best_ms = 1e300, best_pred = 1e300;
  int n_trials = 0;
  pp_batch_schedule cand;
  // most workgroups first (usually fastest, so slow candidates are cut short)
  for (int wpc = std::min(4, b->per_cu_max); wpc >= 1; wpc--)
    for (int res = 1; res >= 0; res--)
    for (int shape = 0; shape <= kShapes; shape += kShapes)  // square tiles, or per-plate fewest tiles
      for (int bal = 1; bal >= 0; bal--) {
        cand.workgroups_per_cu = wpc;
        cand.tile_shape = shape;
        cand.balance_iterations = bal;
        cand.resident_plates = res;
        const int cap = res ? b->per_cu_res : b->per_cu_max;
        if (wpc > cap) continue;
        if ((int)b->P.size() > wpc * b->ctx->cu_count && wpc < cap) continue;
        for (size_t p = 0; p < b->P.size(); p++) b->P[p].last_iterations = its[p];
        if ((rc = apply_schedule(b, &cand))) return rc;
        // model: the slowest plate's iterations x tile sweeps per workgroup
        // (a resident plate's iteration counted as half a sweep: one pass,
        // one all-reduce); candidates predicted 1.5x slower than the best
        // one measured are skipped
        double pred = 0;
        for (size_t p = 0; p < b->P.size(); p++) {
          const BatchPlate& bp = b->P[p];
          const int tl = ceil_div(bp.nx + 1, kShapeTX[bp.shape]) * ceil_div(bp.ny + 1, kShapeTY[bp.shape]);
          const double sweeps = res && bp.nwg == tl ? 0.5 : (double)ceil_div(tl, bp.nwg);
          pred = std::max(pred, (double)its[p] * sweeps);
        }
        if (pred > 1.5 * best_pred) continue;
        // a second run only for candidates within 10% of the best so far
        double ms = 1e300;
        for (int k = 0; k < 2; k++) {
          if ((rc = batch_run(b, &q, &s))) return rc;
          ms = std::min(ms, (double)s.device_ms);
          if (ms > 1.1 * best_ms) break;
        }
        n_trials++;
        if (ms < best_ms) {
          best_ms = ms;
          best_pred = pred;
          *best = cand;
          if (best_stats) {
            *best_stats = s;
            best_stats->device_ms = ms;
          }
        }
      }
  for (size_t p = 0; p < b->P.size(); p++) b->P[p].last_iterations = its[p];
  if ((rc = apply_schedule(b, best))) return rc;
  if (trials) *trials = n_trials;
  return PP_OK;
}

int pp_batch_plate_result(pp_batch* b, int32_t plate, double* x64, float* x32, pp_solve_stats* st) {
  if (!b || plate < 0 || plate >= (int)b->P.size()) return fail(PP_ERR_INVALID, "bad argument");
  PP_LOCK(b->ctx);
  const BatchPlate& q = b->P[plate];
  const int64_t n = 3 * (int64_t)(q.nx + 1) * (q.ny + 1);
  hipStream_t s = b->ctx->stream;
  std::vector<double> h(n);
  PP_HIP(hipMemcpyAsync(h.data(), b->x.p + q.dof_off, n * 8, hipMemcpyDeviceToHost, s));
  DevStats ds;
  PP_HIP(hipMemcpyAsync(&ds, b->st.p + plate, sizeof(ds), hipMemcpyDeviceToHost, s));
  double nr[2];
  PP_HIP(hipMemcpyAsync(nr, b->norms.p + 2 * plate, 16, hipMemcpyDeviceToHost, s));
  PP_HIP(hipStreamSynchronize(s));
  if (x64) std::copy(h.begin(), h.end(), x64);
  if (x32)
    for (int64_t i = 0; i < n; i++) x32[i] = (float)h[i];
  if (st) {
    memset(st, 0, sizeof(*st));
    st->iterations = b->P[plate].last_iterations;
    st->converged = b->P[plate].last_converged;
    st->breakdown = ds.breakdown;
    st->sweeps = 1;
    st->final_residual = ds.final_residual;
    st->true_rel_residual = b->last_checked && nr[1] > 0 ? sqrt(nr[0] / nr[1]) : 0.0;
  }
  return PP_OK;
}

}  // extern "C"
