// Mixed precision on the structured path: FP32 inner PCG whose element
// apply is an MFMA contraction, inside FP64 iterative refinement.
//
// Outer (FP64, exact operator): x accumulates in FP64; each sweep starts
// from the true residual r = F - K x computed by the FP64 matrix-free
// kernel, and the solve ends when |F - K x| / |F| <= tol is measured there.
//
// Inner (FP32 vectors, FP64 dot products, FP64 3x3 block inverses): a
// persistent PCG on the correction e (K~ e = r) that runs until its
// recurrence residual has fallen by the inner factor (or below tol |F|).
// The host then folds e into x, recomputes r in FP64 and relaunches the
// inner solve from p = z on the new residual (iterative refinement).
// PP_FLAG_RELIABLE_UPDATE keeps the previous search direction instead
// (beta = r'.z' / r.z across the replacement).  That keeps the Krylov
// progress only while the FP32 recurrence residual tracks the true one;
// on these plates the gap reaches kappa * u32 > 1 (sweep 1 of the
// 1.47M-DOF plate ends at |F - K x| ~ 100 |F|: the rounding noise of the
// FP32 correction, high-frequency and removed by the next short sweep), and
// a kept direction then stalls, so restarting is the default.
//
// K~ is the shared element matrix of the plate's nominal a x b rectangle,
// split as Ke(t) = (t/t0)^3 Kb + (t/t0) Ks so a per-element thickness stays a
// scaling of the operand:  Y[16 x 16 el] = Kb16 (s3 X) + Ks16 (s1 X),  X the
// element DOF vectors (12 rows + 4 zero rows) as columns, on
// v_mfma_f32_16x16x4f32 with FP32 accumulation.  The float32 node
// coordinates make real element widths differ from the nominal one by
// ~1e-5 relative; the FP64 outer residual uses the exact per-element
// widths, so that difference only slows the refinement, never the answer.

#include <math.h>

#include <algorithm>
#include <vector>

#include "pp_grid.h"

namespace pp {

namespace {

using CfgM = Tile<15, 15, 256>;
static_assert(CfgM::NE % 16 == 0, "whole 16-element MFMA blocks");
typedef float f4 __attribute__((ext_vector_type(4)));

struct MixArgs {
  KGrid G;
  const double* r64;  // FP64 residual the sweep starts from
  float *e, *r, *z, *pb0, *pb1, *Ap;  // pb0 holds the previous p when resuming
  const double* minv;
  const float* kb;  // 16x16 row-major, bending part at t0
  const float* ks;  // 16x16 row-major, shear part at t0
  int uniform;      // 1: one matrix kb + ks, no per-element scaling
  int resume;       // 1: keep the previous search direction (beta = rz / rz_prev)
  double rz_prev;
  double* part;
  GridSync* gs;
  DevStats* st;
  double tol_abs;   // stop: |r| < tol_abs
  double drop;      // stop: |r| < drop * |r0|
  int maxit;
};

// Element apply of one tile on MFMA.  Wave w takes 16-element blocks w,
// w + NT/64, ...; for chunk kk (matrix columns 4kk .. 4kk+3) lane l holds
// A = K[l % 16][4kk + l / 16] and B = DOF 4kk + l / 16 of element l % 16 of
// the block; the accumulator gives lane l rows 4 (l / 16) .. + 3 of element
// l % 16.  Elements outside the grid contribute zero columns.
template <class C>
__device__ __forceinline__ void tile_elements_mfma(const KGrid& G, int i0, int j0, const float* p_s,
                                                   float* ye_s, const float (&ab)[4],
                                                   const float (&as)[4], bool uniform,
                                                   const float (&s1v)[4], const float (&s3v)[4]) {
  static_assert(C::NE / 16 <= 4 * (C::NT / 64), "four element blocks per wave");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int blk = wave + q * (C::NT / 64);
    if (blk >= C::NE / 16) break;
    const int le = blk * 16 + col;
    const int ex = le % C::EW, ey = le / C::EW;
    const int gi = i0 - 1 + ex, gj = j0 - 1 + ey;
    const bool valid = gi >= 0 && gi < G.nx && gj >= 0 && gj < G.ny;
    const int c0 = ey * C::PW + ex;
    const int cn[4] = {c0, c0 + 1, c0 + C::PW + 1, c0 + C::PW};
    const float s3 = s3v[q], s1 = s1v[q];
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      const int k = 4 * kk + kq;
      float b = 0.0f;
      if (valid && k < 12) b = p_s[(k % 3) * C::NP + cn[k / 3]];
      if (uniform) {
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(ab[kk], b, acc, 0, 0, 0);
      } else {
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(ab[kk], b * s3, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(as[kk], b * s1, acc, 0, 0, 0);
      }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int row = 4 * kq + i;
      if (row < 12) ye_s[row * C::NE + le] = acc[i];
    }
  }
}

// Thickness scales (t/t0, (t/t0)^3) of the elements this lane applies in
// tile_elements_mfma (1 on uniform plates and outside the grid).
template <class C>
__device__ __forceinline__ void elem_scales(const KGrid& G, int i0, int j0, bool uniform,
                                            float (&s1v)[4], float (&s3v)[4]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    s1v[q] = s3v[q] = 1.0f;
    const int blk = wave + q * (C::NT / 64);
    if (uniform || !G.elem_t || blk >= C::NE / 16) continue;
    const int le = blk * 16 + (lane & 15);
    const int gi = i0 - 1 + le % C::EW, gj = j0 - 1 + le / C::EW;
    if (gi >= 0 && gi < G.nx && gj >= 0 && gj < G.ny) {
      const double t = (double)G.elem_t[(size_t)gj * G.nx + gi] / G.t0;
      s1v[q] = (float)t;
      s3v[q] = (float)(t * t * t);
    }
  }
}

// Inner FP32 PCG on K~ e = r64, tiles strided over a resident grid, state in
// HBM (the two-reduction recurrence of pcg.ts:146-206 in this order).  Phase
// A loads the next tile's staging values into registers while the current
// tile is applied; the current tile's element scales and node mask are
// loaded before that prefetch, because loads complete in issue order.
template <class C>
__global__ __launch_bounds__(C::NT) void k_pcg32(MixArgs A) {
  static_assert(3 * C::TX * C::TY <= 12 * C::NE, "phase B stages a tile's p and z in ye_s");
  __shared__ float p_s[3 * C::NP];
  __shared__ float ye_s[12 * C::NE];
  __shared__ double red[80];
  const unsigned me = blockIdx.x, nwg = gridDim.x;
  const KGrid& G = A.G;
  const int ntx = tiles_x<C>(G), ntiles = tiles_total<C>(G);
  const TileBand band = xcd_tile_band(me, nwg, ntiles);
  const int tx = threadIdx.x % C::TX, ty = threadIdx.x / C::TX;
  const bool node_thread = threadIdx.x < C::TX * C::TY;
  unsigned epoch = 0;
  auto allreduce = [&](auto& v) -> bool {
    block_sum_published<sizeof(v) / sizeof(double), C::NT>(v, red + 16);  // stores drain during the sum
    return grid_allreduce(v, A.part, A.gs, epoch, red, nwg, me);
  };
  // MFMA A fragments (uniform plates fold bending + shear into one matrix)
  float ab[4], as[4];
  {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      const int idx = (lane & 15) * 16 + 4 * kk + (lane >> 4);
      ab[kk] = A.uniform ? A.kb[idx] + A.ks[idx] : A.kb[idx];
      as[kk] = A.ks[idx];
    }
  }

  // staging values (z, and p_old when there is one) of this thread's nodes
  // of a tile's staged block, loaded ahead of use
  constexpr int SPT = (C::NP + C::NT - 1) / C::NT;
  float sv[SPT][6];
  auto stage_load = [&](int tile, const float* pold) {
    const int i0 = (tile % ntx) * C::TX, j0 = (tile / ntx) * C::TY;
#pragma unroll
    for (int k = 0; k < SPT; k++) {
      const int idx = threadIdx.x + k * C::NT;
#pragma unroll
      for (int q = 0; q < 6; q++) sv[k][q] = 0.0f;
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
  auto stage_store = [&](bool with_pold, float beta) {
#pragma unroll
    for (int k = 0; k < SPT; k++) {
      const int idx = threadIdx.x + k * C::NT;
      if (idx < C::NP)
#pragma unroll
        for (int d = 0; d < 3; d++)
          p_s[d * C::NP + idx] = with_pold ? sv[k][d] + beta * sv[k][3 + d] : sv[k][d];
    }
  };

  // ---------------- init: r = (float) r64, e = 0, z = M^-1 r, r.z, r.r
  double acc[2] = {0, 0};
  for (int tile = band.first; tile < band.last; tile += band.step) {
    const int gi = (tile % ntx) * C::TX + tx, gj = (tile / ntx) * C::TY + ty;
    if (node_thread && gi <= G.nx && gj <= G.ny) {
      const int64_t n = (int64_t)gj * (G.nx + 1) + gi;
      double rv[3], zv[3];
      for (int d = 0; d < 3; d++) {
        const float rf = (float)A.r64[3 * n + d];
        A.r[3 * n + d] = rf;
        A.e[3 * n + d] = 0.0f;
        rv[d] = (double)rf;
      }
      sym3_apply(A.minv + 6 * n, rv, zv);
      for (int d = 0; d < 3; d++) {
        const float zf = (float)zv[d];
        st_coh(&A.z[3 * n + d], zf);
        acc[0] += rv[d] * (double)zf;
        acc[1] += rv[d] * rv[d];
      }
    }
  }
  if (!allreduce(acc)) return;
  double rz = acc[0], rr = acc[1];
  const double r0 = sqrt(rr);
  const bool resume = A.resume && A.rz_prev != 0.0;
  double beta = resume ? rz / A.rz_prev : 0.0;
  double rz_prev = A.rz_prev;
  int it = 0, converged = 0, breakdown = 0, updated = 0;
  if (r0 < A.tol_abs) {
    converged = 1;
  } else {
    for (; it < A.maxit; it++) {
      const float* pold = (it & 1) ? A.pb1 : A.pb0;
      float* pnew = (it & 1) ? A.pb0 : A.pb1;
      const bool fresh = it == 0 && !resume;
      // ---------------- phase A: p = z + beta pold, Ap = K~ p, p.Ap
      double pap[1] = {0};
      const float* po = fresh ? nullptr : pold;
      if (band.first < band.last) stage_load(band.first, po);
      for (int tile = band.first; tile < band.last; tile += band.step) {
        const int i0 = (tile % ntx) * C::TX, j0 = (tile / ntx) * C::TY;
        float s1v[4], s3v[4];
        elem_scales<C>(G, i0, j0, A.uniform != 0, s1v, s3v);
        const int gi = i0 + tx, gj = j0 + ty;
        const bool mine = node_thread && gi <= G.nx && gj <= G.ny;
        const int64_t nm = mine ? (int64_t)gj * (G.nx + 1) + gi : 0;
        const unsigned fixm = mine ? G.cmask[nm] : 0u;
        stage_store(po != nullptr, (float)beta);
        if (tile + band.step < band.last) stage_load(tile + band.step, po);
        __syncthreads();
        tile_elements_mfma<C>(G, i0, j0, p_s, ye_s, ab, as, A.uniform != 0, s1v, s3v);
        __syncthreads();
        if (mine) {
          const int64_t n = nm;
          const unsigned fix = fixm;
          float v[3];
          node_sum<C, float>(tx, ty, ye_s, v);
          const int c = (ty + 1) * C::PW + tx + 1;
          for (int d = 0; d < 3; d++) {
            const float pd = p_s[d * C::NP + c];
            const float ad = ((fix >> d) & 1u) ? pd : v[d];
            A.Ap[3 * n + d] = ad;
            pap[0] += (double)pd * (double)ad;
          }
        }
        // p of the tile's own nodes goes out row by row with consecutive
        // lanes on consecutive words: a device-coherent store is written
        // through per instruction, so node-strided stores (lane = node, one
        // instruction per component) would write each line three times
        {
          const int rw = min(C::TX, G.nx + 1 - i0), rows = min(C::TY, G.ny + 1 - j0);
          for (int t = threadIdx.x; t < 3 * C::TX * rows; t += C::NT) {
            const int r = t / (3 * C::TX), k = t % (3 * C::TX);
            if (k < 3 * rw)
              st_coh(&pnew[3 * ((size_t)(j0 + r) * (G.nx + 1) + i0) + k],
                     p_s[(k % 3) * C::NP + (r + 1) * C::PW + k / 3 + 1]);
          }
        }
        __syncthreads();
      }
      if (!allreduce(pap)) return;
      if (fabs(pap[0]) < 1e-30) { breakdown = 1; it++; break; }
      const float alpha = (float)(rz / pap[0]);
      // ---------------- phase B: e += alpha p, r -= alpha Ap, z = M^-1 r
      // p comes in and z goes out through LDS row by row (whole lines per
      // coherent access); the node's thread works on its own entries
      acc[0] = 0; acc[1] = 0;
      for (int tile = band.first; tile < band.last; tile += band.step) {
        const int i0 = (tile % ntx) * C::TX, j0 = (tile / ntx) * C::TY;
        const int rw = min(C::TX, G.nx + 1 - i0), rows = min(C::TY, G.ny + 1 - j0);
        constexpr int W3 = 3 * C::TX;
        for (int t = threadIdx.x; t < W3 * rows; t += C::NT) {
          const int r = t / W3, k = t % W3;
          if (k < 3 * rw) ye_s[t] = ld_coh(&pnew[3 * ((size_t)(j0 + r) * (G.nx + 1) + i0) + k]);
        }
        __syncthreads();
        const int gi = i0 + tx, gj = j0 + ty;
        if (node_thread && gi <= G.nx && gj <= G.ny) {
          const int64_t n = (int64_t)gj * (G.nx + 1) + gi;
          float* own = ye_s + 3 * (ty * C::TX + tx);
          double rv[3], zv[3];
          for (int d = 0; d < 3; d++) {
            A.e[3 * n + d] = fmaf(alpha, own[d], A.e[3 * n + d]);
            const float rf = fmaf(-alpha, A.Ap[3 * n + d], A.r[3 * n + d]);
            A.r[3 * n + d] = rf;
            rv[d] = (double)rf;
          }
          sym3_apply(A.minv + 6 * n, rv, zv);
          for (int d = 0; d < 3; d++) {
            const float zf = (float)zv[d];
            own[d] = zf;
            acc[0] += rv[d] * (double)zf;
            acc[1] += rv[d] * rv[d];
          }
        }
        __syncthreads();
        for (int t = threadIdx.x; t < W3 * rows; t += C::NT) {
          const int r = t / W3, k = t % W3;
          if (k < 3 * rw) st_coh(&A.z[3 * ((size_t)(j0 + r) * (G.nx + 1) + i0) + k], ye_s[t]);
        }
        __syncthreads();
      }
      if (!allreduce(acc)) return;
      rr = acc[1];
      rz_prev = rz;
      const double rn = sqrt(rr);
      if (rn < A.tol_abs) { converged = 1; it++; break; }
      if (rn < A.drop * r0) { updated = 1; it++; break; }
      beta = acc[0] / rz;
      rz = acc[0];
    }
  }
  if (me == 0 && threadIdx.x == 0) {
    A.st->iterations = it;
    A.st->converged = converged;
    A.st->breakdown = breakdown;
    A.st->updated = updated;
    A.st->final_residual = sqrt(rr);
    A.st->rz_prev = rz_prev;
  }
}

__global__ void k_mix_zero(GridSync* gs, DevStats* st) {
  if (threadIdx.x < 64) {
    gs->count = 0;
    gs->abort = 0;
  }
  if (threadIdx.x == 0) {
    st->iterations = 0;
    st->converged = 0;
    st->breakdown = 0;
    st->updated = 0;
  }
}

// x += e
__global__ void k_fold(int64_t n, const float* e, double* x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += (double)e[i];
}

int nblk(int64_t n) { return (int)((n + 255) / 256); }

// Largest deviation of the column widths / row heights from their mean,
// relative, and the means.
double grid_spread(const pp_plan* pl, double* a, double* b) {
  const int nx = pl->nx, ny = pl->ny, W = nx + 1;
  const double x0 = pl->h_nodes[0], x1 = pl->h_nodes[2 * nx];
  const double y0 = pl->h_nodes[1], y1 = pl->h_nodes[2 * ((size_t)ny * W) + 1];
  *a = (x1 - x0) / nx;
  *b = (y1 - y0) / ny;
  double dev = 0;
  for (int i = 0; i < nx; i++) {
    const double w = (double)pl->h_nodes[2 * (i + 1)] - (double)pl->h_nodes[2 * i];
    dev = std::max(dev, fabs(w - *a) / *a);
  }
  for (int j = 0; j < ny; j++) {
    const double h = (double)pl->h_nodes[2 * ((size_t)(j + 1) * W) + 1] -
                     (double)pl->h_nodes[2 * ((size_t)j * W) + 1];
    dev = std::max(dev, fabs(h - *b) / *b);
  }
  return dev;
}

// Bending and shear element matrices of the nominal a x b rectangle at t0,
// padded to 16 x 16: column c = the matrix-free element operator applied to
// unit vector c (the operator the FP64 path applies).
void nominal_ke(const pp_plan* pl, double a, double b, std::vector<float>& kb,
                std::vector<float>& ks) {
  const double t = pl->t;
  const double D = pl->E * t * t * t / (12.0 * (1.0 - pl->nu * pl->nu));
  const double Ds = (5.0 / 6.0) * pl->E / (2.0 * (1.0 + pl->nu)) * t;
  kb.assign(256, 0.0f);
  ks.assign(256, 0.0f);
  for (int c = 0; c < 12; c++) {
    double u[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, y[12];
    u[c] = 1.0;
    q4_rect_apply<double>(a, b, 1.0 / a, 1.0 / b, D, D * pl->nu, D * (1 - pl->nu) / 2, 0.0, u, y);
    for (int r = 0; r < 12; r++) kb[r * 16 + c] = (float)y[r];
    q4_rect_apply<double>(a, b, 1.0 / a, 1.0 / b, 0.0, 0.0, 0.0, Ds, u, y);
    for (int r = 0; r < 12; r++) ks[r * 16 + c] = (float)y[r];
  }
}

// Widths may differ from the nominal one by float32 rounding of the
// coordinates, not by design (a graded grid needs the FP64 path).
constexpr double kMaxSpread = 1e-3;
constexpr double kDefaultDrop = 1e-3;
constexpr int kDefaultSweeps = 200;

}  // namespace

int struct_solve_mixed(pp_plan* pl, const pp_solve_opts* o, pp_solve_stats* st) {
  hipStream_t s = pl->ctx->stream;
  const int64_t nd = pl->n_dof;
  int rc;
  if (pl->xs.n != (size_t)nd) {
    double a, b;
    const double dev = grid_spread(pl, &a, &b);
    if (dev > kMaxSpread)
      return fail(PP_ERR_UNSUPPORTED,
                  "mixed precision needs a uniform grid (one shared element matrix)");
    if ((rc = pl->xs.alloc(nd)) || (rc = pl->rs.alloc(nd)) || (rc = pl->zs.alloc(nd)) ||
        (rc = pl->ps0.alloc(nd)) || (rc = pl->ps1.alloc(nd)) || (rc = pl->Aps.alloc(nd)) ||
        (rc = pl->kmix.alloc(512)))
      return rc;
    std::vector<float> kb, ks;
    nominal_ke(pl, a, b, kb, ks);
    PP_HIP(hipMemcpyAsync(pl->kmix.p, kb.data(), 256 * 4, hipMemcpyHostToDevice, s));
    PP_HIP(hipMemcpyAsync(pl->kmix.p + 256, ks.data(), 256 * 4, hipMemcpyHostToDevice, s));
    PP_HIP(hipStreamSynchronize(s));
  }
  // persistent grid: three 256-thread workgroups per CU, tiles strided over
  // them (measured at 700x700: 62, 54, 50 and 57 us per iteration at 1-4 per CU)
  const int ntiles =
      ((pl->nx + 1 + CfgM::TX - 1) / CfgM::TX) * ((pl->ny + 1 + CfgM::TY - 1) / CfgM::TY);
  int per_cu = 0;
  PP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg32<CfgM>, CfgM::NT, 0));
  const int grid = std::min(ntiles, std::min(per_cu, 3) * pl->ctx->cu_count);
  if (grid <= 0) return fail(PP_ERR_HIP, "mixed PCG kernel cannot be made resident");
  if (pl->partials.n < (size_t)8 * grid && (rc = pl->partials.alloc(8 * (size_t)grid))) return rc;

  const bool keep_p = (o->flags & PP_FLAG_RELIABLE_UPDATE) != 0;
  const int max_sweeps = o->max_refinements > 0 ? o->max_refinements : kDefaultSweeps;
  MixArgs A;
  A.G = make_kgrid(pl);
  A.r64 = pl->tmp.p;
  A.e = pl->xs.p; A.r = pl->rs.p; A.z = pl->zs.p; A.Ap = pl->Aps.p;
  A.minv = pl->minv.p;
  A.kb = pl->kmix.p;
  A.ks = pl->kmix.p + 256;
  A.uniform = pl->var_t ? 0 : 1;
  A.part = pl->partials.p;
  A.gs = pl->gsync.p;
  A.st = pl->dstats.p;
  A.drop = o->inner_tolerance > 0 ? o->inner_tolerance : kDefaultDrop;

  PP_HIP(hipEventRecord(pl->ctx->ev0, s));
  PP_HIP(hipMemsetAsync(pl->x.p, 0, nd * 8, s));
  PP_HIP(hipMemcpyAsync(pl->tmp.p, pl->F.p, nd * 8, hipMemcpyDeviceToDevice, s));
  double res = 0, fnorm = 0;
  if ((rc = launch_norms(pl, pl->F.p, pl->F.p, &res, &fnorm))) return rc;  // |F|
  int total = 0, sweeps = 0, converged = 0, breakdown = 0;
  double rel = fnorm > 0 ? 1.0 : 0.0, rz_prev = 0.0, last_rr = 0.0;
  float* pcur = pl->ps0.p;  // holds the last search direction
  float* pother = pl->ps1.p;
  if (fnorm < 1e-30) converged = 1;
  while (!converged) {
    if (sweeps >= max_sweeps || total >= o->max_iterations) break;
    A.pb0 = pcur;
    A.pb1 = pother;
    A.resume = (sweeps > 0 && keep_p) ? 1 : 0;
    A.rz_prev = rz_prev;
    A.tol_abs = o->tolerance * fnorm;
    A.maxit = o->max_iterations - total;
    hipLaunchKernelGGL(k_mix_zero, dim3(1), dim3(64), 0, s, pl->gsync.p, pl->dstats.p);
    PP_HIP(hipMemsetAsync(pl->partials.p, 0, 8 * (size_t)grid * sizeof(double), s));
    hipLaunchKernelGGL(k_pcg32<CfgM>, dim3(grid), dim3(CfgM::NT), 0, s, A);
    PP_HIP(hipGetLastError());
    DevStats ds;
    GridSync gsh;
    PP_HIP(hipMemcpyAsync(&ds, pl->dstats.p, sizeof(ds), hipMemcpyDeviceToHost, s));
    PP_HIP(hipMemcpyAsync(&gsh, pl->gsync.p, sizeof(gsh), hipMemcpyDeviceToHost, s));
    PP_HIP(hipStreamSynchronize(s));
    if (gsh.abort) return fail(PP_ERR_TIMEOUT, "grid barrier spin limit reached");
    if (ds.iterations & 1) std::swap(pcur, pother);  // the last p written
    rz_prev = ds.rz_prev;
    last_rr = ds.final_residual;
    if (sweeps < PP_MAX_SWEEPS) st->sweep_iterations[sweeps] = ds.iterations;
    total += ds.iterations;
    breakdown = ds.breakdown;
    // x += e; r = F - K x in FP64 on the exact operator
    hipLaunchKernelGGL(k_fold, dim3(nblk(nd)), dim3(256), 0, s, nd, (const float*)pl->xs.p, pl->x.p);
    PP_HIP(hipGetLastError());
    if ((rc = struct_apply_k(pl, pl->x.p, pl->tmp.p))) return rc;
    if ((rc = launch_norms(pl, pl->F.p, pl->tmp.p, &res, &fnorm))) return rc;
    rel = res / fnorm;
    if (sweeps < PP_MAX_SWEEPS) st->sweep_residual[sweeps] = rel;
    sweeps++;
    if (rel <= o->tolerance) { converged = 1; break; }
    if (breakdown || ds.iterations == 0) break;
    if ((rc = launch_residual(pl, pl->F.p, pl->tmp.p, pl->tmp.p))) return rc;
  }
  PP_HIP(hipEventRecord(pl->ctx->ev1, s));
  PP_HIP(hipEventSynchronize(pl->ctx->ev1));
  float ms = 0;
  PP_HIP(hipEventElapsedTime(&ms, pl->ctx->ev0, pl->ctx->ev1));
  st->iterations = total;
  st->converged = converged;
  st->breakdown = breakdown;
  st->sweeps = sweeps;
  st->final_residual = last_rr;
  st->true_rel_residual = rel;
  st->device_ms = ms;
  return PP_OK;
}

}  // namespace pp
