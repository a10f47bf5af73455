// PCG preconditioned by one Galerkin multigrid V-cycle per iteration
// (PP_PRECOND_MULTIGRID), with the V-cycle and the PCG arithmetic fused into
// few launches.
//
// The hierarchy (levels, stencils, block inverses, coarsest dense inverse)
// is the one pp_mg.hip builds; this file only changes how a PCG iteration is
// launched.  Per iteration, with nu smoothing steps on level 0 and one on
// the coarser stencil levels:
//   A  q = K p with p = z + beta p_old staged per tile (FCG(1) direction),
//      partials of p.q and p.r; the last workgroup to finish sums them in a
//      fixed order and sets alpha (or the breakdown flag, pcg.ts:154)
//   B  x += alpha p, r -= alpha q, partials of r.r (the last workgroup
//      applies the stopping rule of pcg.ts:174), and the V-cycle's first
//      level-0 smoothing step from x = 0
//   V-cycle: level-0 steps 2 .. nu; then per level one kernel computing the
//      residual on the tile plus a one-node ring, restricting it onto the
//      coarse nodes the tile owns and taking the coarse level's first step;
//      the dense coarsest solve; per level one kernel that prolongs the
//      coarse correction while staging its tile and takes the first
//      post-smoothing step; level-0 steps 2 .. nu, the last with the
//      partials of z.q, whose last workgroup sets beta = -z.q / p.q.
// 19 launches per iteration on the 142^2 plate (4 levels, nu = 6) where the
// launch-per-operation form takes 34.  Every kernel returns at once when the
// solve has stopped, so the host can queue iterations ahead of the
// convergence flag it reads back.
//
// Cross-workgroup sums: each workgroup stores its partial with a
// device-scope store and counts itself in with a device-scope acq_rel
// fetch-add; the workgroup that arrives last (count nwg - 1) reads all
// partials with device-scope loads and sums them in a fixed order, so the
// result does not depend on which workgroup is last.  Nothing waits on
// another workgroup.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "pp_cnn.h"

namespace pp {

namespace {

struct MgfState {
  double alpha, beta, pq, rr, fnorm, tol;
  int it, done, converged, breakdown;
};

// Progress record of one iteration, written by the device into coherent
// pinned host memory (slot = iteration % ring) where the stopping rule (or
// the breakdown rule) is applied; seq = iteration + 1 is stored last.
struct MgfRec {
  double rr;
  int it, done, converged, breakdown;
  int seq, pad;
};

__device__ __forceinline__ void write_record(MgfRec* ring, int nring, int iter, const MgfState* S) {
  MgfRec* r = ring + iter % nring;
  r->rr = S->rr;
  r->it = S->it;
  r->done = S->done;
  r->converged = S->converged;
  r->breakdown = S->breakdown;
  __hip_atomic_store(&r->seq, iter + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// One level as the kernels see it.
struct LvView {
  int nx, ny;           // elements per side
  const float* st;      // stencil rows (levels >= 1): 84 floats per node (see k_rows)
  const float* minv;    // FP32 block inverse, 6 per node
  const uint8_t* mask;  // constraint bits (level 0) or dead bits (levels >= 1)
};

// Coarse level written by a residual + restriction kernel.
struct Coarse {
  int nx, ny;
  double* b;          // restricted residual
  const float* minv;  // non-null: also take the coarse level's first step x = c2 B b
  double c2;
  double* cur;        // that step's result
  double* prev;       // zeroed: the iterate before it
};

// Coarse correction prolonged while a post-smoothing step stages its tile.
struct Pro {
  const double* xc;  // coarse solution
  int cnx;
  double* xpro;      // out: the prolonged iterate on own nodes
};

// A fused reduction: per-workgroup partials [value][workgroup] (plain
// stores), summed by every workgroup of the consuming kernel (few
// producers) or by k_finish (many).
struct Red {
  double* part;
};

struct TileArgs {
  MgfRec* rec;        // KP: progress records (breakdown)
  int nrec;
  KGrid G;            // level 0 geometry (matrix-free operator)
  LvView L;
  const double* x;    // iterate to stage (KP: z)
  const double* y;    // KP: previous direction
  double* xo;         // STEP: out, read as the previous iterate when c1 != 0; KP: new direction
  const double* b;    // STEP / RES: right-hand side; KP: r
  double c1, c2;
  Pro pro;
  Coarse co;
  double* q;          // KP: out K p; ZQ: in
  Red red;
  // Partials of the reduction before this kernel, summed by every workgroup
  // itself (null: the producer's last workgroup left the scalar in the
  // state): KP: z.q of the last smoothing step; check: r.r of the update.
  const double* cin;
  int cn;
  int check;          // apply the stopping rule to the update before (with cin)
  MgfState* S;
};

// Sum of NV partials [k * n + i] in a fixed order by wave 0 (lane-strided
// partial sums, then the butterfly); the result is valid in every lane of
// wave 0 and identical in every workgroup.
template <int NV>
__device__ __forceinline__ void wave_sum_parts(const double* part, int n, double (&v)[NV]) {
  const int lane = threadIdx.x & 63;
  double s[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) s[k] = 0.0;
  for (int i = lane; i < n; i += 64)
#pragma unroll
    for (int k = 0; k < NV; k++) s[k] += part[(size_t)k * n + i];
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = wave_sum(s[k]);
}

// The stopping rule of pcg.ts:174 after an update with |r|^2 = rr, applied
// by every workgroup alike; workgroup 0 counts the iteration and records it.
// Returns true when the solve stops here.
__device__ __forceinline__ bool stop_rule(double rr, MgfState* S, MgfRec* rec, int nrec) {
  const double rn = sqrt(rr);
  const bool conv = rn < S->tol || rn / S->fnorm < S->tol;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S->rr = rr;
    S->it++;
    if (conv) {
      S->converged = 1;
      S->done = 1;
    }
    write_record(rec, nrec, S->it - 1, S);
  }
  return conv;
}

enum { M_STEP = 0, M_RES = 1, M_KP = 2 };


// stencil-level tile sides: steps 16 x 16 nodes, residuals 14 x 14 (a
// 2-node halo keeps the computed nodes at 16 x 16: one per thread)
constexpr int kTsS_ = 16, kTrS_ = 14;

template <int TS, int H>
struct TG {
  static constexpr int PW = TS + 2 * H, NP = PW * PW;  // staged nodes
  static constexpr int EW = PW - 1, NE = EW * EW;      // elements touching the computed nodes
  static constexpr int CW = PW - 2, NC = CW * CW;      // computed nodes: tile + (H - 1) ring
};

__device__ __forceinline__ double w1(int f, int c, int n) {
  if (min(2 * c, n) == f) return 1.0;
  const int d = f - 2 * c;
  return ((f & 1) && f < n && (d == 1 || d == -1)) ? 0.5 : 0.0;
}

// The coarse nodes a staged fine block of PW x PW nodes from (fi0, fj0)
// interpolates from: CP x CP nodes from (fi0 / 2, fj0 / 2) (fi0, fj0 >= 0).
template <int PW>
struct CoarseBlock {
  static constexpr int CP = PW / 2 + 2;
};

// Stage the coarse solution's block into LDS (SoA xc_s[d * CP^2 + local],
// zero past the coarse grid), so that prolonging the staged fine nodes reads
// LDS instead of up to four coarse nodes from memory per fine node.
template <int PW, int NT, bool COH>
__device__ __forceinline__ void stage_coarse(int fi0, int fj0, const double* xc, int cnx,
                                             int cny, double* xc_s) {
  constexpr int CP = CoarseBlock<PW>::CP;
  const int ci0 = fi0 / 2, cj0 = fj0 / 2;
  for (int idx = threadIdx.x; idx < CP * CP; idx += NT) {
    const int ci = ci0 + idx % CP, cj = cj0 + idx / CP;
    const bool ok = ci <= cnx && cj <= cny;
    const int64_t m = (int64_t)(ok ? cj : 0) * (cnx + 1) + (ok ? ci : 0);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      double v;
      if constexpr (COH) v = ld_coh(xc + 3 * m + d);
      else v = xc[3 * m + d];
      xc_s[d * CP * CP + idx] = ok ? v : 0.0;
    }
  }
}

// P xc at fine node (fi, fj), bilinear in index space over the coarse
// element's corners (fine node f sits between coarse nodes f/2 and
// (f+1)/2), read from the staged coarse block.
template <int PW>
__device__ __forceinline__ void prolong_node_lds(int fi, int fj, int fnx, int fny, int fi0,
                                                 int fj0, const double* xc_s, double (&s)[3]) {
  constexpr int CP = CoarseBlock<PW>::CP;
  const int i0 = fi / 2, i1 = (fi + 1) / 2, j0 = fj / 2, j1 = (fj + 1) / 2;
  s[0] = s[1] = s[2] = 0.0;
  const int ci[2] = {i0, i1}, cj[2] = {j0, j1};
  const int ni = i0 == i1 ? 1 : 2, nj = j0 == j1 ? 1 : 2;
  for (int b = 0; b < nj; b++)
    for (int a = 0; a < ni; a++) {
      const double w = w1(fi, ci[a], fnx) * w1(fj, cj[b], fny);
      const int l = (cj[b] - fj0 / 2) * CP + (ci[a] - fi0 / 2);
      for (int d = 0; d < 3; d++) s[d] += w * xc_s[d * CP * CP + l];
    }
}

// Stage src + beta y (+ P xc, masked) on the tile plus an H-node halo; zero
// outside the grid.  SoA: u_s[d * NP + local].
template <int TS, int H, int NT>
__device__ __forceinline__ void stage(int i0, int j0, int nx, int ny, const double* src,
                                      const double* y, double beta, const double* xc, int cnx,
                                      const uint8_t* mask, double* u_s, double* xc_s = nullptr) {
  using T = TG<TS, H>;
  const int fi0 = max(i0 - H, 0), fj0 = max(j0 - H, 0);
  if (xc) {
    stage_coarse<T::PW, NT, false>(fi0, fj0, xc, cnx, (ny + 1) / 2, xc_s);
    __syncthreads();
  }
  for (int idx = threadIdx.x; idx < T::NP; idx += NT) {
    const int lx = idx % T::PW, ly = idx / T::PW;
    const int gi = i0 - H + lx, gj = j0 - H + ly;
    double v[3] = {0.0, 0.0, 0.0};
    if (gi >= 0 && gi <= nx && gj >= 0 && gj <= ny) {
      const int64_t n = (int64_t)gj * (nx + 1) + gi;
      for (int d = 0; d < 3; d++) v[d] = src[3 * n + d];
      if (y)
        for (int d = 0; d < 3; d++) v[d] += beta * y[3 * n + d];
      if (xc) {
        double s[3];
        prolong_node_lds<T::PW>(gi, gj, nx, ny, fi0, fj0, xc_s, s);
        const unsigned msk = mask[n];
        for (int d = 0; d < 3; d++)
          if (!(msk >> d & 1)) v[d] += s[d];
      }
    }
    for (int d = 0; d < 3; d++) u_s[d * T::NP + idx] = v[d];
  }
}

// Stage src into u_s and y into y_s (both SoA, tile plus H-node halo, zero
// outside the grid) in one pass, so the loads of both are in flight together.
template <int TS, int H, int NT>
__device__ __forceinline__ void stage_two(int i0, int j0, int nx, int ny, const double* src,
                                          const double* y, double* u_s, double* y_s) {
  using T = TG<TS, H>;
  for (int idx = threadIdx.x; idx < T::NP; idx += NT) {
    const int lx = idx % T::PW, ly = idx / T::PW;
    const int gi = i0 - H + lx, gj = j0 - H + ly;
    double v[3] = {0.0, 0.0, 0.0}, w[3] = {0.0, 0.0, 0.0};
    if (gi >= 0 && gi <= nx && gj >= 0 && gj <= ny) {
      const int64_t n = (int64_t)gj * (nx + 1) + gi;
      for (int d = 0; d < 3; d++) v[d] = src[3 * n + d];
      for (int d = 0; d < 3; d++) w[d] = y[3 * n + d];
    }
    for (int d = 0; d < 3; d++) {
      u_s[d * T::NP + idx] = v[d];
      y_s[d * T::NP + idx] = w[d];
    }
  }
}

// Stencil row of a level >= 1 in the fused form: per node 84 floats (21
// float4), the 9 blocks of its row of A in the order the apply adds them:
// the centre and forward blocks E, NW, N, NE as stored in the planes, then
// for the backward neighbours W, SE, S, SW the block A(n, m) = A(m, n)^T,
// transposed from m's forward block (zero outside the grid); 3 floats pad.
constexpr int kRow = 84;

__global__ void k_rows(int nx, int ny, const float* st, float* rows) {
  const int64_t nn = (int64_t)(nx + 1) * (ny + 1);
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  const int i = (int)(n % (nx + 1)), j = (int)(n / (nx + 1));
  float* R = rows + n * kRow;
  for (int k = 0; k < 45; k++) R[k] = st[k * nn + n];
  const int fdj[5] = {0, 0, 1, 1, 1}, fdi[5] = {0, 1, -1, 0, 1};
  for (int sb = 1; sb < 5; sb++) {
    const int jj = j - fdj[sb], ii = i - fdi[sb];
    const bool ok = jj >= 0 && ii >= 0 && ii <= nx;
    const int64_t m = (int64_t)jj * (nx + 1) + ii;
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++)
        R[45 + (sb - 1) * 9 + 3 * r + c] = ok ? st[(sb * 9 + 3 * c + r) * nn + m] : 0.0f;
  }
  for (int k = 81; k < kRow; k++) R[k] = 0.0f;
}

struct StRow {
  float4 v[kRow / 4];
};

__device__ __forceinline__ void load_st_row(int nx, const float* rows, int gi, int gj, StRow& R) {
  const int64_t n = (int64_t)gj * (nx + 1) + gi;
  const float4* p = (const float4*)(rows + n * kRow);
#pragma unroll
  for (int k = 0; k < kRow / 4; k++) R.v[k] = p[k];
}

// (A x)_n from the prefetched row and the staged x: the 9 blocks in row
// order, block by block as pp_mg.hip's stencil rows add them.
template <int PW, int NP>
__device__ __forceinline__ void stencil_apply(const StRow& R, int sx, int sy, const double* u_s,
                                              double (&y)[3]) {
  const float* f = (const float*)R.v;
  // neighbour offsets (dj, di): centre, E, NW, N, NE, W, SE, S, SW
  const int odj[9] = {0, 0, 1, 1, 1, 0, -1, -1, -1}, odi[9] = {0, 1, -1, 0, 1, -1, 1, 0, -1};
  y[0] = y[1] = y[2] = 0.0;
#pragma unroll
  for (int b = 0; b < 9; b++) {
    const int li = (sy + odj[b]) * PW + sx + odi[b];
    const double x0 = u_s[li], x1 = u_s[NP + li], x2 = u_s[2 * NP + li];
    const float* s = f + 9 * b;
    y[0] += (double)s[0] * x0 + (double)s[1] * x1 + (double)s[2] * x2;
    y[1] += (double)s[3] * x0 + (double)s[4] * x1 + (double)s[5] * x2;
    y[2] += (double)s[6] * x0 + (double)s[7] * x1 + (double)s[8] * x2;
  }
}

// One tile kernel for every fused level-0 phase (the matrix-free Q4 operator
// with the constraint rows); TS x TS nodes per workgroup, staged with an
// H-node halo; one computed node and one element per thread.
//   M_STEP  xo = x + c1 (x - xo) + c2 B (b - A x) on the tile (PRO: x is
//           staged as x + P xc and also written to xpro; ZQ: partials of
//           xo.q, the last workgroup sets beta = -z.q / p.q)
//   M_RES   t = b - A x on the tile and a one-node ring (H = 2), restricted
//           onto the coarse nodes whose fine node min(2I, n) the tile holds,
//           optionally with the coarse first step
//   M_KP    p = z + beta p_old staged, q = K p, p and q stored, partials of
//           p.q and p.r; the last workgroup sets alpha or the breakdown flag
// Every operand a thread needs after the staging (its node's right-hand
// side, block inverse, previous iterate and constraint bits, its element's
// widths and material) is loaded before the staging, so the kernel waits on
// global memory once.
template <int TS, int H, int NT, int MODE, bool PRO, bool ZQ>
__global__ __launch_bounds__(NT) void k_tile(TileArgs A) {
  using T = TG<TS, H>;
  static_assert(T::NC <= NT, "one computed node per thread");
  static_assert(T::NE <= NT, "one element per thread");
  __shared__ double u_s[3 * T::NP];
  __shared__ double y_s[MODE == M_KP ? 3 * T::NP : 1];  // KP: p_old (p = z + beta p_old)
  __shared__ double ye_s[12 * T::NE];
  __shared__ double t_s[MODE == M_RES ? 3 * T::NC : 1];
  __shared__ uint8_t m_s[MODE == M_RES ? T::NC : 1];
  __shared__ double red_s[2 * (NT / 64) + 2];
  __shared__ double in_s[2];
  __shared__ double xc_s[PRO ? 3 * CoarseBlock<T::PW>::CP * CoarseBlock<T::PW>::CP : 1];
  if (A.S->done) return;
  const int nx = A.L.nx, ny = A.L.ny;
  const int ntx = (nx + 1 + TS - 1) / TS;
  const int tile = xcd_band_tile((int)blockIdx.x, (int)gridDim.x);
  const int i0 = (tile % ntx) * TS, j0 = (tile / ntx) * TS;
  // ---- this thread's computed node and its operands
  const int c = threadIdx.x;
  const int cx = c % T::CW, cy = c / T::CW;
  const int sx = cx + 1, sy = cy + 1;
  const int gi = i0 - H + sx, gj = j0 - H + sy;
  const bool in = c < T::NC && gi >= 0 && gi <= nx && gj >= 0 && gj <= ny;
  const int64_t n = in ? (int64_t)gj * (nx + 1) + gi : 0;
  unsigned fix = 0;
  double bv[3] = {0.0, 0.0, 0.0}, xov[3] = {0.0, 0.0, 0.0}, qv[3] = {0.0, 0.0, 0.0};
  float mv[6] = {0, 0, 0, 0, 0, 0};
  if (in) {
    fix = A.L.mask[n];
    for (int d = 0; d < 3; d++) bv[d] = A.b[3 * n + d];
    if (MODE == M_STEP) {
      for (int k = 0; k < 6; k++) mv[k] = A.L.minv[6 * n + k];
      if (A.c1 != 0.0)
        for (int d = 0; d < 3; d++) xov[d] = A.xo[3 * n + d];
      if (ZQ)
        for (int d = 0; d < 3; d++) qv[d] = A.q[3 * n + d];
    }
  }
  // ---- this thread's element (level 0)
  const int le = threadIdx.x;
  const int ex = le % T::EW, ey = le / T::EW;
  const int egi = i0 - H + ex, egj = j0 - H + ey;
  const bool ein = le < T::NE && egi >= 0 && egi < A.G.nx && egj >= 0 && egj < A.G.ny;
  double ea = 0, eb = 0, eia = 0, eib = 0, eD = 0, eDs = 0;
  if (ein) {
    ea = A.G.ax[2 * egi];
    eia = A.G.ax[2 * egi + 1];
    eb = A.G.by[2 * egj];
    eib = A.G.by[2 * egj + 1];
    elem_material(A.G, egi, egj, eD, eDs);
  }
  // ---- (RES) this thread's coarse node and its block inverse
  constexpr int CWc = TS / 2 + 2;
  int cI = -1, cJ = -1;
  float cm[6] = {0, 0, 0, 0, 0, 0};
  if (MODE == M_RES && (int)threadIdx.x < CWc * CWc) {
    const int I = i0 / 2 + threadIdx.x % CWc, J = j0 / 2 + threadIdx.x / CWc;
    const int fi0 = min(2 * I, nx), fj0 = min(2 * J, ny);
    if (I <= A.co.nx && J <= A.co.ny && fi0 >= i0 && fi0 < i0 + TS && fj0 >= j0 &&
        fj0 < j0 + TS) {
      cI = I;
      cJ = J;
      if (A.co.minv)
        for (int q = 0; q < 6; q++) cm[q] = A.co.minv[6 * ((int64_t)J * (A.co.nx + 1) + I) + q];
    }
  }
  // ---- consumed partials (wave 0; the sum reaches the other waves through LDS)
  double cv[1] = {0.0};
  if (A.cin && threadIdx.x < 64) wave_sum_parts<1>(A.cin, A.cn, cv);
  if constexpr (MODE == M_KP)
    stage_two<TS, H, NT>(i0, j0, nx, ny, A.x, A.y, u_s, y_s);
  else
    stage<TS, H, NT>(i0, j0, nx, ny, A.x, nullptr, 0.0, PRO ? A.pro.xc : nullptr, A.pro.cnx,
                     A.L.mask, u_s, xc_s);
  if (threadIdx.x == 0) in_s[0] = cv[0];
  __syncthreads();
  if (A.check && stop_rule(in_s[0], A.S, A.rec, A.nrec)) return;
  // KP: beta = -z.q / p.q of the last direction (0 on the first iteration)
  const double beta =
      MODE == M_KP ? (A.cin ? -in_s[0] / A.S->pq : A.S->beta) : 0.0;
  {
    if (le < T::NE) {
      double y[12];
      if (ein) {
        const int c0 = ey * T::PW + ex, c1 = c0 + 1, c2 = c0 + T::PW + 1, c3 = c0 + T::PW;
        constexpr int NP = T::NP;
        const int cc[4] = {c0, c1, c2, c3};
        double u[12];
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
          for (int d = 0; d < 3; d++)
            u[3 * k + d] = MODE == M_KP ? u_s[d * NP + cc[k]] + beta * y_s[d * NP + cc[k]]
                                        : u_s[d * NP + cc[k]];
        q4_rect_apply_fast<double>(ea, eb, eia, eib, eD, eD * A.G.nu, eD * (1.0 - A.G.nu) / 2.0,
                                   eDs, u, y);
      } else {
#pragma unroll
        for (int k = 0; k < 12; k++) y[k] = 0.0;
      }
#pragma unroll
      for (int k = 0; k < 12; k++) ye_s[k * T::NE + le] = y[k];
    }
    __syncthreads();
  }
  double acc[2] = {0.0, 0.0};
  if (c < T::NC) {
    if (!in) {
      if (MODE == M_RES) {
        for (int d = 0; d < 3; d++) t_s[d * T::NC + c] = 0.0;
        m_s[c] = 0;
      }
    } else {
      const int si = sy * T::PW + sx;
      double own[3];  // the staged value of this node (KP: p = z + beta p_old)
      for (int d = 0; d < 3; d++)
        own[d] = MODE == M_KP ? u_s[d * T::NP + si] + beta * y_s[d * T::NP + si]
                              : u_s[d * T::NP + si];
      double Kx[3];
      {
        constexpr int EW = T::EW, NE = T::NE;
        const int sw = cy * EW + cx, se = sw + 1, ne = sw + EW + 1, nw = sw + EW;
        for (int d = 0; d < 3; d++) {
          const double v = ye_s[(6 + d) * NE + sw] + ye_s[(9 + d) * NE + se] + ye_s[d * NE + ne] +
                           ye_s[(3 + d) * NE + nw];
          Kx[d] = (fix >> d) & 1u ? own[d] : v;
        }
      }
      if (MODE == M_RES) {
        for (int d = 0; d < 3; d++) t_s[d * T::NC + c] = bv[d] - Kx[d];
        m_s[c] = (uint8_t)fix;
      } else if (MODE == M_KP) {
        for (int d = 0; d < 3; d++) {
          const double p = own[d];
          A.xo[3 * n + d] = p;
          A.q[3 * n + d] = Kx[d];
          acc[0] += p * Kx[d];
          acc[1] += p * bv[d];
        }
      } else {
        double r[3], z[3];
        for (int d = 0; d < 3; d++) r[d] = bv[d] - Kx[d];
        sym3_apply(mv, r, z);
        for (int d = 0; d < 3; d++) {
          const double xc = own[d];
          const double dn = (A.c1 != 0.0 ? A.c1 * (xc - xov[d]) : 0.0) + A.c2 * z[d];
          const double xn = xc + dn;
          A.xo[3 * n + d] = xn;
          if (PRO) A.pro.xpro[3 * n + d] = xc;
          if (ZQ) acc[0] += xn * qv[d];
        }
      }
    }
  }
  if constexpr (MODE == M_RES) {
    __syncthreads();
    const Coarse& co = A.co;
    const int fnx = nx, fny = ny, cnx = co.nx;
    static_assert(CWc * CWc <= NT, "one coarse node per thread");
    if (cI >= 0) {
      const int I = cI, J = cJ;
      const int64_t nc = (int64_t)J * (cnx + 1) + I;
      double s[3] = {0.0, 0.0, 0.0};
      for (int b = -1; b <= 1; b++) {
        const int fj = 2 * J + b;
        if (fj < 0 || fj > fny) continue;
        for (int a = -1; a <= 1; a++) {
          const int fi = 2 * I + a;
          if (fi < 0 || fi > fnx) continue;
          const double w = w1(fi, I, fnx) * w1(fj, J, fny);
          const int ci = (fj - (j0 - H + 1)) * T::CW + (fi - (i0 - H + 1));
          const unsigned msk = m_s[ci];
          for (int d = 0; d < 3; d++)
            if (!(msk >> d & 1)) s[d] += w * t_s[d * T::NC + ci];
        }
      }
      for (int d = 0; d < 3; d++) co.b[3 * nc + d] = s[d];
      if (co.minv) {
        double z[3];
        sym3_apply(cm, s, z);
        for (int d = 0; d < 3; d++) {
          co.cur[3 * nc + d] = 0.0 + co.c2 * z[d];
          co.prev[3 * nc + d] = 0.0;
        }
      }
    }
    return;
  }
  if constexpr (MODE == M_KP || ZQ) {
    constexpr int NV = MODE == M_KP ? 2 : 1;
    double v[NV];
    for (int k = 0; k < NV; k++) v[k] = acc[k];
    block_sum<NV, NT>(v, red_s);
    if (threadIdx.x == 0)
      for (int k = 0; k < NV; k++) A.red.part[(size_t)k * gridDim.x + blockIdx.x] = v[k];
  }
}

// ----------------------------------------------------------- stencil levels
//
// One tile of a stencil level (levels >= 1), as k_tile does it for level 0.
// COH: vectors are read with device-coherent loads and stored with
// device-coherent stores, for a caller whose workgroups exchange them within
// one launch; plain loads and stores otherwise.  (A persistent launch of the
// whole coarse part of the V-cycle, levels 1 and 2 and the dense solve of
// the 142^2 plate behind grid barriers, was measured slower than separate
// launches: 45.7 against 32.2 us per V-cycle, the coherent loads and the
// barrier of each phase costing more than a launch.)

template <bool COH>
__device__ __forceinline__ double ldv(const double* p) {
  if constexpr (COH) return ld_coh(p);
  else return *p;
}
template <bool COH>
__device__ __forceinline__ void stv(double* p, double v) {
  if constexpr (COH) st_coh(p, v);
  else *p = v;
}

// LDS of a stencil tile: the staged x (tile + 2-node halo at most), the
// residual and the constraint bits of the computed nodes (tile + 1 ring).
struct StSmem {
  double u[3 * 18 * 18];
  double t[3 * 16 * 16];
  double xc[3 * CoarseBlock<18>::CP * CoarseBlock<18>::CP];  // PRO: the coarse block
  uint8_t m[16 * 16];
};

template <int TS, int H, int MODE, bool PRO, bool COH>
__device__ __forceinline__ void st_tile(const TileArgs& A, int tile, StSmem& sm) {
  using T = TG<TS, H>;
  static_assert(T::NP <= 18 * 18 && T::NC <= 16 * 16, "tile larger than the LDS plan");
  static_assert(T::NC <= 256, "one computed node per thread");
  const int nx = A.L.nx, ny = A.L.ny;
  const int ntx = (nx + 1 + TS - 1) / TS;
  const int i0 = (tile % ntx) * TS, j0 = (tile / ntx) * TS;
  const int c = threadIdx.x;
  const int cx = c % T::CW, cy = c / T::CW;
  const int sx = cx + 1, sy = cy + 1;
  const int gi = i0 - H + sx, gj = j0 - H + sy;
  const bool in = c < T::NC && gi >= 0 && gi <= nx && gj >= 0 && gj <= ny;
  const int64_t n = in ? (int64_t)gj * (nx + 1) + gi : 0;
  unsigned fix = 0;
  double bv[3] = {0.0, 0.0, 0.0}, xov[3] = {0.0, 0.0, 0.0};
  float mv[6] = {0, 0, 0, 0, 0, 0};
  StRow R;
  if (in) {
    if (MODE == M_RES) fix = A.L.mask[n];
    for (int d = 0; d < 3; d++) bv[d] = ldv<COH>(A.b + 3 * n + d);
    if (MODE == M_STEP) {
      for (int k = 0; k < 6; k++) mv[k] = A.L.minv[6 * n + k];
      if (A.c1 != 0.0)
        for (int d = 0; d < 3; d++) xov[d] = ldv<COH>(A.xo + 3 * n + d);
    }
    load_st_row(nx, A.L.st, gi, gj, R);
  }
  constexpr int CWc = TS / 2 + 2;
  static_assert(CWc * CWc <= 256, "one coarse node per thread");
  int cI = -1, cJ = -1;
  float cm[6] = {0, 0, 0, 0, 0, 0};
  if (MODE == M_RES && (int)threadIdx.x < CWc * CWc) {
    const int I = i0 / 2 + threadIdx.x % CWc, J = j0 / 2 + threadIdx.x / CWc;
    const int fi0 = min(2 * I, nx), fj0 = min(2 * J, ny);
    if (I <= A.co.nx && J <= A.co.ny && fi0 >= i0 && fi0 < i0 + TS && fj0 >= j0 &&
        fj0 < j0 + TS) {
      cI = I;
      cJ = J;
      if (A.co.minv)
        for (int q = 0; q < 6; q++) cm[q] = A.co.minv[6 * ((int64_t)J * (A.co.nx + 1) + I) + q];
    }
  }
  // stage x (+ P xc, masked) on the tile and its halo
  const int fi0 = max(i0 - H, 0), fj0 = max(j0 - H, 0);
  if (PRO) {
    stage_coarse<T::PW, 256, COH>(fi0, fj0, A.pro.xc, A.pro.cnx, (ny + 1) / 2, sm.xc);
    __syncthreads();
  }
  for (int idx = threadIdx.x; idx < T::NP; idx += 256) {
    const int lx = idx % T::PW, ly = idx / T::PW;
    const int hi = i0 - H + lx, hj = j0 - H + ly;
    double v[3] = {0.0, 0.0, 0.0};
    if (hi >= 0 && hi <= nx && hj >= 0 && hj <= ny) {
      const int64_t m = (int64_t)hj * (nx + 1) + hi;
      for (int d = 0; d < 3; d++) v[d] = ldv<COH>(A.x + 3 * m + d);
      if (PRO) {
        double sp[3];
        prolong_node_lds<T::PW>(hi, hj, nx, ny, fi0, fj0, sm.xc, sp);
        const unsigned msk = A.L.mask[m];
        for (int d = 0; d < 3; d++)
          if (!(msk >> d & 1)) v[d] += sp[d];
      }
    }
    for (int d = 0; d < 3; d++) sm.u[d * T::NP + idx] = v[d];
  }
  __syncthreads();
  if (c < T::NC) {
    if (!in) {
      if (MODE == M_RES) {
        for (int d = 0; d < 3; d++) sm.t[d * T::NC + c] = 0.0;
        sm.m[c] = 0;
      }
    } else {
      const int si = sy * T::PW + sx;
      double Kx[3];
      stencil_apply<T::PW, T::NP>(R, sx, sy, sm.u, Kx);
      if (MODE == M_RES) {
        for (int d = 0; d < 3; d++) sm.t[d * T::NC + c] = bv[d] - Kx[d];
        sm.m[c] = (uint8_t)fix;
      } else {
        double r[3], z[3];
        for (int d = 0; d < 3; d++) r[d] = bv[d] - Kx[d];
        sym3_apply(mv, r, z);
        for (int d = 0; d < 3; d++) {
          const double xc = sm.u[d * T::NP + si];
          const double dn = (A.c1 != 0.0 ? A.c1 * (xc - xov[d]) : 0.0) + A.c2 * z[d];
          stv<COH>(A.xo + 3 * n + d, xc + dn);
          if (PRO) stv<COH>(A.pro.xpro + 3 * n + d, xc);
        }
      }
    }
  }
  if (MODE == M_RES) {
    __syncthreads();
    const Coarse& co = A.co;
    if (cI >= 0) {
      const int I = cI, J = cJ;
      const int64_t nc = (int64_t)J * (co.nx + 1) + I;
      double sr[3] = {0.0, 0.0, 0.0};
      for (int bb = -1; bb <= 1; bb++) {
        const int fj = 2 * J + bb;
        if (fj < 0 || fj > ny) continue;
        for (int aa = -1; aa <= 1; aa++) {
          const int fi = 2 * I + aa;
          if (fi < 0 || fi > nx) continue;
          const double w = w1(fi, I, nx) * w1(fj, J, ny);
          const int ci = (fj - (j0 - H + 1)) * T::CW + (fi - (i0 - H + 1));
          const unsigned msk = sm.m[ci];
          for (int d = 0; d < 3; d++)
            if (!(msk >> d & 1)) sr[d] += w * sm.t[d * T::NC + ci];
        }
      }
      for (int d = 0; d < 3; d++) stv<COH>(co.b + 3 * nc + d, sr[d]);
      if (co.minv) {
        double z[3];
        sym3_apply(cm, sr, z);
        for (int d = 0; d < 3; d++) {
          stv<COH>(co.cur + 3 * nc + d, 0.0 + co.c2 * z[d]);
          stv<COH>(co.prev + 3 * nc + d, 0.0);
        }
      }
    }
  }
}

// One launch per stencil-level phase.
template <int TS, int H, int MODE, bool PRO>
__global__ __launch_bounds__(256) void k_st(TileArgs A) {
  __shared__ StSmem sm;
  if (A.S->done) return;
  st_tile<TS, H, MODE, PRO, false>(A, xcd_band_tile((int)blockIdx.x, (int)gridDim.x), sm);
}

// x += alpha p, r -= alpha q with alpha = p.r / p.q, the partials of r.r,
// and the V-cycle's first level-0 step from x = 0 on the new r: cur = c2 B r,
// prev = 0.  Consumer form (cin): every workgroup sums the K p partials
// [p.q | p.r] itself and applies the breakdown rule (pcg.ts:154), and the
// kernel after this one applies the stopping rule; otherwise alpha comes
// from the state and k_finish applies the stopping rule.
__global__ __launch_bounds__(256) void k_update(int64_t nn, double* x, double* r, const double* p,
                                                const double* q, const float* minv, double c2,
                                                double* cur, double* prev, const double* cin,
                                                int cn, Red red, MgfState* S, MgfRec* rec,
                                                int nrec) {
  __shared__ double red_s[8];
  __shared__ double in_s[2];
  if (S->done) return;
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double xv[3] = {0, 0, 0}, rv[3] = {0, 0, 0}, pv[3] = {0, 0, 0}, qv[3] = {0, 0, 0};
  float mv[6] = {0, 0, 0, 0, 0, 0};
  if (n < nn) {
    for (int d = 0; d < 3; d++) {
      xv[d] = x[3 * n + d];
      rv[d] = r[3 * n + d];
      pv[d] = p[3 * n + d];
      qv[d] = q[3 * n + d];
    }
    for (int k = 0; k < 6; k++) mv[k] = minv[6 * n + k];
  }
  double a;
  if (cin) {
    double pp[2] = {0.0, 0.0};
    if (threadIdx.x < 64) wave_sum_parts<2>(cin, cn, pp);
    if (threadIdx.x == 0) {
      in_s[0] = pp[0];
      in_s[1] = pp[1];
    }
    __syncthreads();
    const double pq = in_s[0];
    if (fabs(pq) < 1e-30) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        S->pq = pq;
        S->breakdown = 1;
        S->done = 1;
        write_record(rec, nrec, S->it, S);
      }
      return;
    }
    a = in_s[1] / pq;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      S->pq = pq;
      S->alpha = a;
    }
  } else {
    a = S->alpha;
  }
  double v[1] = {0.0};
  if (n < nn) {
    double rn[3];
    for (int d = 0; d < 3; d++) {
      x[3 * n + d] = xv[d] + a * pv[d];
      const double ri = rv[d] - a * qv[d];
      r[3 * n + d] = ri;
      rn[d] = ri;
      v[0] += ri * ri;
    }
    double z[3];
    sym3_apply(mv, rn, z);
    for (int d = 0; d < 3; d++) {
      cur[3 * n + d] = 0.0 + c2 * z[d];
      prev[3 * n + d] = 0.0;
    }
  }
  block_sum<1, 256>(v, red_s);
  if (threadIdx.x == 0) red.part[blockIdx.x] = v[0];
}

// The scalar step after a reduction with many producers, in one 1024-thread
// workgroup: a fixed-order sum of the n partials, then
//   kind 0 (after K p): alpha = p.r / p.q or the breakdown rule (pcg.ts:154);
//   kind 1 (after the update): the stopping rule (pcg.ts:174), the iteration
//           count and its progress record;
//   kind 2 (after the last smoothing step): beta = -z.q / p.q.
__global__ __launch_bounds__(1024) void k_finish(int kind, const double* part, int n, MgfState* S,
                                                 MgfRec* rec, int nrec) {
  __shared__ double lds[2 * 16];
  if (S->done) return;
  double v[2] = {0.0, 0.0};
  const int nv = kind == 0 ? 2 : 1;
  for (int i = threadIdx.x; i < n; i += 1024)
    for (int k = 0; k < nv; k++) v[k] += part[(size_t)k * n + i];
  block_sum<2, 1024>(v, lds);
  if (threadIdx.x) return;
  if (kind == 0) {
    S->pq = v[0];
    if (fabs(v[0]) < 1e-30) {
      S->breakdown = 1;
      S->done = 1;
      write_record(rec, nrec, S->it, S);
    } else {
      S->alpha = v[1] / v[0];
    }
  } else if (kind == 1) {
    S->rr = v[0];
    S->it++;
    const double rn = sqrt(v[0]);
    if (rn < S->tol || rn / S->fnorm < S->tol) {
      S->converged = 1;
      S->done = 1;
    }
    write_record(rec, nrec, S->it - 1, S);
  } else {
    S->beta = -v[0] / S->pq;
  }
}

// The V-cycle's first level-0 step from x = 0 without an update (start of a run).
__global__ void k_point(int64_t nn, const float* minv, double c2, const double* b, double* cur,
                        double* prev) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  double rv[3] = {b[3 * n], b[3 * n + 1], b[3 * n + 2]}, z[3];
  sym3_apply(minv + 6 * n, rv, z);
  for (int d = 0; d < 3; d++) {
    cur[3 * n + d] = 0.0 + c2 * z[d];
    prev[3 * n + d] = 0.0;
  }
}

// x = C b on the coarsest level (C dense symmetric nd x nd), one wave per row.
__global__ void k_dense(int nd, const double* C, const double* b, double* x, const MgfState* S) {
  if (S->done) return;
  const int row = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64, lane = threadIdx.x & 63;
  if (row >= nd) return;
  // four independent partial sums per lane keep four loads in flight
  const double* Cr = C + (int64_t)row * nd;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int k = lane;
  for (; k + 192 < nd; k += 256) {
    s0 += Cr[k] * b[k];
    s1 += Cr[k + 64] * b[k + 64];
    s2 += Cr[k + 128] * b[k + 128];
    s3 += Cr[k + 192] * b[k + 192];
  }
  for (; k < nd; k += 64) s0 += Cr[k] * b[k];
  double s = (s0 + s1) + (s2 + s3);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) x[row] = s;
}

__global__ void k_state_init(MgfState* S, double fnorm, double tol) {
  if (threadIdx.x) return;
  *S = MgfState{};
  S->fnorm = fnorm > 0.0 ? fnorm : 1.0;
  S->tol = tol;
}

__global__ void k_fill0(int64_t n, double* a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

int nblk(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }
int ntiles(int nx, int ny, int ts) { return ((nx + 1 + ts - 1) / ts) * ((ny + 1 + ts - 1) / ts); }

// tile shapes: level-0 steps and K p on 15 x 15 nodes (16 x 16 elements: one
// element per thread), level-0 residuals on 13 x 13 (a 2-node halo: again
// 16 x 16 elements); stencil levels 16 x 16 and 14 x 14
constexpr int kTs0 = 15, kTr0 = 13, kTsS = kTsS_, kTrS = kTrS_;

struct Coef {
  double c1, c2;
};

// Host side of the fused iteration.
struct Fused {
  pp_plan* pl = nullptr;
  Mg* mg = nullptr;
  hipStream_t s = nullptr;
  KGrid G;
  int nl = 0;
  struct Lv {
    LvView v;
    int64_t nn = 0;
    double* buf[3] = {nullptr, nullptr, nullptr};  // smoothing iterates
    double* b = nullptr;
    double* out = nullptr;                         // where the cycle left x
    std::vector<Coef> cs;                          // step coefficients of a smoothing pass
  } L[Mg::kMaxLevels];
  MgfState* S = nullptr;
  MgfRec* rec = nullptr;  // coherent pinned progress records
  int nrec = 0;
  double* part = nullptr;  // partials: K p (2 x maxwg), update (maxwg), z.q (maxwg)
  int maxwg = 0;
  // consumer form of the reductions (few workgroups): every workgroup of the
  // next kernel sums the partials itself, so no kernel ends with a last-arrival
  // tail; with many workgroups the last arrival finishes each sum instead
  bool consumer = false;
  bool pending_check = false;  // the next V-cycle kernel applies the stopping rule
  // buffers the level-0 first step writes (cur, prev), chosen so that the
  // cycle ends in buf[0] = z: (1, 0) for an odd step count, (2, 1) for even
  int pc0 = 1, pp0 = 0;

  int setup(pp_plan* p, Mg* m, MgfState* st, double* pt, int mw) {
    // the fused stencil rows of the stencil levels, built once per hierarchy
    for (int l = 1; l + 1 < m->nl; l++) {
      MgLevel& M = m->lv[l];
      if (M.rows.n == (size_t)kRow * M.nn) continue;
      int rc = M.rows.alloc((size_t)kRow * M.nn);
      if (rc) return rc;
      hipLaunchKernelGGL(k_rows, dim3(nblk(M.nn, 256)), dim3(256), 0, p->ctx->stream, M.nx, M.ny,
                         (const float*)M.st.p, M.rows.p);
      PP_HIP(hipGetLastError());
    }
    pl = p;
    mg = m;
    s = p->ctx->stream;
    G = make_kgrid(p);
    nl = m->nl;
    S = st;
    part = pt;
    maxwg = mw;
    pc0 = m->nu % 2 ? 1 : 2;
    pp0 = m->nu % 2 ? 0 : 1;
    for (int l = 0; l < nl; l++) {
      MgLevel& M = m->lv[l];
      Lv& V = L[l];
      V.v.nx = M.nx;
      V.v.ny = M.ny;
      V.v.st = l > 0 ? M.rows.p : nullptr;
      V.v.minv = M.minv.p;
      V.v.mask = l == 0 ? (const uint8_t*)p->cmask.p : (const uint8_t*)M.dead.p;
      V.nn = M.nn;
      if (l == 0) {
        V.buf[0] = p->z.p;
        V.buf[1] = M.x2.p;
        V.buf[2] = M.t.p;
        V.b = p->r.p;
      } else {
        V.buf[0] = M.x.p;
        V.buf[1] = M.x2.p;
        V.buf[2] = M.t.p;
        V.b = M.b.p;
      }
      if (l == nl - 1) continue;
      const int ns = l == 0 ? m->nu : m->nuc;
      V.cs.resize(ns);
      double theta = 1.0, delta = 1.0, rho = 0.0;
      if (!m->jacobi) {
        const double hi = 4.0 / (3.0 * M.omega), lo = hi / m->cheb_ratio;
        theta = 0.5 * (hi + lo);
        delta = 0.5 * (hi - lo);
      }
      const double sigma = theta / delta;
      for (int k = 0; k < ns; k++) {
        if (m->jacobi) {
          V.cs[k] = {0.0, M.omega};
        } else if (k == 0) {
          rho = 1.0 / sigma;
          V.cs[k] = {0.0, 1.0 / theta};
        } else {
          const double rn = 1.0 / (2.0 * sigma - rho);
          V.cs[k] = {rn * rho, 2.0 * rn / delta};
          rho = rn;
        }
      }
    }
    return PP_OK;
  }

  TileArgs args(int l) {
    TileArgs A;
    memset(&A, 0, sizeof(A));
    A.G = G;
    A.L = L[l].v;
    A.S = S;
    A.rec = rec;
    A.nrec = nrec;
    if (pending_check) {
      A.check = 1;
      A.cin = part + 2 * (size_t)maxwg;
      A.cn = nblk(L[0].nn, 256);
      pending_check = false;
    }
    return A;
  }

  // one smoothing step on level l: x -> xo (xo holds the previous iterate)
  int step(int l, const double* x, double* xo, Coef c, const Pro* pro, bool zq) {
    TileArgs A = args(l);
    A.x = x;
    A.xo = xo;
    A.b = L[l].b;
    A.c1 = c.c1;
    A.c2 = c.c2;
    if (pro) A.pro = *pro;
    if (zq) {
      A.q = pl->Ap.p;
      A.red = {part + 3 * (size_t)maxwg};
    }
    const int nx = L[l].v.nx, ny = L[l].v.ny;
    if (l == 0) {
      const int g = ntiles(nx, ny, kTs0);
      if (pro && zq)
        hipLaunchKernelGGL((k_tile<kTs0, 1, 256, M_STEP, true, true>), dim3(g), dim3(256), 0, s, A);
      else if (pro)
        hipLaunchKernelGGL((k_tile<kTs0, 1, 256, M_STEP, true, false>), dim3(g), dim3(256), 0, s, A);
      else if (zq)
        hipLaunchKernelGGL((k_tile<kTs0, 1, 256, M_STEP, false, true>), dim3(g), dim3(256), 0, s, A);
      else
        hipLaunchKernelGGL((k_tile<kTs0, 1, 256, M_STEP, false, false>), dim3(g), dim3(256), 0, s, A);
      PP_HIP(hipGetLastError());
      if (zq && !consumer) return finish(2, A.red.part, g);
    } else {
      const int g = ntiles(nx, ny, kTsS);
      if (pro)
        hipLaunchKernelGGL((k_st<kTsS, 1, M_STEP, true>), dim3(g), dim3(256), 0, s, A);
      else
        hipLaunchKernelGGL((k_st<kTsS, 1, M_STEP, false>), dim3(g), dim3(256), 0, s, A);
    }
    PP_HIP(hipGetLastError());
    return PP_OK;
  }

  // residual of level l at x, restricted to level l + 1, with its first step
  // when l + 1 is not the coarsest (into buf[1], buf[0] zeroed)
  int res(int l, const double* x) {
    TileArgs A = args(l);
    A.x = x;
    A.b = L[l].b;
    Lv& C = L[l + 1];
    A.co.nx = C.v.nx;
    A.co.ny = C.v.ny;
    A.co.b = C.b;
    if (l + 1 < nl - 1) {
      A.co.minv = C.v.minv;
      A.co.c2 = C.cs[0].c2;
      A.co.cur = C.buf[1];
      A.co.prev = C.buf[0];
    }
    const int nx = L[l].v.nx, ny = L[l].v.ny;
    if (l == 0) {
      hipLaunchKernelGGL((k_tile<kTr0, 2, 256, M_RES, false, false>),
                         dim3(ntiles(nx, ny, kTr0)), dim3(256), 0, s, A);
    } else {
      hipLaunchKernelGGL((k_st<kTrS, 2, M_RES, false>), dim3(ntiles(nx, ny, kTrS)), dim3(256), 0,
                         s, A);
    }
    PP_HIP(hipGetLastError());
    return PP_OK;
  }

  // V-cycle on level l whose first step is already done (level 0: cur =
  // buf[pc0], prev = buf[pp0]; coarser: buf[1], buf[0]); the result is left
  // in L[l].out (level 0: buf[0] = z).  zq: the last level-0
  // step also sums z.q for beta.
  int cycle(int l, bool zq) {
    Lv& V = L[l];
    if (l == nl - 1) {
      hipLaunchKernelGGL(k_dense, dim3(nblk(mg->nd_coarse, 4)), dim3(256), 0, s, mg->nd_coarse,
                         (const double*)mg->cinv.p, (const double*)V.b, V.buf[0],
                         (const MgfState*)S);
      PP_HIP(hipGetLastError());
      V.out = V.buf[0];
      return PP_OK;
    }
    const int ns = (int)V.cs.size();
    int cur = l == 0 ? pc0 : 1, prev = l == 0 ? pp0 : 0, rc;
    for (int k = 1; k < ns; k++) {
      if ((rc = step(l, V.buf[cur], V.buf[prev], V.cs[k], nullptr, false))) return rc;
      std::swap(cur, prev);
    }
    if ((rc = res(l, V.buf[cur]))) return rc;
    if ((rc = cycle(l + 1, false))) return rc;
    // prolongation with the first post-smoothing step: out and the prolonged
    // iterate go to the two buffers other than cur (which other tiles still
    // read as their halo); pick them so that the pass ends in buf[0] (level
    // 0: z)
    const int o1 = (cur + 1) % 3, o2 = (cur + 2) % 3;
    int out, pro;
    if ((ns - 1) % 2 == 0) {  // the pass ends where the first post step writes
      out = o1 == 0 ? o1 : (o2 == 0 ? o2 : o1);
      pro = out == o1 ? o2 : o1;
    } else {                  // ... or where the prolonged iterate goes
      pro = o1 == 0 ? o1 : (o2 == 0 ? o2 : o1);
      out = pro == o1 ? o2 : o1;
    }
    Pro P{L[l + 1].out, L[l + 1].v.nx, V.buf[pro]};
    if ((rc = step(l, V.buf[cur], V.buf[out], V.cs[0], &P, zq && ns == 1))) return rc;
    cur = out;
    prev = pro;
    for (int k = 1; k < ns; k++) {
      if ((rc = step(l, V.buf[cur], V.buf[prev], V.cs[k], nullptr, zq && k == ns - 1))) return rc;
      std::swap(cur, prev);
    }
    V.out = V.buf[cur];
    return PP_OK;
  }

  // K p with p = z + beta p_old, partials for alpha; first: beta = 0
  int kp(const double* pold, double* pnew, bool first) {
    TileArgs A = args(0);
    A.x = pl->z.p;
    A.y = pold;
    A.xo = pnew;
    A.b = pl->r.p;
    A.q = pl->Ap.p;
    A.red = {part};
    if (consumer && !first) {
      A.cin = part + 3 * (size_t)maxwg;
      A.cn = ntiles(L[0].v.nx, L[0].v.ny, kTs0);
    }
    const int g = ntiles(L[0].v.nx, L[0].v.ny, kTs0);
    hipLaunchKernelGGL((k_tile<kTs0, 1, 256, M_KP, false, false>), dim3(g), dim3(256), 0, s, A);
    PP_HIP(hipGetLastError());
    return consumer ? PP_OK : finish(0, part, g);
  }

  // the scalar step of a reduction with many producers (k_finish)
  int finish(int kind, const double* pt, int n) {
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(1024), 0, s, kind, pt, n, S, rec, nrec);
    PP_HIP(hipGetLastError());
    return PP_OK;
  }

  int update(const double* p) {
    const Lv& V = L[0];
    const double* cin = consumer ? part : nullptr;
    const int cn = ntiles(V.v.nx, V.v.ny, kTs0);
    Red red{part + 2 * (size_t)maxwg};
    hipLaunchKernelGGL(k_update, dim3(nblk(V.nn, 256)), dim3(256), 0, s, V.nn, pl->x.p, pl->r.p,
                       p, (const double*)pl->Ap.p, V.v.minv, V.cs[0].c2, V.buf[pc0], V.buf[pp0],
                       cin, cn, red, S, rec, nrec);
    PP_HIP(hipGetLastError());
    pending_check = consumer;
    return consumer ? PP_OK : finish(1, red.part, nblk(V.nn, 256));
  }

  // z = V r from the current r (start of a run)
  int first_cycle() {
    const Lv& V = L[0];
    hipLaunchKernelGGL(k_point, dim3(nblk(V.nn, 256)), dim3(256), 0, s, V.nn, V.v.minv,
                       V.cs[0].c2, (const double*)V.b, V.buf[pc0], V.buf[pp0]);
    PP_HIP(hipGetLastError());
    return cycle(0, false);
  }
};

// largest producer grid for which every consumer workgroup sums the partials itself
constexpr int kConsumerMax = 512;

// workgroups of the largest fused reduction
int max_workgroups(const pp_plan* pl) {
  return std::max(ntiles(pl->nx, pl->ny, kTs0), nblk(pl->n_nodes, 256));
}

}  // namespace

// ------------------------------------------------ V-cycle outside PCG

struct MgfCycle {
  Fused F;
  DevBuf<MgfState> S;  // never stops: only the V-cycle kernels read it
  DevBuf<double> part;
};

int mgf_cycle_create(pp_plan* pl, Mg* mg, MgfCycle** out) {
  *out = nullptr;
  MgfCycle* c = new MgfCycle();
  int rc;
  if ((rc = c->S.alloc(1)) || (rc = c->part.alloc(4)) || (rc = c->F.setup(pl, mg, c->S.p, c->part.p, 1))) {
    delete c;
    return rc;
  }
  hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, pl->ctx->stream, c->S.p, 1.0, 0.0);
  if (hipGetLastError() != hipSuccess) {
    delete c;
    return fail(PP_ERR_HIP, "multigrid: state initialisation failed");
  }
  *out = c;
  return PP_OK;
}

// x = V b: the pointwise first step from x = 0, then the fused cycle; level 0
// reads b and leaves its result in x (the cycle's buffers are chosen so that
// it ends in buf[0]), using the hierarchy's own level-0 x2 and t as the
// other two iterates.
int mgf_cycle_apply(MgfCycle* c, const double* b, double* x) {
  Fused& F = c->F;
  Fused::Lv& V = F.L[0];
  if (b == x || x == V.buf[1] || x == V.buf[2] || b == V.buf[1] || b == V.buf[2])
    return fail(PP_ERR_INVALID, "multigrid: V-cycle vectors alias");
  V.b = const_cast<double*>(b);
  V.buf[0] = x;
  int rc = F.first_cycle();
  if (rc) return rc;
  if (V.out != x) return fail(PP_ERR_INVALID, "multigrid: V-cycle did not end in its output");
  return PP_OK;
}

void mgf_cycle_destroy(MgfCycle* c) { delete c; }

constexpr int kRecRing = 8;

// FCG(1) with the fused V-cycle from x (zero on a cold start) with r = F - K x
// in pl->r; stops by the rule of pcg.ts:174 or at maxit iterations.  Every
// iteration that starts before the stop leaves one progress record in
// coherent pinned host memory; the host spins on the record of the oldest
// queued iteration and keeps up to kAhead iterations queued.  From the last
// two recorded residual norms it predicts the iteration at which the rule
// will hold and queues nothing beyond it until that is confirmed (queued
// iterations after the stop return at once but still cost a dispatch each).
static int mgf_run(pp_plan* pl, Fused& F, double fnorm, double tol, int maxit, MgfState* hs) {
  constexpr int kAhead = 2;
  hipStream_t s = pl->ctx->stream;
  if (!pl->mgf_pinned)
    PP_HIP(hipHostMalloc(&pl->mgf_pinned, kRecRing * sizeof(MgfRec), hipHostMallocCoherent));
  MgfRec* rec = (MgfRec*)pl->mgf_pinned;
  memset(rec, 0, kRecRing * sizeof(MgfRec));  // the stream is idle: earlier runs were waited for
  F.rec = rec;
  F.nrec = kRecRing;
  hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, s, F.S, fnorm, tol);
  // p_old = 0 (beta = 0 on the first iteration makes p = z)
  hipLaunchKernelGGL(k_fill0, dim3(nblk(pl->n_dof, 256)), dim3(256), 0, s, pl->n_dof, pl->p0.p);
  hipLaunchKernelGGL(k_fill0, dim3(nblk(pl->n_dof, 256)), dim3(256), 0, s, pl->n_dof, pl->p1.p);
  PP_HIP(hipGetLastError());
  int rc;
  if ((rc = F.first_cycle())) return rc;
  if (F.L[0].out != pl->z.p) return fail(PP_ERR_INVALID, "multigrid: V-cycle did not end in z");
  const double fn = fnorm > 0.0 ? fnorm : 1.0;
  auto converges = [&](double rn) { return rn < tol || rn / fn < tol; };
  int queued = 0, checked = 0;
  double rn_last = -1.0, rho = -1.0;  // |r| of the last checked iteration, and its ratio to the one before
  rc = PP_OK;
  for (;;) {
    while (queued < maxit && queued - checked < kAhead) {
      // predicted |r| after iteration queued - 1 (slack of 1.5 for the estimate)
      if (queued > checked && rn_last > 0.0 && rho > 0.0 && rho < 1.0 &&
          converges(rn_last * pow(rho, queued - checked) / 1.5))
        break;
      double* pold = (queued & 1) ? pl->p0.p : pl->p1.p;
      double* pnew = (queued & 1) ? pl->p1.p : pl->p0.p;
      if ((rc = F.kp(pold, pnew, queued == 0)) || (rc = F.update(pnew)) ||
          (rc = F.cycle(0, true)))
        break;
      queued++;
    }
    if (rc || checked == queued) break;
    // wait for the record of iteration `checked` (or for the stream to drain:
    // then no record is coming and the device state is final)
    const MgfRec* r = rec + checked % kRecRing;
    bool got = false;
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned spin = 0;; spin++) {
      if (__atomic_load_n(&r->seq, __ATOMIC_ACQUIRE) == checked + 1) {
        got = true;
        break;
      }
      if ((spin & 255) == 255) {
        if (hipStreamQuery(s) == hipSuccess) {
          got = __atomic_load_n(&r->seq, __ATOMIC_ACQUIRE) == checked + 1;
          break;
        }
        if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(60)) {
          rc = fail(PP_ERR_TIMEOUT, "multigrid: no progress record within 60 s");
          break;
        }
      }
    }
    if (rc || !got) break;
    const double rn = sqrt(r->rr);
    const int done = r->done;
    if (rn_last > 0.0) rho = rn / rn_last;
    rn_last = rn;
    checked++;
    if (done) break;
  }
  hipError_t e = hipStreamSynchronize(s);
  if (rc) return rc;
  if (e != hipSuccess) return fail(PP_ERR_HIP, std::string("multigrid: ") + hipGetErrorString(e));
  PP_HIP(hipMemcpy(hs, F.S, sizeof(MgfState), hipMemcpyDeviceToHost));
  return PP_OK;
}

// The fused multigrid PCG with residual replacement (the restart rule of the
// launch-per-operation form in pp_fcg.hip).
int mgf_solve(pp_plan* pl, Mg* mg, const pp_solve_opts* o, pp_solve_stats* st) {
  constexpr int kMaxRestarts = 16;
  hipStream_t s = pl->ctx->stream;
  const int64_t n = pl->n_dof;
  const int mw = max_workgroups(pl);
  DevBuf<MgfState> S;
  DevBuf<double> part;
  int rc;
  if ((rc = S.alloc(1)) || (rc = part.alloc(4 * (size_t)mw))) return rc;
  Fused F;
  if ((rc = F.setup(pl, mg, S.p, part.p, mw))) return rc;
  F.consumer = mw <= kConsumerMax;

  PP_HIP(hipEventRecord(pl->ctx->ev0, s));
  hipLaunchKernelGGL(k_fill0, dim3(nblk(n, 256)), dim3(256), 0, s, n, pl->x.p);
  PP_HIP(hipMemcpyAsync(pl->r.p, pl->F.p, n * 8, hipMemcpyDeviceToDevice, s));
  double res = 0, fnorm = 0;
  if ((rc = launch_norms(pl, pl->F.p, pl->F.p, &res, &fnorm))) return rc;
  MgfState hs;
  memset(&hs, 0, sizeof(hs));
  int total = 0, sweeps = 0;
  double rel = 0.0;
  const bool check = !(o->flags & PP_FLAG_NO_TRUE_RESIDUAL);
  if (fnorm < 1e-30) {
    hs.converged = 1;
  } else {
    for (int k = 0; k <= kMaxRestarts; k++) {
      const double tol = k <= 1 ? o->tolerance : 0.5 * o->tolerance;
      if ((rc = mgf_run(pl, F, fnorm, tol, o->max_iterations - total, &hs))) return rc;
      if (sweeps < PP_MAX_SWEEPS) st->sweep_iterations[sweeps] = hs.it;
      total += hs.it;
      if (!check) {
        sweeps++;
        break;
      }
      if ((rc = struct_apply_k(pl, pl->x.p, pl->tmp.p))) return rc;
      if ((rc = launch_norms(pl, pl->F.p, pl->tmp.p, &res, &fnorm))) return rc;
      rel = res / fnorm;
      if (sweeps < PP_MAX_SWEEPS) st->sweep_residual[sweeps] = rel;
      sweeps++;
      if (!hs.converged || rel <= o->tolerance || total >= o->max_iterations) break;
      if ((rc = launch_residual(pl, pl->F.p, pl->tmp.p, pl->r.p))) return rc;
    }
  }
  PP_HIP(hipEventRecord(pl->ctx->ev1, s));
  PP_HIP(hipEventSynchronize(pl->ctx->ev1));
  float ms = 0;
  PP_HIP(hipEventElapsedTime(&ms, pl->ctx->ev0, pl->ctx->ev1));
  st->iterations = total;
  st->converged = hs.converged && (!check || rel <= o->tolerance);
  st->breakdown = hs.breakdown;
  st->sweeps = sweeps;
  st->final_residual = sqrt(hs.rr);
  st->true_rel_residual = check ? rel : 0.0;
  st->device_ms = ms;
  return PP_OK;
}

}  // namespace pp
