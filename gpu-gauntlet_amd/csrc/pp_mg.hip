// Galerkin geometric multigrid V-cycle on the structured node grid: the
// smoother S around the learned network when the configuration asks for it.
//
// Hierarchy (set up on the device):
//   level 0      the plan's operator: matrix-free FP64 K with the reference's
//                identity rows on constrained DOFs, its FP64 3x3 block-Jacobi
//                inverse, and the constraint bits as excluded DOFs;
//   level l + 1  elements (I, J) of 2 x 2 level-l elements; coarse node I is
//                level-l node min(2I, n) for n level-l elements along that
//                side, so an odd count ends in a coarse element one level-l
//                element wide (n -> (n + 1) / 2).  P_l interpolates each DOF
//                component bilinearly in index space over the coarse
//                element's level-l nodes (weights 1, 1/2, 1/4) and is zero
//                on the excluded DOFs of level l, so that
//                A_{l+1} = P_l^T A_l P_l.  Because a level-l node on the
//                closure of coarse element E interpolates from E's corners
//                only, A_{l+1} is the assembly of the coarse element matrices
//                K_E = sum over its (up to 4) children c of P_c^T (D K_c D) P_c
//                (P_c the 12 x 12 corner weights, D the exclusion mask).
//                A DOF whose assembled diagonal is exactly zero (its basis
//                function vanishes) is excluded on that level: identity row,
//                masked in P.
// Levels >= 1 keep A as an FP32 9-point block stencil: by symmetry only the
// centre and the four forward neighbour blocks (180 B per node, as 45 planes
// of one float per node so that a wave's loads coalesce).  Every level but
// the coarsest smooths with an FP32 copy of its block-Jacobi inverse; the
// coarsest (at most kDenseMax DOFs) is solved by its dense FP64 inverse,
// formed on the host by Cholesky.
//
// V-cycle on level l for right-hand side b (x starts at 0):
//   x = w B b;  nu - 1 times x += w B (b - A x);   (pre-smoothing, nu sweeps)
//   b_{l+1} = P^T (b - A x);  x_{l+1} = V_{l+1} b_{l+1};  x += P x_{l+1};
//   nu times x += w B (b - A x)                   (post-smoothing)
// With the same sweeps before and after, R = P^T and an exact coarsest solve
// the cycle is a symmetric operator, positive definite for w < 2 / lambda_max.

#include <math.h>

#include <algorithm>
#include <vector>

#include "pp_cnn.h"
#include "pp_ke.h"

namespace pp {

namespace {

constexpr int kDenseMax = 2200;   // DOFs of the coarsest level (dense inverse)
constexpr int kDenseStop = 1200;  // stop coarsening at or below this many DOFs
constexpr int kMgDotBlocks = 128;

int mg_blocks(int64_t n, int bs = 256) { return (int)((n + bs - 1) / bs); }

// Corner k of an element at (ci, cj) offsets: SW, SE, NE, NW.
__host__ __device__ __forceinline__ int corner_dx(int k) { return k == 1 || k == 2; }
__host__ __device__ __forceinline__ int corner_dy(int k) { return k >= 2; }

// Coarse element matrices of level l + 1 (elements ncx x ncy) from the
// level-l elements.  Level 0 (kf == nullptr): the Q4 Mindlin element of the
// tensor grid, from its corner coordinates and thickness, with the
// constraint bits of level 0 as the exclusion mask; level >= 1: the stored
// element matrices kf [fine element][144] with the dead bits of level l.
__global__ void k_mg_coarse_elems(int ncx, int ncy, int fnx, int fny, const double* kf, KGrid G,
                                  double E, double nu, const uint8_t* mask, double* kc) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)ncx * ncy) return;
  const int I = (int)(e % ncx), J = (int)(e / ncx);
  // children per side: 2, or 1 for the last coarse element of an odd count
  const int sx = min(2, fnx - 2 * I), sy = min(2, fny - 2 * J);
  double Kc[144];
  for (int k = 0; k < 144; k++) Kc[k] = 0.0;
  double Kf[144];
  for (int c = 0; c < 4; c++) {
    if (corner_dx(c) >= sx || corner_dy(c) >= sy) continue;
    const int fi = 2 * I + corner_dx(c), fj = 2 * J + corner_dy(c);  // child element (fi, fj)
    if (kf) {
      const double* src = kf + ((int64_t)fj * fnx + fi) * 144;
      for (int k = 0; k < 144; k++) Kf[k] = src[k];
    } else {
      double xy[8];
      for (int k = 0; k < 4; k++) {
        xy[2 * k] = G.xc[fi + corner_dx(k)];
        xy[2 * k + 1] = G.yc[fj + corner_dy(k)];
      }
      const double t = G.elem_t ? (double)G.elem_t[(int64_t)fj * G.nx + fi] : G.t0;
      ke_q4(xy, E, nu, t, Kf);
    }
    // exclusion mask on the child's 12 DOFs
    unsigned m[4];
    for (int k = 0; k < 4; k++)
      m[k] = mask[(int64_t)(fj + corner_dy(k)) * (fnx + 1) + fi + corner_dx(k)];
    // corner weights: child corner k sits at (u, v) in {0, 1/2, 1} of E (in
    // {0, 1} along a side with one child)
    double w[4][4];
    for (int k = 0; k < 4; k++) {
      const double u = (double)(corner_dx(c) + corner_dx(k)) / sx;
      const double v = (double)(corner_dy(c) + corner_dy(k)) / sy;
      w[k][0] = (1 - u) * (1 - v);
      w[k][1] = u * (1 - v);
      w[k][2] = u * v;
      w[k][3] = (1 - u) * v;
    }
    for (int k = 0; k < 4; k++)
      for (int d = 0; d < 3; d++) {
        if (m[k] >> d & 1) continue;
        for (int l = 0; l < 4; l++)
          for (int f = 0; f < 3; f++) {
            if (m[l] >> f & 1) continue;
            const double a = Kf[(3 * k + d) * 12 + 3 * l + f];
            if (a == 0.0) continue;
            for (int K = 0; K < 4; K++) {
              const double wa = w[k][K] * a;
              if (wa == 0.0) continue;
              for (int L = 0; L < 4; L++) Kc[(3 * K + d) * 12 + 3 * L + f] += wa * w[l][L];
            }
          }
      }
  }
  double* dst = kc + e * 144;
  for (int k = 0; k < 144; k++) dst[k] = Kc[k];
}

// Stencil, dead bits and block-Jacobi inverse of a level >= 1 from its
// element matrices: node (i, j) sums its elements SW, SE, NE, NW of it in
// that fixed order.
__global__ void k_mg_stencil(int nx, int ny, const double* ke, float* st, uint8_t* dead,
                             float* minv) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)(nx + 1) * (ny + 1)) return;
  const int i = (int)(n % (nx + 1)), j = (int)(n / (nx + 1));
  double S[81];
  for (int k = 0; k < 81; k++) S[k] = 0.0;
  // element (i - 1 + ex, j - 1 + ey); the node is its corner c
  const int ex[4] = {0, 1, 1, 0}, ey[4] = {0, 0, 1, 1};
  for (int q = 0; q < 4; q++) {
    const int ei = i - 1 + ex[q], ej = j - 1 + ey[q];
    if (ei < 0 || ej < 0 || ei >= nx || ej >= ny) continue;
    const double* K = ke + ((int64_t)ej * nx + ei) * 144;
    int c = 0;
    for (int k = 0; k < 4; k++)
      if (ei + corner_dx(k) == i && ej + corner_dy(k) == j) c = k;
    for (int k = 0; k < 4; k++) {
      const int nb = (ej + corner_dy(k) - j + 1) * 3 + (ei + corner_dx(k) - i + 1);
      for (int d = 0; d < 3; d++)
        for (int f = 0; f < 3; f++) S[nb * 9 + 3 * d + f] += K[(3 * c + d) * 12 + 3 * k + f];
    }
  }
  unsigned bits = 0;
  for (int d = 0; d < 3; d++)
    if (S[4 * 9 + 4 * d] == 0.0) {
      bits |= 1u << d;
      S[4 * 9 + 4 * d] = 1.0;
    }
  dead[n] = (uint8_t)bits;
  // A is symmetric: keep the centre block and the forward neighbours E, NW,
  // N, NE (S blocks 4..8); a backward block is the transpose of that
  // neighbour's forward block
  const int64_t nn = (int64_t)(nx + 1) * (ny + 1);
  for (int k = 0; k < 45; k++) st[k * nn + n] = (float)S[36 + k];
  double A[9], m6[6];
  for (int k = 0; k < 9; k++) A[k] = S[4 * 9 + k];
  block_inverse_sym(A, bits, m6);
  for (int k = 0; k < 6; k++) minv[6 * n + k] = (float)m6[k];
}

// (A x)_n on a stencil level.  Stored blocks (planes of one float per
// node): 0 centre, 1 E (0, +1), 2 NW (+1, -1), 3 N (+1, 0), 4 NE (+1, +1)
// as (dj, di).  A forward neighbour outside the grid has exactly zero entries
// (no element contributes to them), so its index is clamped into the grid
// instead of branching; a backward block A(n, m) = A(m, n)^T is read from
// neighbour m's forward block and masked to zero outside the grid.  Every
// load is unconditional, so a thread's loads are all in flight together.
__device__ __forceinline__ void stencil_row(int nx, int ny, const float* st, const double* x,
                                            int64_t n, double (&y)[3]) {
  const int i = (int)(n % (nx + 1)), j = (int)(n / (nx + 1));
  const int64_t nn = (int64_t)(nx + 1) * (ny + 1);
  const int fdj[5] = {0, 0, 1, 1, 1}, fdi[5] = {0, 1, -1, 0, 1};
  y[0] = y[1] = y[2] = 0.0;
#pragma unroll
  for (int sb = 0; sb < 5; sb++) {
    const int64_t m = (int64_t)min(j + fdj[sb], ny) * (nx + 1) + min(max(i + fdi[sb], 0), nx);
    const double x0 = x[3 * m], x1 = x[3 * m + 1], x2 = x[3 * m + 2];
    const float* s = st + (int64_t)sb * 9 * nn + n;
    y[0] += (double)s[0] * x0 + (double)s[nn] * x1 + (double)s[2 * nn] * x2;
    y[1] += (double)s[3 * nn] * x0 + (double)s[4 * nn] * x1 + (double)s[5 * nn] * x2;
    y[2] += (double)s[6 * nn] * x0 + (double)s[7 * nn] * x1 + (double)s[8 * nn] * x2;
  }
#pragma unroll
  for (int sb = 1; sb < 5; sb++) {
    const int jj = j - fdj[sb], ii = i - fdi[sb];
    const double v = (jj >= 0 && ii >= 0 && ii <= nx) ? 1.0 : 0.0;
    const int64_t m = (int64_t)max(jj, 0) * (nx + 1) + min(max(ii, 0), nx);
    const double x0 = v * x[3 * m], x1 = v * x[3 * m + 1], x2 = v * x[3 * m + 2];
    const float* s = st + (int64_t)sb * 9 * nn + m;  // A(m, n) = A(n, m)^T
    y[0] += (double)s[0] * x0 + (double)s[3 * nn] * x1 + (double)s[6 * nn] * x2;
    y[1] += (double)s[nn] * x0 + (double)s[4 * nn] * x1 + (double)s[7 * nn] * x2;
    y[2] += (double)s[2 * nn] * x0 + (double)s[5 * nn] * x1 + (double)s[8 * nn] * x2;
  }
}

// y = A x on a stencil level
__global__ void k_mg_apply(int nx, int ny, const float* st, const double* x, double* y) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)(nx + 1) * (ny + 1)) return;
  double v[3];
  stencil_row(nx, ny, st, x, n, v);
  for (int d = 0; d < 3; d++) y[3 * n + d] = v[d];
}

// Fused sweep on a stencil level: MODE 1: xo = x + c1 (x - xo) + c2 B (b - A x)
// (damped block Jacobi when c1 = 0, xo then not read; a Chebyshev step
// otherwise, xo holding the previous iterate; xo must not alias x); MODE 2:
// xo = b - A x.  Launched with 256 threads: the bound lets MODE 1 keep its
// stencil loads in flight without spilling (under the 1,024-thread default it
// needs scratch).
template <int MODE>
__global__ __launch_bounds__(256) void k_mg_sweep(int nx, int ny, const float* st,
                                                  const float* minv, double c1, double c2,
                                                  const double* x, const double* b, double* xo) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)(nx + 1) * (ny + 1)) return;
  double v[3], r[3];
  stencil_row(nx, ny, st, x, n, v);
  for (int d = 0; d < 3; d++) r[d] = b[3 * n + d] - v[d];
  if (MODE == 2) {
    for (int d = 0; d < 3; d++) xo[3 * n + d] = r[d];
    return;
  }
  double z[3];
  sym3_apply(minv + 6 * n, r, z);
  for (int d = 0; d < 3; d++) {
    const double xc = x[3 * n + d];
    const double dn = (c1 != 0.0 ? c1 * (xc - xo[3 * n + d]) : 0.0) + c2 * z[d];
    xo[3 * n + d] = xc + dn;
  }
}

// mode 0: x = w B b (and x0 = 0 when x0 is given: the iterate before it, for
// the Chebyshev step that follows);  mode 1: x += w B (b - t);  mode 2: t = b - t
__global__ void k_mg_point(int64_t nn, int mode, const float* minv, double omega, const double* b,
                           double* t, double* x, double* x0) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  double r[3], z[3];
  for (int d = 0; d < 3; d++) r[d] = mode == 0 ? b[3 * n + d] : b[3 * n + d] - t[3 * n + d];
  if (mode == 2) {
    for (int d = 0; d < 3; d++) t[3 * n + d] = r[d];
    return;
  }
  sym3_apply(minv + 6 * n, r, z);
  for (int d = 0; d < 3; d++) {
    const double v = (mode == 0 ? 0.0 : x[3 * n + d]) + omega * z[d];
    x[3 * n + d] = v;
    if (x0) x0[3 * n + d] = 0.0;
  }
}

// Interpolation weight of level-l node index f from coarse node index c along
// a side of n level-l elements: coarse node c sits on level-l node
// min(2c, n); an odd f < n between two coarse nodes takes half of each.
__device__ __forceinline__ double w1(int f, int c, int n) {
  if (min(2 * c, n) == f) return 1.0;
  const int d = f - 2 * c;
  return ((f & 1) && f < n && (d == 1 || d == -1)) ? 0.5 : 0.0;
}

// b_c = P^T r on coarse node (I, J): the level-l nodes (2I + a, 2J + b),
// a, b in {-1, 0, 1}, excluded DOFs masked.
__global__ void k_mg_restrict(int fnx, int fny, const uint8_t* fmask, const double* r, int cnx,
                              int cny, double* bc) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)(cnx + 1) * (cny + 1)) return;
  const int I = (int)(n % (cnx + 1)), J = (int)(n / (cnx + 1));
  double s[3] = {0.0, 0.0, 0.0};
  for (int b = -1; b <= 1; b++) {
    const int fj = 2 * J + b;
    if (fj < 0 || fj > fny) continue;
    for (int a = -1; a <= 1; a++) {
      const int fi = 2 * I + a;
      if (fi < 0 || fi > fnx) continue;
      const double w = w1(fi, I, fnx) * w1(fj, J, fny);
      const int64_t m = (int64_t)fj * (fnx + 1) + fi;
      const unsigned msk = fmask[m];
      for (int d = 0; d < 3; d++)
        if (!(msk >> d & 1)) s[d] += w * r[3 * m + d];
    }
  }
  for (int d = 0; d < 3; d++) bc[3 * n + d] = s[d];
}

// x += P x_c on level-l node (fi, fj), excluded DOFs untouched.
__global__ void k_mg_prolong(int fnx, int fny, const uint8_t* fmask, const double* xc, int cnx,
                             double* x) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)(fnx + 1) * (fny + 1)) return;
  const int fi = (int)(n % (fnx + 1)), fj = (int)(n / (fnx + 1));
  const int i0 = fi / 2, i1 = (fi + 1) / 2, j0 = fj / 2, j1 = (fj + 1) / 2;
  double s[3] = {0.0, 0.0, 0.0};
  const int ci[2] = {i0, i1}, cj[2] = {j0, j1};
  const int ni = i0 == i1 ? 1 : 2, nj = j0 == j1 ? 1 : 2;
  for (int b = 0; b < nj; b++)
    for (int a = 0; a < ni; a++) {
      const double w = w1(fi, ci[a], fnx) * w1(fj, cj[b], fny);
      const int64_t m = (int64_t)cj[b] * (cnx + 1) + ci[a];
      for (int d = 0; d < 3; d++) s[d] += w * xc[3 * m + d];
    }
  const unsigned msk = fmask[n];
  for (int d = 0; d < 3; d++)
    if (!(msk >> d & 1)) x[3 * n + d] += s[d];
}

// x = C b, C dense symmetric [nd][nd]: one wave per row.
__global__ void k_mg_dense(int nd, const double* C, const double* b, double* x) {
  const int row = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64, lane = threadIdx.x & 63;
  if (row >= nd) return;
  double s = 0.0;
  for (int k = lane; k < nd; k += 64) s += C[(int64_t)row * nd + k] * b[k];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) x[row] = s;
}

// partials of a.a and b.b (fixed order)
__global__ void k_mg_dot2(int64_t n, const double* a, const double* b, double* part) {
  __shared__ double sa[256], sb[256];
  double x = 0.0, y = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    x += a[i] * a[i];
    y += b[i] * b[i];
  }
  sa[threadIdx.x] = x;
  sb[threadIdx.x] = y;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sa[threadIdx.x] += sa[threadIdx.x + s];
      sb[threadIdx.x] += sb[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sa[0];
    part[2 * blockIdx.x + 1] = sb[0];
  }
}

__global__ void k_mg_f32(int64_t n, const double* a, float* b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

__global__ void k_mg_scale(int64_t n, double* v, double f) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] *= f;
}

// In-place inverse of a symmetric positive definite matrix (row-major n x n)
// by Cholesky A = L L^T, then A^-1 = L^-T L^-1.  Returns false if a pivot is
// not positive.
bool spd_inverse(std::vector<double>& A, int n) {
  std::vector<double> L(A);
  for (int j = 0; j < n; j++) {
    double d = L[(size_t)j * n + j];
    for (int k = 0; k < j; k++) d -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
    if (!(d > 0.0)) return false;
    d = sqrt(d);
    L[(size_t)j * n + j] = d;
    for (int i = j + 1; i < n; i++) {
      double s = L[(size_t)i * n + j];
      for (int k = 0; k < j; k++) s -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
      L[(size_t)i * n + j] = s / d;
    }
  }
  // Linv (lower triangular), column by column
  std::vector<double> Li((size_t)n * n, 0.0);
  for (int c = 0; c < n; c++) {
    Li[(size_t)c * n + c] = 1.0 / L[(size_t)c * n + c];
    for (int i = c + 1; i < n; i++) {
      double s = 0.0;
      for (int k = c; k < i; k++) s += L[(size_t)i * n + k] * Li[(size_t)k * n + c];
      Li[(size_t)i * n + c] = -s / L[(size_t)i * n + i];
    }
  }
  // A^-1 = Li^T Li
  for (int i = 0; i < n; i++)
    for (int j = 0; j <= i; j++) {
      double s = 0.0;
      for (int k = i; k < n; k++) s += Li[(size_t)k * n + i] * Li[(size_t)k * n + j];
      A[(size_t)i * n + j] = A[(size_t)j * n + i] = s;
    }
  return true;
}

}  // namespace

int Mg::setup(pp_plan* plan, double omega0, int sweeps, int coarse_sweeps, bool damped_jacobi) {
  pl = plan;
  stream = plan->ctx->stream;
  jacobi = damped_jacobi;
  // measured at 1024^2: 6 Chebyshev steps (105 flexible-PCG iterations) beat 4 (164) and 8
  // (99, each dearer); damped Jacobi is best at 3
  nu = sweeps > 0 ? sweeps : (jacobi ? 3 : 6);
  nuc = coarse_sweeps > 0 ? coarse_sweeps : 1;
  if (nu > 8 || nuc > 8) return fail(PP_ERR_INVALID, "mg_sweeps must be at most 8");
  // hierarchy shape
  std::vector<std::pair<int, int>> dims{{pl->nx, pl->ny}};
  while (true) {
    const int nx = dims.back().first, ny = dims.back().second;
    const int64_t nd = 3LL * (nx + 1) * (ny + 1);
    if (dims.size() > 1 && nd <= kDenseStop) break;
    if (nx < 2 || ny < 2 || (int)dims.size() == kMaxLevels) break;
    dims.push_back({(nx + 1) / 2, (ny + 1) / 2});
  }
  nl = (int)dims.size();
  const int64_t ndc = 3LL * (dims.back().first + 1) * (dims.back().second + 1);
  if (nl < 2 || ndc > kDenseMax)
    return fail(PP_ERR_INVALID,
                "multigrid: halving the element counts per side (odd counts rounded up) must "
                "reach a coarsest grid of at most 2,200 DOFs");
  nd_coarse = (int)ndc;
  int rc;
  for (int l = 0; l < nl; l++) {
    MgLevel& L = lv[l];
    L.nx = dims[l].first;
    L.ny = dims[l].second;
    L.nn = (int64_t)(L.nx + 1) * (L.ny + 1);
    if ((rc = L.t.alloc(3 * L.nn))) return rc;
    if (l < nl - 1 && ((rc = L.x2.alloc(3 * L.nn)) || (rc = L.minv.alloc(6 * L.nn)))) return rc;
    if (l > 0 && ((rc = L.b.alloc(3 * L.nn)) || (rc = L.x.alloc(3 * L.nn)))) return rc;
    if (l > 0 && l < nl - 1 &&
        ((rc = L.st.alloc(45 * L.nn)) || (rc = L.dead.alloc(L.nn))))
      return rc;
  }
  lv[0].omega = omega0;
  // level 0 smooths with an FP32 copy of the plan's block-Jacobi inverse
  hipLaunchKernelGGL(k_mg_f32, dim3(mg_blocks(6 * lv[0].nn)), dim3(256), 0, stream, 6 * lv[0].nn,
                     (const double*)pl->minv.p, lv[0].minv.p);
  PP_HIP(hipGetLastError());
  if ((rc = dotp.alloc(2 * kMgDotBlocks))) return rc;
  // coarse element matrices level by level (two alive at a time)
  KGrid G = make_kgrid(pl);
  DevBuf<double> kfine, kcoarse;
  for (int l = 0; l + 1 < nl; l++) {
    const MgLevel &F = lv[l], &Cl = lv[l + 1];
    const int64_t nec = (int64_t)Cl.nx * Cl.ny;
    if ((rc = kcoarse.alloc(144 * nec))) return rc;
    const uint8_t* fmask = l == 0 ? (const uint8_t*)pl->cmask.p : (const uint8_t*)F.dead.p;
    hipLaunchKernelGGL(k_mg_coarse_elems, dim3(mg_blocks(nec, 64)), dim3(64), 0, stream, Cl.nx,
                       Cl.ny, F.nx, F.ny, l == 0 ? (const double*)nullptr : (const double*)kfine.p,
                       G, pl->E, pl->nu, fmask, kcoarse.p);
    PP_HIP(hipGetLastError());
    if (l + 1 < nl - 1) {
      MgLevel& C = lv[l + 1];
      hipLaunchKernelGGL(k_mg_stencil, dim3(mg_blocks(C.nn, 64)), dim3(64), 0, stream, C.nx, C.ny,
                         (const double*)kcoarse.p, C.st.p, C.dead.p, C.minv.p);
      PP_HIP(hipGetLastError());
    }
    std::swap(kfine.p, kcoarse.p);
    std::swap(kfine.n, kcoarse.n);
    kcoarse.release();
  }
  // coarsest: dense assembly on the host, identity on zero-diagonal DOFs
  {
    const MgLevel& C = lv[nl - 1];
    const int64_t ne = (int64_t)C.nx * C.ny;
    std::vector<double> ke(144 * ne);
    PP_HIP(hipMemcpyAsync(ke.data(), kfine.p, ke.size() * 8, hipMemcpyDeviceToHost, stream));
    PP_HIP(hipStreamSynchronize(stream));
    const int n = nd_coarse;
    std::vector<double> A((size_t)n * n, 0.0);
    for (int64_t e = 0; e < ne; e++) {
      const int ei = (int)(e % C.nx), ej = (int)(e / C.nx);
      int dof[12];
      for (int k = 0; k < 4; k++)
        for (int d = 0; d < 3; d++)
          dof[3 * k + d] = 3 * ((ej + corner_dy(k)) * (C.nx + 1) + ei + corner_dx(k)) + d;
      for (int a = 0; a < 12; a++)
        for (int b = 0; b < 12; b++) A[(size_t)dof[a] * n + dof[b]] += ke[e * 144 + a * 12 + b];
    }
    for (int i = 0; i < n; i++)
      if (A[(size_t)i * n + i] == 0.0) A[(size_t)i * n + i] = 1.0;
    if (!spd_inverse(A, n)) return fail(PP_ERR_INVALID, "multigrid: coarsest operator not SPD");
    if ((rc = cinv.alloc((int64_t)n * n))) return rc;
    PP_HIP(hipMemcpyAsync(cinv.p, A.data(), A.size() * 8, hipMemcpyHostToDevice, stream));
    PP_HIP(hipStreamSynchronize(stream));
  }
  kfine.release();
  for (int l = 1; l + 1 < nl; l++)
    if ((rc = estimate_omega(l))) return rc;
  return PP_OK;
}

int Mg::apply_level(int l, const double* x, double* y) {
  if (l == 0) return struct_apply_k(pl, x, y);
  const MgLevel& L = lv[l];
  hipLaunchKernelGGL(k_mg_apply, dim3(mg_blocks(L.nn)), dim3(256), 0, stream, L.nx, L.ny,
                     (const float*)L.st.p, x, y);
  PP_HIP(hipGetLastError());
  return PP_OK;
}

// w = 4 / (3 lambda_max(B A)) by 30 power iterations from a fixed start
// (x 1.05 against underestimation), as for level 0.
int Mg::estimate_omega(int l) {
  MgLevel& L = lv[l];
  const int64_t nd = 3 * L.nn;
  std::vector<double> h(nd);
  uint64_t s = 0x243F6A8885A308D3ull + l;
  std::vector<uint8_t> dead(L.nn);
  PP_HIP(hipMemcpyAsync(dead.data(), L.dead.p, L.nn, hipMemcpyDeviceToHost, stream));
  PP_HIP(hipStreamSynchronize(stream));
  for (int64_t i = 0; i < nd; i++) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    h[i] = dead[i / 3] >> (i % 3) & 1 ? 0.0 : (double)(s >> 11) * (1.0 / 9007199254740992.0) - 0.5;
  }
  PP_HIP(hipMemcpyAsync(L.x.p, h.data(), nd * 8, hipMemcpyHostToDevice, stream));
  double lam = 0.0;
  for (int it = 0; it < 30; it++) {
    int rc;
    if ((rc = apply_level(l, L.x.p, L.t.p))) return rc;
    hipLaunchKernelGGL(k_mg_point, dim3(mg_blocks(L.nn)), dim3(256), 0, stream, L.nn, 0,
                       (const float*)L.minv.p, 1.0, (const double*)L.t.p, (double*)nullptr, L.b.p,
                       (double*)nullptr);
    hipLaunchKernelGGL(k_mg_dot2, dim3(kMgDotBlocks), dim3(256), 0, stream, nd,
                       (const double*)L.b.p, (const double*)L.x.p, dotp.p);
    double hp[2 * kMgDotBlocks];
    PP_HIP(hipMemcpyAsync(hp, dotp.p, sizeof(hp), hipMemcpyDeviceToHost, stream));
    PP_HIP(hipStreamSynchronize(stream));
    double a = 0, b = 0;
    for (int k = 0; k < kMgDotBlocks; k++) {
      a += hp[2 * k];
      b += hp[2 * k + 1];
    }
    if (!(a > 0) || !(b > 0)) break;
    lam = sqrt(a / b);
    hipLaunchKernelGGL(k_mg_scale, dim3(mg_blocks(nd)), dim3(256), 0, stream, nd, L.b.p,
                       1.0 / sqrt(a));
    std::swap(L.x.p, L.b.p);
  }
  PP_HIP(hipGetLastError());
  L.omega = lam > 0 ? 4.0 / (3.0 * 1.05 * lam) : 0.5;
  return PP_OK;
}

// Chebyshev smoothing interval of a level: [hi / ratio, hi], hi = 1.05
// lambda_max(B A), lambda_max recovered from the level's weight
// w = 4 / (3 * 1.05 lambda_max).
static void cheb_interval(double omega, double ratio, double& theta, double& delta) {
  const double hi = 1.05 * 4.0 / (3.0 * 1.05 * omega), lo = hi / ratio;
  theta = 0.5 * (hi + lo);
  delta = 0.5 * (hi - lo);
}

int Mg::sweep(int l, int mode, const double* x, double* xo, double c1, double c2) {
  MgLevel& L = lv[l];
  if (l == 0) return struct_sweep(pl, mode, x, L.b.p, (const float*)L.minv.p, c1, c2, xo);
  if (mode == 1)
    hipLaunchKernelGGL(k_mg_sweep<1>, dim3(mg_blocks(L.nn)), dim3(256), 0, stream, L.nx, L.ny,
                       (const float*)L.st.p, (const float*)L.minv.p, c1, c2, x,
                       (const double*)L.b.p, xo);
  else
    hipLaunchKernelGGL(k_mg_sweep<2>, dim3(mg_blocks(L.nn)), dim3(256), 0, stream, L.nx, L.ny,
                       (const float*)L.st.p, (const float*)L.minv.p, c1, c2, x,
                       (const double*)L.b.p, xo);
  PP_HIP(hipGetLastError());
  return PP_OK;
}

// Smoothing: ns damped block-Jacobi sweeps (x += w B (b - A x)), or, by
// default, ns steps of the Chebyshev semi-iteration for B A x = B b on
// [hi / cheb_ratio, hi] (the first step d = B (b - A x) / theta, then
// d = rho_k rho_{k-1} d + (2 rho_k / delta) B (b - A x), rho_k = 1 / (2 sigma
// - rho_{k-1}), rho_0 = 1 / sigma; x += d): the same polynomial before and
// after the coarse correction keeps the cycle symmetric.  The sweeps
// ping-pong between L.x and L.x2: 2 ns - 1 of them follow the first (from
// x = 0), an odd count, so starting in x2 the cycle ends in L.x.
int Mg::cycle(int l) {
  MgLevel& L = lv[l];
  if (l == nl - 1) {
    hipLaunchKernelGGL(k_mg_dense, dim3(mg_blocks(nd_coarse, 4)), dim3(256), 0, stream, nd_coarse,
                       (const double*)cinv.p, (const double*)L.b.p, L.x.p);
    PP_HIP(hipGetLastError());
    return PP_OK;
  }
  const float* B = L.minv.p;
  const uint8_t* mask = l == 0 ? (const uint8_t*)pl->cmask.p : (const uint8_t*)L.dead.p;
  const int g = mg_blocks(L.nn);
  int rc;
  const int ns = l == 0 ? nu : nuc;
  double theta = 1.0, delta = 1.0;
  if (!jacobi) cheb_interval(L.omega, cheb_ratio, theta, delta);
  const double sigma = theta / delta;
  // x ping-pongs between L.x2 and L.x (2 ns - 1 steps end in L.x); a
  // Chebyshev step reads the previous iterate from the buffer it overwrites
  double* cur = L.x2.p;
  auto other = [&](double* p) { return p == L.x.p ? L.x2.p : L.x.p; };
  // step k of a smoothing pass (k = 0 starts it from the current x)
  double rho = 0.0;
  auto step = [&](int k) -> int {
    double c1 = 0.0, c2 = L.omega;
    if (!jacobi) {
      if (k == 0) {
        rho = 1.0 / sigma;
        c2 = 1.0 / theta;
      } else {
        const double rn = 1.0 / (2.0 * sigma - rho);
        c1 = rn * rho;
        c2 = 2.0 * rn / delta;
        rho = rn;
      }
    }
    const int r = sweep(l, 1, cur, other(cur), c1, c2);
    cur = other(cur);
    return r;
  };
  // pre-smoothing from x = 0: the first step is pointwise
  rho = 1.0 / sigma;
  hipLaunchKernelGGL(k_mg_point, dim3(g), dim3(256), 0, stream, L.nn, 0, B,
                     jacobi ? L.omega : 1.0 / theta, (const double*)L.b.p, (double*)nullptr, cur,
                     jacobi ? nullptr : other(cur));
  for (int s = 1; s < ns; s++)
    if ((rc = step(s))) return rc;
  if ((rc = sweep(l, 2, cur, L.t.p, 0.0, 0.0))) return rc;  // residual
  MgLevel& C = lv[l + 1];
  hipLaunchKernelGGL(k_mg_restrict, dim3(mg_blocks(C.nn)), dim3(256), 0, stream, L.nx, L.ny, mask,
                     (const double*)L.t.p, C.nx, C.ny, C.b.p);
  PP_HIP(hipGetLastError());
  if ((rc = cycle(l + 1))) return rc;
  hipLaunchKernelGGL(k_mg_prolong, dim3(g), dim3(256), 0, stream, L.nx, L.ny, mask,
                     (const double*)C.x.p, C.nx, cur);
  for (int s = 0; s < ns; s++)
    if ((rc = step(s))) return rc;
  PP_HIP(hipGetLastError());
  return PP_OK;
}

void mg_destroy(Mg* mg) { delete mg; }

int Mg::vcycle(const double* b, double* x) {
  MgLevel& L = lv[0];
  // level 0 reads b and writes x in place of its own buffers
  double* keep_b = L.b.p;
  double* keep_x = L.x.p;
  L.b.p = const_cast<double*>(b);
  L.x.p = x;
  const int rc = cycle(0);
  L.b.p = keep_b;
  L.x.p = keep_x;
  return rc;
}

}  // namespace pp
