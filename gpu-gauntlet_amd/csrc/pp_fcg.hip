// Flexible PCG with the learned preconditioner, the training loop, and the
// C ABI of networks.
//
// FCG(m) (Notay's flexible CG truncated to the last m directions):
//   r = F, z = M(r), p_0 = z
//   loop k: q_k = K p_k; alpha = p_k.r / p_k.q_k; x += alpha p_k; r -= alpha q_k
//           stop when |r| < tol or |r| / |F| < tol (the rule of pcg.ts:174)
//           z = M(r); p_{k+1} = z - sum over the last m directions j of
//                               (z.q_j / p_j.q_j) p_j
// The K-orthogonalisation of z against recent directions keeps the method
// convergent for a preconditioner that is not a fixed SPD matrix (a trained
// network is neither symmetric nor, with the energy-optimal step, linear);
// with m = 1 the orthogonality lost to such a preconditioner stalls it on
// large plates.  Directions and their K-products live in rings of m + 1
// vectors.  Scalars stay on the device; the host reads the convergence flag
// every few iterations, and the update kernels stop changing x and r once it
// is set.

#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <functional>
#include <vector>

#include "pp_cnn.h"

namespace pp {

namespace {

constexpr int kMaxWin = 16;

struct FcgState {
  double pq, pr, alpha, rr, fnorm, tol;
  double pqr[kMaxWin + 1];  // p.q of the directions in the ring
  double beta[kMaxWin];     // the coefficients of the next direction
  int it, done, breakdown, converged;
};

// the directions a new one is orthogonalised against: ring slots, count
struct Win {
  int slot[kMaxWin];
  int w;
};

constexpr int kBlocks = 1024;  // partial sums per reduction: 16 waves per CU stream the vectors
constexpr int kCheckEvery = 4;
// FP64 restarts after the first run (learned and multigrid solves).  On the
// 1024^2 log-normal plate 1e-8 is at the FP64 floor of the relative residual
// (the rounding of x and of K x, about eps |K| |x| / |F|): there restarts of
// 1-4 iterations each land between 0.95e-8 and 1.05e-8, and it takes several
// to draw one below the tolerance
constexpr int kMaxRestarts = 16;

int nb(int64_t n) { return (int)((n + 255) / 256); }

// part[2b] = a.b, part[2b+1] = c.d per block (fixed order)
__global__ void k_dots(int64_t n, const double* a, const double* b, const double* c,
                       const double* d, double* part) {
  __shared__ double r0[256], r1[256];
  double s0 = 0.0, s1 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    s0 += a[i] * b[i];
    if (c) s1 += c[i] * d[i];
  }
  r0[threadIdx.x] = s0;
  r1[threadIdx.x] = s1;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      r0[threadIdx.x] += r0[threadIdx.x + s];
      r1[threadIdx.x] += r1[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r0[0];
    part[2 * blockIdx.x + 1] = r1[0];
  }
}

// Fixed-order sum of the kBlocks pairs of partials by one kBlocks-thread
// block (every thread gets the result).
__device__ void sum2(const double* part, double& a, double& b) {
  __shared__ double sa[kBlocks], sb[kBlocks];
  const int t = threadIdx.x;
  sa[t] = part[2 * t];
  sb[t] = part[2 * t + 1];
  __syncthreads();
  for (int k = kBlocks / 2; k > 0; k >>= 1) {
    if (t < k) {
      sa[t] += sa[t + k];
      sb[t] += sb[t + k];
    }
    __syncthreads();
  }
  a = sa[0];
  b = sb[0];
}

// after q = K p: pq, pr -> alpha (breakdown when |p.q| < 1e-30, pcg.ts:154)
__global__ void k_fcg_alpha(const double* part, FcgState* S, int slot) {
  if (S->done) return;  // uniform over the block
  double pq, pr;
  sum2(part, pq, pr);
  if (threadIdx.x) return;
  S->pq = pq;
  S->pr = pr;
  S->pqr[slot] = pq;
  if (fabs(S->pq) < 1e-30) {
    S->breakdown = 1;
    S->done = 1;
    return;
  }
  S->alpha = S->pr / S->pq;
}

// x += alpha p; r -= alpha q; partial |r|^2
__global__ void k_fcg_update(int64_t n, double* x, double* r, const double* p, const double* q,
                             const FcgState* S, double* part) {
  __shared__ double red[256];
  double s = 0.0;
  if (!S->done) {
    const double a = S->alpha;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
      x[i] += a * p[i];
      const double ri = r[i] - a * q[i];
      r[i] = ri;
      s += ri * ri;
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = red[0];
    part[2 * blockIdx.x + 1] = 0.0;
  }
}

// |r| test (the OR rule) after an update
__global__ void k_fcg_check(const double* part, FcgState* S) {
  if (S->done) return;
  double rr, unused;
  sum2(part, rr, unused);
  if (threadIdx.x) return;
  S->rr = rr;
  S->it++;
  const double rn = sqrt(rr);
  if (rn < S->tol || rn / S->fnorm < S->tol) {
    S->converged = 1;
    S->done = 1;
  }
}

// after z = M(r): per-block partials of z.q_j for the window (block row j)
__global__ void k_fcg_zq(int64_t n, const double* z, const double* Q, Win win, const FcgState* S,
                         double* part) {
  __shared__ double red[256];
  if (S->done) return;
  const double* q = Q + (size_t)win.slot[blockIdx.y] * n;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    acc += z[i] * q[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)blockIdx.y * kBlocks + blockIdx.x] = red[0];
}

// beta_j = -(z.q_j) / (p_j.q_j), partials summed in a fixed order
__global__ void k_fcg_betas(const double* part, Win win, FcgState* S) {
  __shared__ double sa[kBlocks];
  if (S->done) return;
  for (int j = 0; j < win.w; j++) {
    sa[threadIdx.x] = part[(size_t)j * kBlocks + threadIdx.x];
    __syncthreads();
    for (int k = kBlocks / 2; k > 0; k >>= 1) {
      if (threadIdx.x < k) sa[threadIdx.x] += sa[threadIdx.x + k];
      __syncthreads();
    }
    if (threadIdx.x == 0) S->beta[j] = -sa[0] / S->pqr[win.slot[j]];
    __syncthreads();
  }
}

// p_out = z + sum_j beta_j p_j (no window: p_out = z)
__global__ void k_fcg_dir(int64_t n, const double* z, double* P, Win win, int out,
                          const FcgState* S, int first) {
  if (S->done && !first) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = z[i];
  for (int j = 0; j < win.w; j++) v += S->beta[j] * P[(size_t)win.slot[j] * n + i];
  P[(size_t)out * n + i] = v;
}

// dst = (float)(a - b) (b may be null)
__global__ void k_err32(int64_t n, const double* a, const double* b, float* dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (float)(a[i] - (b ? b[i] : 0.0));
}

__global__ void k_widen(int64_t n, const float* a, double* b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}

double splitmix_uni(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

__global__ void k_fcg_init(FcgState* S, double fnorm, double tol, double rr) {
  if (threadIdx.x) return;
  *S = FcgState{};
  S->fnorm = fnorm > 0.0 ? fnorm : 1.0;
  S->tol = tol;
  S->rr = rr;
}

}  // namespace

// z = M(r) on device FP64 DOF vectors
using Precond = std::function<int(const double* r, double* z)>;

// FCG(m) from x (zero on a cold start) with r = F - K x given in pl->r; P and
// Q are rings of m + 1 vectors each (directions, their K-products).
static int fcg_run(pp_plan* pl, const Precond& prec, int m, double fnorm, double tol, int maxit,
                   FcgState* hs, DevBuf<FcgState>& S, DevBuf<double>& part, double* P,
                   double* Q) {
  hipStream_t s = pl->ctx->stream;
  const int64_t n = pl->n_dof;
  const int R = m + 1;
  int rc;
  hipLaunchKernelGGL(k_fcg_init, dim3(1), dim3(64), 0, s, S.p, fnorm, tol, 0.0);
  // z = M(r), p_0 = z
  if ((rc = prec(pl->r.p, pl->z.p))) return rc;
  Win none{};
  hipLaunchKernelGGL(k_fcg_dir, dim3(nb(n)), dim3(256), 0, s, n, (const double*)pl->z.p, P, none,
                     0, (const FcgState*)S.p, 1);
  PP_HIP(hipGetLastError());
  for (int k = 0; k < maxit; k++) {
    const int cur = k % R;
    double* p = P + (size_t)cur * n;
    double* q = Q + (size_t)cur * n;
    if ((rc = struct_apply_k(pl, p, q))) return rc;
    hipLaunchKernelGGL(k_dots, dim3(kBlocks), dim3(256), 0, s, n, (const double*)p,
                       (const double*)q, (const double*)p, (const double*)pl->r.p, part.p);
    hipLaunchKernelGGL(k_fcg_alpha, dim3(1), dim3(kBlocks), 0, s, (const double*)part.p, S.p, cur);
    hipLaunchKernelGGL(k_fcg_update, dim3(kBlocks), dim3(256), 0, s, n, pl->x.p, pl->r.p,
                       (const double*)p, (const double*)q, (const FcgState*)S.p, part.p);
    hipLaunchKernelGGL(k_fcg_check, dim3(1), dim3(kBlocks), 0, s, (const double*)part.p, S.p);
    PP_HIP(hipGetLastError());
    if ((k + 1) % kCheckEvery == 0 || k + 1 == maxit) {
      PP_HIP(hipMemcpyAsync(hs, S.p, sizeof(FcgState), hipMemcpyDeviceToHost, s));
      PP_HIP(hipStreamSynchronize(s));
      if (hs->done) break;
    }
    if ((rc = prec(pl->r.p, pl->z.p))) return rc;
    Win win{};
    win.w = std::min(k + 1, m);
    for (int j = 0; j < win.w; j++) win.slot[j] = (k - j) % R;
    hipLaunchKernelGGL(k_fcg_zq, dim3(kBlocks, win.w), dim3(256), 0, s, n,
                       (const double*)pl->z.p, (const double*)Q, win, (const FcgState*)S.p,
                       part.p);
    hipLaunchKernelGGL(k_fcg_betas, dim3(1), dim3(kBlocks), 0, s, (const double*)part.p, win, S.p);
    hipLaunchKernelGGL(k_fcg_dir, dim3(nb(n)), dim3(256), 0, s, n, (const double*)pl->z.p, P, win,
                       (k + 1) % R, (const FcgState*)S.p, 0);
    PP_HIP(hipGetLastError());
  }
  PP_HIP(hipMemcpyAsync(hs, S.p, sizeof(FcgState), hipMemcpyDeviceToHost, s));
  PP_HIP(hipStreamSynchronize(s));
  return PP_OK;
}

// Flexible PCG with residual replacement: after convergence the true FP64
// residual is measured; while it is above tol, FCG restarts from x with that
// residual (at most kMaxRestarts times).  Replacing r inside the run instead
// was measured slower: near 1e-8 the true residual is mostly rounding noise.
static int fcg_solve(pp_plan* pl, const pp_solve_opts* o, pp_solve_stats* st,
                     const Precond& prec, int m) {
  hipStream_t s = pl->ctx->stream;
  const int64_t n = pl->n_dof;
  DevBuf<FcgState> S;
  DevBuf<double> part, ring;
  const int R = m + 1;
  int rc;
  if ((rc = S.alloc(1)) || (rc = part.alloc((size_t)std::max(2, kMaxWin) * kBlocks)) ||
      (rc = ring.alloc((size_t)2 * R * n)))
    return rc;
  PP_HIP(hipEventRecord(pl->ctx->ev0, s));
  PP_HIP(hipMemsetAsync(pl->x.p, 0, n * 8, s));
  PP_HIP(hipMemcpyAsync(pl->r.p, pl->F.p, n * 8, hipMemcpyDeviceToDevice, s));
  double res = 0, fnorm = 0;
  if ((rc = launch_norms(pl, pl->F.p, pl->F.p, &res, &fnorm))) return rc;
  FcgState hs;
  memset(&hs, 0, sizeof(hs));
  int total = 0, sweeps = 0;
  double rel = 0.0;
  if (fnorm < 1e-30) {
    hs.converged = 1;
  } else {
    for (int k = 0; k <= kMaxRestarts; k++) {
      // from the second restart on, aim the recursive residual at half the
      // tolerance: stopped just under it, restarts near the FP64 floor took one
      // iteration each and left the true residual where it was (1.01e-8 to
      // 1.04e-8 on the 1024^2 log-normal plate); the first restart keeps the
      // tolerance (with half of it the linear preconditioners took 25% more
      // iterations there)
      const double tol = k <= 1 ? o->tolerance : 0.5 * o->tolerance;
      if ((rc = fcg_run(pl, prec, m, fnorm, tol, o->max_iterations - total, &hs, S, part, ring.p,
                        ring.p + (size_t)R * n)))
        return rc;
      if (sweeps < PP_MAX_SWEEPS) st->sweep_iterations[sweeps] = hs.it;
      total += hs.it;
      if (o->flags & PP_FLAG_NO_TRUE_RESIDUAL) {  // one run, no FP64 check
        sweeps++;
        rel = 0.0;
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
  st->converged = hs.converged && ((o->flags & PP_FLAG_NO_TRUE_RESIDUAL) || rel <= o->tolerance);
  st->breakdown = hs.breakdown;
  st->sweeps = sweeps;
  st->final_residual = sqrt(hs.rr);
  st->true_rel_residual = rel;
  st->device_ms = ms;
  return PP_OK;
}

int learned_solve(pp_plan* pl, const pp_solve_opts* o, pp_solve_stats* st) {
  Cnn& net = *pl->learned;
  const bool ls = net.line_search != 0;
  return fcg_solve(
      pl, o, st, [&](const double* r, double* z) { return net.precondition(r, z, ls); },
      net.fcg_window);
}

// PCG with one Galerkin multigrid V-cycle per iteration (symmetric and
// linear, so FCG(1) is plain PCG); the hierarchy is built on the first such
// solve of the plan, with level 0 smoothing at w = 4 / (3 1.05 lambda_max(B K)),
// o->poly_degree Chebyshev steps on level 0 on [lambda_max / 32, lambda_max].
// The default of 5 steps was measured with the fused launches at 142^2 (49
// iterations, 4.96 ms; 6: 46, 5.08 ms; 3: 63, 5.28 ms) and with one launch per
// operation at 700^2 (64.1 ms; 6: 68.2) and 1024^2 (120.9 ms; 4: 123.7; 6: 133.8).
int mg_solve(pp_plan* pl, const pp_solve_opts* o, pp_solve_stats* st) {
  const int nu = o->poly_degree > 0 ? o->poly_degree : 5;
  if (nu > 8) return fail(PP_ERR_INVALID, "multigrid: poly_degree (smoothing steps) must be at most 8");
  int rc;
  if (!pl->mg_pc || pl->mg_pc->nu != nu) {
    mg_destroy(pl->mg_pc);
    pl->mg_pc = nullptr;
    if (pl->cheb_lmax <= 0 && (rc = cheb_lambda_max(pl))) return rc;
    Mg* mg = new Mg;
    mg->cheb_ratio = 32.0;
    if ((rc = mg->setup(pl, 4.0 / (3.0 * 1.05 * pl->cheb_lmax), nu, 0, false))) {
      delete mg;
      return rc;
    }
    pl->mg_pc = mg;
  }
  Mg* mg = pl->mg_pc;
  if (!(o->flags & PP_FLAG_MG_UNFUSED)) return mgf_solve(pl, mg, o, st);
  return fcg_solve(pl, o, st, [mg](const double* r, double* z) { return mg->vcycle(r, z); }, 1);
}

}  // namespace pp

using namespace pp;

extern "C" {

int pp_cnn_create(pp_plan* pl, const pp_cnn_config* cfg, uint64_t seed, pp_cnn** out) {
  if (!pl || !out) return fail(PP_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (!pl->structured) return fail(PP_ERR_UNSUPPORTED, "the learned preconditioner needs a structured grid");
  PP_LOCK(pl->ctx);
  PP_HIP(hipSetDevice(pl->ctx->device));
  pp_cnn* c = new pp_cnn();
  pl->refs++;
  int rc = c->net.setup(pl, cfg, seed);
  if (rc) {
    delete c;
    return rc;
  }
  *out = c;
  return PP_OK;
}

int pp_cnn_destroy(pp_cnn* c) {
  if (!c) return PP_OK;
  (void)hipSetDevice(c->net.pl->ctx->device);
  delete c;
  return PP_OK;
}

int64_t pp_cnn_param_count(const pp_cnn* c) { return c ? c->net.n_params : 0; }

double pp_cnn_smoother_weight(const pp_cnn* c) { return c ? c->net.omega : 0.0; }

int pp_cnn_get_params(pp_cnn* c, float* out) {
  if (!c || !out) return fail(PP_ERR_INVALID, "NULL argument");
  PP_LOCK(c->net.pl->ctx);
  Cnn& n = c->net;
  PP_HIP(hipSetDevice(n.pl->ctx->device));
  PP_HIP(hipMemcpyAsync(out, n.params.p, n.n_params * 4, hipMemcpyDeviceToHost, n.stream));
  PP_HIP(hipStreamSynchronize(n.stream));
  return PP_OK;
}

int pp_cnn_set_params(pp_cnn* c, const float* in) {
  if (!c || !in) return fail(PP_ERR_INVALID, "NULL argument");
  PP_LOCK(c->net.pl->ctx);
  Cnn& n = c->net;
  PP_HIP(hipSetDevice(n.pl->ctx->device));
  PP_HIP(hipMemcpyAsync(n.params.p, in, n.n_params * 4, hipMemcpyHostToDevice, n.stream));
  int rc = n.pack_weights();
  if (rc) return rc;
  PP_HIP(hipStreamSynchronize(n.stream));
  return PP_OK;
}

int pp_cnn_apply(pp_cnn* c, const double* r, double* z) {
  if (!c || !r || !z) return fail(PP_ERR_INVALID, "NULL argument");
  PP_LOCK(c->net.pl->ctx);
  Cnn& n = c->net;
  pp_plan* pl = n.pl;
  PP_HIP(hipSetDevice(pl->ctx->device));
  PP_HIP(hipMemcpyAsync(n.evec.p, r, pl->n_dof * 8, hipMemcpyHostToDevice, n.stream));
  int rc = n.precondition(n.evec.p, n.rvec.p, false);
  if (rc) return rc;
  PP_HIP(hipMemcpyAsync(z, n.rvec.p, pl->n_dof * 8, hipMemcpyDeviceToHost, n.stream));
  PP_HIP(hipStreamSynchronize(n.stream));
  return PP_OK;
}

int pp_cnn_forward(pp_cnn* c, const double* v, double* u) {
  if (!c || !v || !u) return fail(PP_ERR_INVALID, "NULL argument");
  PP_LOCK(c->net.pl->ctx);
  Cnn& n = c->net;
  pp_plan* pl = n.pl;
  PP_HIP(hipSetDevice(pl->ctx->device));
  PP_HIP(hipMemcpyAsync(n.evec.p, v, pl->n_dof * 8, hipMemcpyHostToDevice, n.stream));
  int rc = n.apply(n.evec.p, n.rvec.p);
  if (rc) return rc;
  PP_HIP(hipMemcpyAsync(u, n.rvec.p, pl->n_dof * 8, hipMemcpyDeviceToHost, n.stream));
  PP_HIP(hipStreamSynchronize(n.stream));
  return PP_OK;
}

int pp_cnn_mg_levels(const pp_cnn* c, int32_t* nx, int32_t* ny, double* omega, int32_t cap) {
  if (!c) return fail(PP_ERR_INVALID, "NULL argument");
  const Mg* mg = c->net.mg;
  if (!mg) return 0;
  for (int l = 0; l < mg->nl && l < cap; l++) {
    if (nx) nx[l] = mg->lv[l].nx;
    if (ny) ny[l] = mg->lv[l].ny;
    if (omega) omega[l] = l + 1 < mg->nl ? mg->lv[l].omega : 0.0;
  }
  return mg->nl;
}

int pp_cnn_mg_apply(pp_cnn* c, int32_t level, const double* x, double* y) {
  if (!c || !x || !y) return fail(PP_ERR_INVALID, "NULL argument");
  PP_LOCK(c->net.pl->ctx);
  Cnn& n = c->net;
  Mg* mg = n.mg;
  if (!mg) return fail(PP_ERR_UNSUPPORTED, "the network was created without multigrid");
  if (level < 0 || level >= mg->nl - 1)
    return fail(PP_ERR_INVALID, "level must be below the coarsest");
  PP_HIP(hipSetDevice(n.pl->ctx->device));
  MgLevel& L = mg->lv[level];
  double* dx = level ? L.x.p : n.evec.p;
  double* dy = level ? L.t.p : n.rvec.p;
  const int64_t nd = 3 * L.nn;
  PP_HIP(hipMemcpyAsync(dx, x, nd * 8, hipMemcpyHostToDevice, n.stream));
  int rc = mg->apply_level(level, dx, dy);
  if (rc) return rc;
  PP_HIP(hipMemcpyAsync(y, dy, nd * 8, hipMemcpyDeviceToHost, n.stream));
  PP_HIP(hipStreamSynchronize(n.stream));
  return PP_OK;
}

int pp_cnn_mg_cycle(pp_cnn* c, const double* b, double* x) {
  if (!c || !b || !x) return fail(PP_ERR_INVALID, "NULL argument");
  PP_LOCK(c->net.pl->ctx);
  Cnn& n = c->net;
  if (!n.mg) return fail(PP_ERR_UNSUPPORTED, "the network was created without multigrid");
  pp_plan* pl = n.pl;
  PP_HIP(hipSetDevice(pl->ctx->device));
  PP_HIP(hipMemcpyAsync(n.evec.p, b, pl->n_dof * 8, hipMemcpyHostToDevice, n.stream));
  int rc = mgf_cycle_apply(n.mgc, n.evec.p, n.rvec.p);  // the preconditioner's own cycle
  if (rc) return rc;
  PP_HIP(hipMemcpyAsync(x, n.rvec.p, pl->n_dof * 8, hipMemcpyDeviceToHost, n.stream));
  PP_HIP(hipStreamSynchronize(n.stream));
  return PP_OK;
}

int pp_cnn_loss_grad(pp_cnn* c, const double* e, double* loss, float* grad) {
  if (!c || !e || !loss) return fail(PP_ERR_INVALID, "NULL argument");
  PP_LOCK(c->net.pl->ctx);
  Cnn& n = c->net;
  pp_plan* pl = n.pl;
  PP_HIP(hipSetDevice(pl->ctx->device));
  PP_HIP(hipMemcpyAsync(n.evec.p, e, pl->n_dof * 8, hipMemcpyHostToDevice, n.stream));
  int rc = n.loss_grad(n.evec.p, n.scal.p + 6);
  if (rc) return rc;
  PP_HIP(hipMemcpyAsync(loss, n.scal.p + 6, 8, hipMemcpyDeviceToHost, n.stream));
  if (grad) PP_HIP(hipMemcpyAsync(grad, n.grads.p, n.n_params * 4, hipMemcpyDeviceToHost, n.stream));
  PP_HIP(hipStreamSynchronize(n.stream));
  return PP_OK;
}

int pp_cnn_gen_error(pp_cnn* c, uint64_t seed, double* e) {
  if (!c || !e) return fail(PP_ERR_INVALID, "NULL argument");
  PP_LOCK(c->net.pl->ctx);
  Cnn& n = c->net;
  PP_HIP(hipSetDevice(n.pl->ctx->device));
  int rc = n.gen_error(seed, n.evec.p);
  if (rc) return rc;
  PP_HIP(hipMemcpyAsync(e, n.evec.p, n.pl->n_dof * 8, hipMemcpyDeviceToHost, n.stream));
  PP_HIP(hipStreamSynchronize(n.stream));
  return PP_OK;
}

int pp_cnn_train(pp_cnn* c, const pp_train_opts* o, pp_train_stats* st) {
  if (!c || !o || !st) return fail(PP_ERR_INVALID, "NULL argument");
  PP_LOCK(c->net.pl->ctx);
  if (o->pairs <= 0 || o->epochs <= 0 || !(o->lr > 0)) return fail(PP_ERR_INVALID, "bad training options");
  if (o->source != PP_TRAIN_SYNTHETIC && o->source != PP_TRAIN_SOLVER &&
      o->source != PP_TRAIN_LEARNED)
    return fail(PP_ERR_INVALID, "unknown training source");
  Cnn& n = c->net;
  pp_plan* pl = n.pl;
  PP_HIP(hipSetDevice(pl->ctx->device));
  memset(st, 0, sizeof(*st));
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t nd = pl->n_dof;
  int rc;
  // solver pairs: per load case the FP64 solution x* (to 1e-12) and the
  // errors x* - x_k of its iterates k = 0, then geometric up to the iteration
  // count, kept on the device in FP32; block-Jacobi PCG (PP_TRAIN_SOLVER) or
  // flexible PCG with this network's current preconditioner (PP_TRAIN_LEARNED)
  const bool solver = o->source == PP_TRAIN_SOLVER || o->source == PP_TRAIN_LEARNED;
  Cnn* const attached = pl->learned;
  auto solve = [&](pp_solve_opts* so, pp_solve_stats* ss) {
    if (o->source == PP_TRAIN_SOLVER) return struct_solve_fp64(pl, so, ss);
    pl->learned = &n;
    const int r = learned_solve(pl, so, ss);
    pl->learned = attached;
    return r;
  };
  const int snaps = solver ? std::max(1, o->snapshots) : 1;
  const int per_epoch = o->pairs * snaps;
  DevBuf<float> bank;
  if (solver) {
    if ((rc = bank.alloc((size_t)per_epoch * nd))) return rc;
    const double x0 = pl->h_nodes[0], x1 = pl->h_nodes[2 * pl->nx];
    const double y0 = pl->h_nodes[1], y1 = pl->h_nodes[2 * ((size_t)pl->ny * (pl->nx + 1)) + 1];
    std::vector<float> F(nd);
    pp_solve_opts so;
    memset(&so, 0, sizeof(so));
    so.tolerance = 1e-12;
    so.max_iterations = 1000000;
    so.flags = PP_FLAG_NO_TRUE_RESIDUAL;
    pp_solve_stats ss;
    for (int l = 0; l < o->pairs; l++) {
      uint64_t sd = o->seed0 + (uint64_t)l;
      const int np = 1 + (int)(splitmix_uni(sd) * 3.0);
      std::vector<double> pt(4 * 3);
      for (int k = 0; k < np; k++) {
        pt[4 * k] = 1e4 + 4e4 * splitmix_uni(sd);
        pt[4 * k + 1] = x0 + (x1 - x0) * splitmix_uni(sd);
        pt[4 * k + 2] = y0 + (y1 - y0) * splitmix_uni(sd);
        pt[4 * k + 3] = 0.5 + 1.5 * splitmix_uni(sd);
      }
      if ((rc = pp_pressure_load(pl, 1e4, pt.data(), np, F.data()))) return rc;
      if ((rc = pp_set_rhs(pl, F.data()))) return rc;
      so.max_iterations = 1000000;
      if ((rc = solve(&so, &ss))) return rc;
      const int kstar = std::max(1, ss.iterations);
      PP_HIP(hipMemcpyAsync(n.evec.p, pl->x.p, nd * 8, hipMemcpyDeviceToDevice, n.stream));
      for (int j = 0; j < snaps; j++) {
        const int k = j == 0 ? 0 : (int)llround(pow((double)kstar, (double)j / snaps));
        float* dst = bank.p + ((size_t)l * snaps + j) * nd;
        if (k == 0) {
          hipLaunchKernelGGL(k_err32, dim3(nb(nd)), dim3(256), 0, n.stream, nd,
                             (const double*)n.evec.p, (const double*)nullptr, dst);
        } else {
          so.max_iterations = k;
          if ((rc = solve(&so, &ss))) return rc;
          hipLaunchKernelGGL(k_err32, dim3(nb(nd)), dim3(256), 0, n.stream, nd,
                             (const double*)n.evec.p, (const double*)pl->x.p, dst);
        }
        PP_HIP(hipGetLastError());
      }
    }
    PP_HIP(hipStreamSynchronize(n.stream));
  }
  const int steps = per_epoch * o->epochs;
  const int nbuck = std::min(steps, PP_TRAIN_CURVE);
  std::vector<double> bucket(nbuck, 0.0);
  std::vector<int> bcount(nbuck, 0);
  DevBuf<double> losses;
  if ((rc = losses.alloc(steps))) return rc;
  const double lr1 = o->lr_final > 0 ? o->lr_final : o->lr;
  for (int step = 0; step < steps; step++) {
    const int pair = step % per_epoch;
    if (solver) {
      hipLaunchKernelGGL(k_widen, dim3(nb(nd)), dim3(256), 0, n.stream, nd,
                         (const float*)bank.p + (size_t)pair * nd, n.evec.p);
      PP_HIP(hipGetLastError());
    } else if ((rc = n.gen_error(o->seed0 + (uint64_t)pair, n.evec.p))) {
      return rc;
    }
    if ((rc = n.loss_grad(n.evec.p, losses.p + step))) return rc;
    // cosine decay from lr to lr_final
    const double f = steps > 1 ? (double)step / (steps - 1) : 0.0;
    const double lr = lr1 + 0.5 * (o->lr - lr1) * (1.0 + cos(3.141592653589793 * f));
    if ((rc = n.adam_step((float)lr, (float)o->max_grad_norm))) return rc;
  }
  std::vector<double> h(steps);
  PP_HIP(hipMemcpyAsync(h.data(), losses.p, steps * 8, hipMemcpyDeviceToHost, n.stream));
  PP_HIP(hipStreamSynchronize(n.stream));
  for (int i = 0; i < steps; i++) {
    const int b = (int)((int64_t)i * nbuck / steps);
    bucket[b] += h[i];
    bcount[b]++;
  }
  for (int b = 0; b < nbuck; b++) st->loss_curve[b] = bcount[b] ? bucket[b] / bcount[b] : 0.0;
  st->curve_points = nbuck;
  st->steps = steps;
  st->loss_first = h[0];
  st->loss_last = h[steps - 1];
  st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return PP_OK;
}

int pp_plan_set_learned(pp_plan* pl, pp_cnn* c) {
  if (!pl) return fail(PP_ERR_INVALID, "NULL plan");
  if (c && c->net.pl != pl) return fail(PP_ERR_INVALID, "network was made for another plan");
  PP_LOCK(pl->ctx);
  pl->learned = c ? &c->net : nullptr;
  return PP_OK;
}

}  // extern "C"
