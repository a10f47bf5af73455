// Learned preconditioner: network state shared by pp_cnn.hip (kernels,
// training) and pp_fcg.hip (flexible PCG).
#pragma once

#include <vector>

#include "pp_grid.h"

namespace pp {

// channels-last BF16 tensor view: (pixel, c) at p[pixel * cs + c0 + c]
struct TV {
  uint16_t* p;
  int cs, c0;
};

struct Layer {
  int cin, cout, relu, dil;
  int coutP, coutA, cinP, kp, kpT;  // padded sizes (16 / 4 / 16 / K of forward / K of transpose)
  int64_t w_off, b_off;             // FP32 master weights [cout][9][cin], bias [cout]
  int64_t wf_off, wt_off;           // packed BF16 forward / backward-data weights
};

// Galerkin geometric multigrid on the structured node grid, used as the
// smoother around the network when the configuration asks for it
// (pp_mg.hip).  Level 0 is the plan's own operator (matrix-free FP64 K, its
// block-Jacobi inverse and constraint bits); level l+1 halves the element
// counts of level l (vertex-centred: coarse node I is fine node 2I) with
// bilinear prolongation P_l (index space, per DOF component, zero on the
// excluded DOFs of level l) and the coarse operator A_{l+1} = P_l^T A_l P_l,
// assembled from coarse element matrices (each the sum over its four child
// elements of P_loc^T K_child P_loc, which is exact for bilinear P).  Coarse
// levels keep A as an FP32 9-point block stencil; the coarsest is solved by
// a dense FP64 inverse.
struct MgLevel {
  int nx = 0, ny = 0;    // elements per side
  int64_t nn = 0;        // nodes (nx + 1) (ny + 1)
  double omega = 0.5;    // block-Jacobi smoothing weight 4 / (3 lambda_max(B A))
  DevBuf<float> st;      // level >= 1: [5 blocks][3][3][node]: centre, E, NW, N, NE
  DevBuf<float> minv;    // 6 per node (symmetric 3x3 block inverse, FP32 for smoothing)
  DevBuf<uint8_t> dead;  // level >= 1: per-node bits of DOFs with a zero diagonal
  DevBuf<double> b, x, x2, t;  // right-hand side, iterate and its sweep partner, work
  DevBuf<float> rows;    // levels >= 1, fused form: each node's full 9-block row, 84 floats
};

struct Mg {
  pp_plan* pl = nullptr;
  hipStream_t stream = nullptr;
  int nu = 6;   // smoothing steps before and after each coarse correction, level 0
  int nuc = 1;  // the same on levels >= 1
  static constexpr int kMaxLevels = 16;
  MgLevel lv[kMaxLevels];
  int nl = 0;
  int nd_coarse = 0;
  DevBuf<double> cinv;  // dense inverse of the coarsest operator [nd][nd]
  DevBuf<double> dotp;
  bool jacobi = false;  // smoothing by damped block Jacobi instead of Chebyshev steps
  // Chebyshev steps target [lambda_max / cheb_ratio, lambda_max] of B A.  Measured at 1024^2:
  // 16 suits the learned preconditioner's two V-cycles (105 flexible-PCG iterations against
  // 122-127 at 24-32); PCG with one V-cycle takes 32 (181 against 195 iterations)
  double cheb_ratio = 16.0;
  int setup(pp_plan* plan, double omega0, int sweeps, int coarse_sweeps, bool damped_jacobi);
  int vcycle(const double* b, double* x);  // x = V b on level 0 (device FP64 DOF vectors)
  int apply_level(int l, const double* x, double* y);  // y = A_l x
 private:
  int cycle(int l);
  int sweep(int l, int mode, const double* x, double* xo, double c1, double c2);
  int estimate_omega(int l);
};

// One multigrid V-cycle x = V b with the fused kernels of pp_mgf.hip, for use
// outside PCG (the learned preconditioner's smoother); the hierarchy is mg's.
struct MgfCycle;
int mgf_cycle_create(pp_plan* pl, Mg* mg, MgfCycle** out);
int mgf_cycle_apply(MgfCycle* c, const double* b, double* x);  // b and x must not alias
void mgf_cycle_destroy(MgfCycle* c);

struct Cnn {
  pp_plan* pl = nullptr;
  hipStream_t stream = nullptr;
  int cu = 0, nx = 0, ny = 0;
  size_t lds_max = 65536;  // dynamic LDS per workgroup of the device
  int H[3] = {0, 0, 0}, W[3] = {0, 0, 0};
  double sw = 1.0, st = 0.3;  // input scales of the w and rotation channels
  double omega = 0.5;         // block-Jacobi smoothing weight (S = w B, and level 0 of the V-cycle)
  int cat0C = 0, cat1C = 0;
  std::vector<Layer> layers;
  int64_t n_params = 0;
  int adam_t = 0;
  double last_grad_norm = 0.0;
  int linear = 1;       // no ReLU, no biases
  int line_search = 1;  // energy-optimal step on the network's correction (else unit step)
  int fcg_window = 8;   // directions kept by flexible PCG (FCG(m))
  Mg* mg = nullptr;     // V-cycle smoother in place of w B (null: w B)
  MgfCycle* mgc = nullptr;  // the same V-cycle on the fused kernels

  DevBuf<float> params, grads, adam_m, adam_v;
  DevBuf<uint16_t> wf, wt;
  DevBuf<uint16_t> xin, a0, cat0, r1, a1, cat1, r2, a2, bb, c1b, d1, c0, c0b;
  // the skip and prolonged tensors stored alone, for the forward pass without
  // backward (the decoder's first layers read their two input parts in place)
  DevBuf<uint16_t> s0a, s1a, p0a, p1a;
  bool split_fwd = false;  // the network's shapes allow that forward pass
  DevBuf<uint16_t> g_c0b, g_c0, g_cat0, g_d1, g_c1, g_cat1, g_bb, g_a2, g_r2, g_a1, g_r1, g_a0,
      g_out;
  DevBuf<float> out32, wpart, bpart, logt;
  DevBuf<double> zbj, mout, dvec, kvec, evec, rvec, scal, dotp, bumps, z0v, z1v, r1v, mgz, kbv;

  const TV* mk(const TV& v) const { return linear ? nullptr : &v; }  // ReLU mask or none
  int setup(pp_plan* plan, const pp_cnn_config* cfg, uint64_t seed);
  int pack_weights();
  // network alone, device vectors; keep = the activations the backward pass reads
  int apply(const double* r, double* out, bool keep = false);
  // smoothing + network + smoothing
  int precondition(const double* r, double* out, bool line_search, bool keep = false);
  int smooth(const double* in, const double* base, double* out);  // out = base + S in (base may be null)
  int estimate_omega();
  int backward();
  int loss_grad(const double* e, double* loss_dev);  // e on the device
  int gen_error(uint64_t seed, double* e);
  int adam_step(float lr, float max_norm);
  int conv_fwd(const Layer& L, const TV& in, const TV& out, float* out32, int out32_cs, int lev,
               bool backward, const TV* mask = nullptr, int mask_from = 0,
               const TV* first = nullptr, int first_ch = 0);
  // the first two convolutions (4 -> 16 -> 16 channels) in one launch, for
  // the forward pass without backward (k_conv2)
  static constexpr int kConv2Rows = 4;
  bool fuse_pairs = true;
  bool conv2_ok(const Layer& La, const Layer& Lb) const;
  int conv2_fwd(const Layer& La, const Layer& Lb, const TV& in, const TV& out, int lev);
  int convp_pick(int cin, int nb) const;   // k_convp shape index or -1
  size_t convp_lds(int cin, int k) const;  // its dynamic LDS bytes
  int conv_wgrad(const Layer& L, const TV& in, const TV& dout, int lev);
  int resample(bool restrict_op, const TV& in, int lin, const TV& out, int lout, int C,
               float scale, const TV* mask, bool accumulate);
  ~Cnn();
};

// flexible PCG with the network as preconditioner (pp_fcg.hip)
int learned_solve(pp_plan* pl, const pp_solve_opts* o, pp_solve_stats* st);

}  // namespace pp

struct pp_cnn {
  pp::Cnn net;
};
