// Internal declarations shared by the platepcg translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/platepcg.h"

namespace pp {

// ------------------------------------------------------------ errors

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define PP_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return ::pp::fail(PP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ----------------------------------------------------- device buffers

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  int alloc(size_t count) {
    release();
    if (count == 0) return PP_OK;
    hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      return fail(PP_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    n = count;
    return PP_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

// Grid barrier state for the persistent kernels: a monotonically increasing
// arrival counter and an abort flag raised by a workgroup that gave up
// waiting.  Zeroed before every persistent launch.
struct GridSync {
  unsigned int count;
  unsigned int abort;
  unsigned int pad[62];
};

// Device-resident solve statistics written by the persistent kernels.
struct DevStats {
  int iterations;
  int converged;
  int breakdown;
  int updated;  // mixed inner solve: stopped because |r| fell by the update factor
  double final_residual;
  double bnorm;
  double rz_prev;  // mixed inner solve: r.z of the iteration before the exit
  // wall-clock ticks (100 MHz) spent by workgroup 0 in each part of an
  // iteration, summed over iterations: stage, elements, node sums,
  // all-reduce A, phase B, all-reduce B
  unsigned long long ticks[6];
};

struct Cnn;
struct Mg;

}  // namespace pp

struct pp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int cu_count = 0;
  std::atomic<int> refs{1};  // the owner plus one per live plan or batch
  // Every entry point that works on this context's stream, events or the
  // state of its plans holds this lock, so calls from several host threads
  // (the Node addon runs solves on the libuv pool) are serialised per context.
  std::recursive_mutex mu;
};

// Holds a context's lock for the rest of the scope.
#define PP_LOCK(ctx) std::lock_guard<std::recursive_mutex> pp_lock_guard_((ctx)->mu)

struct pp_plan {
  pp_ctx* ctx = nullptr;
  int refs = 1;             // the owner plus one per live network on this plan
  pp::Cnn* learned = nullptr;  // preconditioner for PP_PRECOND_LEARNED (not owned)
  pp::Mg* mg_pc = nullptr;     // V-cycle of PP_PRECOND_MULTIGRID (owned, built on first use)
  int64_t n_nodes = 0, n_elems = 0, n_dof = 0;
  int npe = 4;
  bool structured = false;
  int nx = 0, ny = 0;
  double E = 0, nu = 0, t = 0;
  bool var_t = false;

  // host copies
  std::vector<float> h_nodes;
  std::vector<uint32_t> h_elems;
  std::vector<uint32_t> h_colour_off;
  std::vector<uint8_t> h_cmask;
  std::vector<float> h_elem_t;

  // device data
  pp::DevBuf<double> xc, yc;          // structured: grid coordinates
  pp::DevBuf<float> elem_t;           // per-element thickness (optional)
  pp::DevBuf<uint8_t> cmask;          // per-node PP_FIX bits
  pp::DevBuf<uint8_t> cmask_dof;      // per-DOF 0/1 (coloured path)
  pp::DevBuf<double> minv;            // 6 per node (symmetric 3x3 inverse)
  double cheb_lmax = 0.0;             // lambda_max(B K) for PP_PRECOND_CHEBYSHEV (0: not yet)
  int cheb_degree = 8;                // degree of the solve being launched
  pp::DevBuf<double> ke;              // coloured path: per-element FP64 Ke
  pp::DevBuf<uint32_t> elems;         // coloured path: connectivity
  pp::DevBuf<uint32_t> by_colour;     // coloured path: elements by colour
  pp::DevBuf<float> gen_nodes;        // coloured path: f32 coordinates
  pp::DevBuf<int64_t> gen_off;        // coloured path: node -> (element, corner) list,
  pp::DevBuf<uint32_t> gen_adj;       //   CSR, entries element * npe + corner
  pp::DevBuf<int64_t> ell_slc;        // coloured path: K as 3x3 node blocks, sliced ELL
  pp::DevBuf<uint32_t> ell_col;       //   (64 nodes per slice): slice offsets and widths,
  pp::DevBuf<double> ell_blk;         //   neighbour per slot, 9 doubles per slot

  // solve state (FP64)
  pp::DevBuf<double> F, x, r, z, p0, p1, Ap, tmp;
  pp::DevBuf<float> F32;
  pp::DevBuf<double> partials;        // grid-reduction partials
  pp::DevBuf<pp::GridSync> gsync;
  pp::DevBuf<pp::DevStats> dstats;
  pp::DevBuf<double> scal;            // coloured path device scalars
  pp::DevBuf<double> normp;           // norm-reduction partials

  // post-processing (built on first use): mesh copies, node -> element
  // list (CSR), element moments, node moments, argmax partials
  pp::DevBuf<float> post_nodes;
  pp::DevBuf<uint32_t> post_elems;
  pp::DevBuf<int64_t> post_off;
  pp::DevBuf<uint32_t> post_adj;
  pp::DevBuf<double> post_em;
  pp::DevBuf<float> post_out;
  pp::DevBuf<double> post_red;

  // mixed precision state (FP32 inner solve)
  pp::DevBuf<float> xs, rs, zs, ps0, ps1, Aps;
  pp::DevBuf<float> kmix;             // shared element matrix: bending | shear, 16x16 each

  bool rhs_set = false;
  double rhs_norm = 0;
  void* mgf_pinned = nullptr;  // pinned host ring of the fused multigrid PCG's state
};

namespace pp {

// structured kernels (pp_struct.hip)
int struct_setup(pp_plan* pl);
int struct_apply_k(pp_plan* pl, const double* dx, double* dy);
// mode 1: xo = x + c1 (x - xo) + c2 B (b - K x) (B: 6 floats per node, a copy
// of the plan's block-Jacobi inverse; xo, the previous iterate, is read only
// when c1 != 0); mode 2: xo = b - K x.  xo must not alias x.
int struct_sweep(pp_plan* pl, int mode, const double* x, const double* b, const float* minv,
                 double c1, double c2, double* xo);
int struct_apply_precond(pp_plan* pl, const double* dr, double* dz);
int struct_pressure(pp_plan* pl, double q0, const double* d_patches, int np, float* dF);
int struct_solve_fp64(pp_plan* pl, const pp_solve_opts* o, pp_solve_stats* st);
int cheb_lambda_max(pp_plan* pl);  // lambda_max(B K) by power iteration -> pl->cheb_lmax
// PCG preconditioned by a Galerkin multigrid V-cycle (pp_fcg.hip, pp_mg.hip;
// the fused launches in pp_mgf.hip)
int mg_solve(pp_plan* pl, const pp_solve_opts* o, pp_solve_stats* st);
int mgf_solve(pp_plan* pl, Mg* mg, const pp_solve_opts* o, pp_solve_stats* st);
void mg_destroy(Mg* mg);
int struct_solve_mixed(pp_plan* pl, const pp_solve_opts* o, pp_solve_stats* st);
int struct_true_residual(pp_plan* pl, const double* dx, double* rel);

// coloured kernels (pp_general.hip)
int general_setup(pp_plan* pl, const pp_colouring* col);
int general_apply_k(pp_plan* pl, const double* dx, double* dy, int colour, bool bc_rows);
int general_apply_precond(pp_plan* pl, const double* dr, double* dz);
int general_solve_fp64(pp_plan* pl, const pp_solve_opts* o, pp_solve_stats* st);
int general_true_residual(pp_plan* pl, const double* dx, double* rel);
int general_pressure(pp_plan* pl, double q0, const double* d_patches, int np, float* dF);
// host FP64 element stiffness: Q4 12x12 (xy: 4 corners), DKT 9x9 (xy: 3 corners)
void element_ke_q4(const double* xy, double E, double nu, double t, double* Ke);
void element_ke_dkt(const double* xy, double E, double nu, double t, double* Ke);

// device address of the no-allocation scratch array on `device` (current),
// looked up once per device (pp_noalloc.hip)
int noalloc_scratch_init(int device);

// lifetime (pp_api.hip)
void ctx_release(pp_ctx* c);
void plan_release(pp_plan* pl);

// Buffers of the structured FP64 solve: right-hand side, solution, the two
// exchange vectors of the persistent kernel, a DOF vector for the true
// residual, grid-reduction partials (part_n doubles), norm partials
// (kNormParts doubles), the grid barrier words and the device statistics.
// From the plan (pp_solve_resident) or from the no-allocation scratch array.
struct StructWs {
  const double* F;
  double* x;
  double *pb0, *pb1, *tmp;
  double* part;
  size_t part_n;
  double* normp;
  GridSync* gs;
  DevStats* ds;
};
constexpr int kNormBlocks = 256;
constexpr int kNormParts = 2 * kNormBlocks;
// The FP64 structured solve on explicit buffers; resident_only: fail with
// PP_ERR_UNSUPPORTED unless the single-reduction resident kernel is used
// (the only one that needs no further plan vectors).
int struct_solve_fp64_ws(pp_plan* pl, const pp_solve_opts* o, pp_solve_stats* st,
                         const StructWs& ws, bool resident_only);
// Grid-reduction partials the persistent kernel of this plan needs, in doubles.
int struct_partials_needed(pp_plan* pl, int variant, size_t* n);

// shared helpers (pp_api.hip)
// launch_norms with the partials in part (kNormParts doubles)
int launch_norms_into(pp_plan* pl, double* part, const double* dF, const double* dKx, double* res,
                      double* fnorm);
int launch_norm_ratio(pp_plan* pl, const double* dF, const double* dKx, double* rel);
// |F - Kx| and |F| (FP64, fixed-order reduction)
int launch_norms(pp_plan* pl, const double* dF, const double* dKx, double* res, double* fnorm);
// out = F - Kx (out may alias Kx)
int launch_residual(pp_plan* pl, const double* dF, const double* dKx, double* out);

}  // namespace pp
