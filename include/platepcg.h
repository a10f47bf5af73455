/*
 * platepcg: MI355X (gfx950) plate-bending PCG solve behind a C ABI.
 *
 * Every entry point replaces a piece of the reference's GPU solve path
 * (cseifert512/gpu-gauntlet, src/lib/plate/gpu).  Plain pointers and sizes
 * only; all host pointers are borrowed for the duration of the call and every
 * output buffer is caller-allocated.  Functions return PP_OK (0) or a
 * negative PP_ERR_* code; pp_last_error() then holds a message for the
 * calling thread.
 *
 * Vectors use the reference's DOF numbering: 3 DOFs per node [w, thx, thy],
 * DOF index = 3*node + k (types.ts:175-188).
 */
#ifndef PLATEPCG_H
#define PLATEPCG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PP_OK 0
#define PP_ERR_INVALID (-1)     /* bad argument or inconsistent mesh */
#define PP_ERR_HIP (-2)         /* HIP runtime error */
#define PP_ERR_UNSUPPORTED (-3) /* option not available for this plan */
#define PP_ERR_NOMEM (-4)       /* device allocation failed */
#define PP_ERR_TIMEOUT (-5)     /* a persistent kernel's grid barrier gave up */
#define PP_ERR_NODEVICE (-6)    /* no HIP device */

/* Constrained-DOF bits per node in the mask arrays (solver.ts:92-134). */
#define PP_FIX_W 1u
#define PP_FIX_THX 2u
#define PP_FIX_THY 4u

/* Mesh, as PlateMesh (types.ts:73-89).  boundaryNodes is not needed here:
 * constraints arrive as a per-DOF mask (gpu/solver.ts:80-87 receives them as
 * a Set<number>; buffers.ts:225-235 turns that into a per-DOF mask). */
typedef struct {
  const float* nodes;        /* 2 per node: x, y (float32, as mesher.ts) */
  const uint32_t* elements;  /* nodes_per_element per element, CCW */
  uint32_t node_count;
  uint32_t element_count;
  uint32_t nodes_per_element; /* 4 = Q4 Mindlin, 3 = DKT triangle */
  uint32_t grid_nx;           /* structured grid size, 0 when unstructured */
  uint32_t grid_ny;
} pp_mesh;

/* Material, as PlateMaterial (types.ts:14-21).  element_t is an extension
 * with no reference counterpart: per-element thickness, NULL = uniform t. */
typedef struct {
  double E;
  double nu;
  double t;
  const float* element_t;
} pp_material;

/* Element colouring, as ElementColoring (types.ts:95-102): colour c holds
 * elements_by_colour[colour_offsets[c] .. colour_offsets[c+1]). */
typedef struct {
  const uint32_t* elements_by_colour;
  const uint32_t* colour_offsets; /* colour_count + 1 entries */
  uint32_t colour_count;
} pp_colouring;

#define PP_PRECISION_FP64 0  /* FP64 PCG throughout */
#define PP_PRECISION_MIXED 1 /* FP32 inner PCG, FP64 iterative refinement */

#define PP_PRECOND_BLOCK_JACOBI 0 /* 3x3 nodal block Jacobi (solver.ts:188-339) */
#define PP_PRECOND_LEARNED 1      /* CNN preconditioner, flexible PCG */
#define PP_PRECOND_CHEBYSHEV 2    /* block Jacobi inside a Chebyshev polynomial of degree
                                     poly_degree in B K (no reference counterpart): about
                                     poly_degree times fewer PCG iterations and all-reduces for
                                     about the same number of K applies; structured FP64 plates
                                     whose tiles are co-resident (up to about 1,000^2 nodes) */
#define PP_PRECOND_MULTIGRID 3    /* one Galerkin geometric-multigrid V-cycle per PCG iteration
                                     (no reference counterpart): Chebyshev smoothing, bilinear
                                     prolongation, dense coarsest solve; the hierarchy is built
                                     on the plan's first such solve.  Structured FP64 plates
                                     whose coarsest level (element counts halved, odd counts
                                     rounded up) has at most 2,200 DOFs */

#define PP_PATH_AUTO 0       /* structured kernels when the mesh is a grid */
#define PP_PATH_STRUCTURED 1 /* tensor-grid tile kernels (error if not a grid) */
#define PP_PATH_COLOURED 2   /* stored-Ke element kernels, one pass per colour */

/* Solver options, as GPUSolverOptions (gpu/solver.ts:35-42) plus the
 * precision / preconditioner / kernel-path choices of this library. */
typedef struct {
  double tolerance;      /* stop when |r| < tol or |r|/|F| < tol (pcg.ts:174) */
  int32_t max_iterations;
  int32_t precision;     /* PP_PRECISION_* */
  int32_t precond;       /* PP_PRECOND_* */
  int32_t path;          /* PP_PATH_* */
  double inner_tolerance;  /* mixed: a sweep ends when the FP32 residual has
                              fallen by this factor (<= 0: 1e-3) */
  int32_t max_refinements; /* mixed: most FP64 refinement sweeps (<= 0: 200) */
  int32_t flags;           /* PP_FLAG_* */
  int32_t poly_degree;     /* PP_PRECOND_CHEBYSHEV: degree (<= 0: 8, at most 32);
                              PP_PRECOND_MULTIGRID: smoothing steps on the finest level
                              (<= 0: 5; at most 8) */
} pp_solve_opts;

/* Skip the FP64 |F - K x| / |F| recomputation after the solve, and with it
 * the residual-replacement restarts (stats.true_rel_residual is then 0 and
 * converged reflects the recurrence alone).  Mixed precision ignores it: its
 * refinement sweeps are built on that FP64 residual. */
#define PP_FLAG_NO_TRUE_RESIDUAL 1
/* Structured path: use the 15x15-node tiles even where the plate fits the
 * large-tile persistent kernel (for comparison and testing). */
#define PP_FLAG_SMALL_TILES 2
/* Structured path: the two-reduction PCG kernel (the recurrence order of
 * pcg.ts) instead of the default single-reduction kernel. */
#define PP_FLAG_CLASSIC_PCG 4
/* Mixed precision: keep the inner search direction across the FP64
 * residual replacement ("reliable update") instead of restarting the inner
 * PCG from p = z at every refinement sweep. */
#define PP_FLAG_RELIABLE_UPDATE 8
/* PP_PRECOND_MULTIGRID: one launch per smoothing step, transfer and PCG
 * operation (34 per iteration on the 142^2 plate) instead of the fused
 * launches (19), for comparison; same arithmetic up to the order of the
 * dot-product sums. */
#define PP_FLAG_MG_UNFUSED 16
/* Coloured path: the launch-per-operation PCG loop (a launch per colour of
 * every K apply, device scalars, the host polling a stop flag every 64
 * iterations) instead of the persistent kernel, which applies K from 3x3
 * node blocks assembled at set-up; the same iterates up to rounding. */
#define PP_FLAG_COLOURED_HOST_LOOP 32

#define PP_MAX_SWEEPS 32

/* Solve statistics, as GPUSolveResult (gpu/solver.ts:47-60) minus the
 * solution, plus what the FP64 check needs. */
typedef struct {
  int32_t iterations;         /* total PCG iterations (all sweeps) */
  int32_t converged;
  int32_t breakdown;          /* |pAp| < 1e-30 (pcg.ts:154-161) */
  int32_t sweeps;             /* mixed: refinement sweeps done */
  double final_residual;      /* |r| of the recurrence at exit */
  double true_rel_residual;   /* |F - K x| / |F| recomputed in FP64 */
  double device_ms;           /* HIP-event time of the resident solve */
  int32_t sweep_iterations[PP_MAX_SWEEPS];
  double sweep_residual[PP_MAX_SWEEPS]; /* true FP64 rel. residual after each sweep */
  /* Wall time (us) workgroup 0 of the resident persistent kernel spent in
   * each part of an iteration, summed over iterations: halo/stage, element
   * apply, node sums, all-reduce of p.Ap, vector updates, all-reduce of
   * r.z and r.r; on the coloured path's persistent kernel: vector updates
   * and M r, grid barrier, K u, all-reduce.  Zero where not measured. */
  double phase_us[6];
} pp_solve_stats;

typedef struct pp_ctx pp_ctx;
typedef struct pp_plan pp_plan;

/* Capability probe; replaces isWebGPUAvailable (gpu/context.ts:19-21).
 * Returns the number of HIP devices (0 = none). */
int pp_device_count(void);

/* Message of the last failing call on this thread. */
const char* pp_last_error(void);

/* Device + stream; replaces initGPU / the cached context
 * (gpu/context.ts:31-106). */
int pp_ctx_create(int device, pp_ctx** out);
int pp_ctx_destroy(pp_ctx* ctx);

/* One-time set-up for a mesh: replaces the set-up half of solveGPUInternal
 * (gpu/solver.ts:118-150: block diagonal + inverse, buffer creation and
 * uploads, colour offsets).  The FP64 3x3 block inverses are built on the
 * device.  constrained_mask holds PP_FIX_* bits per node. */
int pp_plan_create(pp_ctx* ctx, const pp_mesh* mesh, const pp_material* material,
                   const pp_colouring* colouring, const uint8_t* constrained_mask,
                   pp_plan** out);
int pp_plan_destroy(pp_plan* plan);

/* 1 when the plan runs the structured tile kernels, 0 for the coloured path. */
int pp_plan_is_structured(const pp_plan* plan);

/* Upload F and make it the resident right-hand side: replaces
 * device.queue.writeBuffer(buffers.r, F) (gpu/solver.ts:153). */
int pp_set_rhs(pp_plan* plan, const float* F);

/* PCG on the resident F; x stays resident.  Replaces runGPUPCG
 * (gpu/solver.ts:190-553), with the whole loop on the device. */
int pp_solve_resident(pp_plan* plan, const pp_solve_opts* opts, pp_solve_stats* stats);

/* Read the resident solution: replaces readFromBuffer(x) (buffers.ts:315).
 * Either pointer may be NULL. */
int pp_get_solution(pp_plan* plan, double* x64, float* x32);

/* set_rhs + solve_resident + get_solution: the whole of solveGPU
 * (gpu/solver.ts:80-185) for one load case. */
int pp_solve(pp_plan* plan, const float* F, const pp_solve_opts* opts, float* x32, double* x64,
             pp_solve_stats* stats);

/* The FP64 block-Jacobi solve of pp_solve_resident (structured plan, the
 * single-reduction PCG with every 18x18 tile resident, residual-replacement
 * restarts and the FP64 true-residual check) on caller device buffers:
 * F_dev and x_dev are device pointers to n_dof = 3 * node_count doubles; F is
 * read, the solution written to x.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It calls no
 * hipMalloc / hipFree: every scratch buffer (exchange vectors, residual
 * vector, reduction partials, barrier words, statistics) lives in one
 * 16 MiB device array per GPU, held until the process ends, whose device
 * address is looked up when a context is created.  Plates above about
 * 690,000 DOFs (the three scratch vectors must fit) are rejected with
 * PP_ERR_INVALID; other preconditioners, precisions and the strided
 * kernel with PP_ERR_UNSUPPORTED.  Calls on one GPU share the array and
 * are serialised by a lock; the array keeps the last call's contents.
 * The plan's own resident F and x are not touched. */
int pp_solve_device_noalloc(pp_plan* plan, const double* F_dev, double* x_dev,
                            const pp_solve_opts* opts, pp_solve_stats* stats);

/* y = K x with the reference's BC rows (y[c] = x[c] on constrained DOFs):
 * replaces gpuApplyK (gpu/solver.ts:413-461) + apply_k_q4/apply_k_dkt +
 * apply_bc.  path picks the kernels (PP_PATH_*). */
int pp_apply_k(pp_plan* plan, const double* x, double* y, int path);

/* One colour only, scatter skipping constrained DOFs and no BC rows: the
 * GPU side of applySingleColorKx (gpu/single_color_test.ts:44-171). */
int pp_apply_k_colour(pp_plan* plan, int colour, const double* x, double* y);

/* z = M^-1 r with the FP64 block inverses: replaces the block
 * preconditioner shader (gpu/shaders/index.ts:214-262). */
int pp_apply_precond(pp_plan* plan, const double* r, double* z);

/* The FP64 block inverses, 9 per node row-major (solver.ts:231-304). */
int pp_get_block_inverse(pp_plan* plan, double* blocks);

/* Consistent nodal pressure load for p(x,y) = q0 + sum_k A_k
 * exp(-|x-c_k|^2 / (2 s_k^2)); patches holds (A, cx, cy, s) per patch.  F is
 * zero on constrained DOFs and on the rotation DOFs.  No reference
 * counterpart: the reference has only point loads (solver.ts:363-392). */
int pp_pressure_load(pp_plan* plan, double q0, const double* patches, int n_patches, float* F);

/* Log-normal element thickness of the learned-preconditioner workload,
 * synthesised on the device: g is a unit-variance Gaussian random field on
 * the element centres of an nx x ny grid of a width x height plate with
 * squared-exponential covariance exp(-d^2 / (2 corr_len^2)) (spectral
 * synthesis from the SplitMix64 normals of `seed` on a grid padded to
 * 2nx x 2ny, cropped, centred and scaled to unit sample variance), and
 * t_e = t_median exp(sigma_log g_e) as float32.  field (nx * ny doubles) or
 * element_t (nx * ny floats) may be NULL, not both; both are host pointers,
 * row-major with x fastest (element j * nx + i).  No reference counterpart:
 * the reference's plates have uniform thickness (types.ts PlateMaterial). */
int pp_lognormal_thickness(pp_ctx* ctx, int32_t nx, int32_t ny, double width, double height,
                           double corr_len, uint64_t seed, double t_median, double sigma_log,
                           double* field, float* element_t);

/* Bending moments at the nodes: replaces computeMoments
 * (postprocess.ts:40-270).  Curvatures at each element's centroid (Q4: the
 * isoparametric derivatives at xi = eta = 0; triangles: linear shape
 * functions, degenerate ones skipped), M = D [k_x + nu k_y, k_y + nu k_x,
 * (1 - nu) k_xy / 2] with D = E t^3 / (12 (1 - nu^2)) (per element with
 * per-element thickness), averaged over each node's elements.  x: DOF vector
 * [w, thx, thy] per node, or NULL for the plan's current solution on the
 * device.  Mx, My, Mxy: node_count floats each (any may be NULL).  FP64
 * arithmetic, float32 results. */
int pp_moments(pp_plan* plan, const double* x, float* Mx, float* My, float* Mxy);

/* Largest |w| (findMaxDisplacement, postprocess.ts:296-319): the signed w,
 * its node (the lowest index among equal |w|) and that node's coordinates.
 * x as for pp_moments.  Any output may be NULL. */
int pp_max_displacement(pp_plan* plan, const double* x, double* w, int64_t* node, float* node_x,
                        float* node_y);

/* ------------------------------------------------------------------------
 * Learned preconditioner (no reference counterpart: the reference has only
 * block Jacobi, solver.ts:188-339).  A 3-level U-net on the nodal grid of a
 * structured plan, BF16 MFMA with FP32 accumulation, trained on the device
 * on pairs (r = K e, e) built from the plan's own operator, and used by
 * pp_solve_resident with precond = PP_PRECOND_LEARNED inside flexible PCG.
 * A network holds a reference to its plan; destroy it before or after the
 * plan, either order is safe.
 * ---------------------------------------------------------------------- */

typedef struct pp_cnn pp_cnn;

typedef struct {
  int32_t channels[3];         /* per level; 0 = 16 / 32 / 64 */
  int32_t bottleneck_dilation; /* dilation of the last coarse conv; 0 = 1 */
  double rotation_scale;       /* input scale of thx/thy relative to w (metres); 0 = 0.3 */
  double smoother_weight;      /* block-Jacobi smoothing weight; 0 = 4 / (3 lambda_max(B K)) */
  int32_t relu;                /* 0: no activation and no biases, so the network is a fixed
                                  linear operator (flexible PCG then converges like PCG);
                                  1: ReLU after every conv but the last, with biases */
  int32_t fixed_step;          /* 0: scale the network's correction by its energy-optimal
                                  step in every application, so an application is never
                                  worse than the smoothing alone (the preconditioner is then
                                  nonlinear: use fcg_window > 1); 1: unit step (with relu = 0
                                  a fixed linear map, as in training, but a trained network
                                  need not then be a convergent preconditioner) */
  int32_t fcg_window;          /* directions flexible PCG K-orthogonalises against (FCG(m));
                                  0 = 8 (4 with multigrid = 1), at most 16.  With the energy-optimal step, FCG(1)
                                  stalls on large plates (1024^2, zero network: 5e-3 after
                                  30,000 iterations) where FCG(8) converges (29,897) */
  int32_t multigrid;           /* 0: the smoother S around the network is w B (block Jacobi);
                                  1: S is a Galerkin geometric-multigrid V-cycle (bilinear
                                  prolongation, Chebyshev or damped block-Jacobi smoothing, dense
                                  coarsest solve).  Needs the element counts per side to halve
                                  down to a coarsest grid of at most 2,200 DOFs */
  int32_t mg_sweeps;           /* smoothing steps before and after each coarse correction on
                                  the finest level, at most 8; 0 = 6 (Chebyshev) or 3 (Jacobi) */
  int32_t mg_coarse_sweeps;    /* the same on the coarser stencil levels; 0 = 1 */
  int32_t mg_jacobi;           /* smoothing steps: 0 = Chebyshev semi-iteration for B A on
                                  [lambda_max / 16, lambda_max] (the default), 1 = damped block
                                  Jacobi w B, w = 4 / (3 lambda_max) */
  int32_t unfused_forward;     /* 1: every convolution of the solver's forward pass runs as its
                                  own launch; 0 (default): the first two (4 -> 16 -> 16
                                  channels) run as one launch with the intermediate tensor in
                                  LDS (the same bits) */
} pp_cnn_config;

#define PP_TRAIN_CURVE 64

#define PP_TRAIN_SYNTHETIC 0 /* e: seeded multi-scale smooth fields, r = K e */
#define PP_TRAIN_SOLVER 1    /* e = x* - x_k: block-Jacobi PCG iterates on seeded pressure loads */
#define PP_TRAIN_LEARNED 2   /* e = x* - x_k: iterates of flexible PCG with the network's current
                                preconditioner (x* by the same solve to 1e-12), so the pairs are
                                the errors that preconditioner leaves */

typedef struct {
  int32_t pairs;        /* synthetic: pairs (seeds seed0 ..); solver: load cases (seeds seed0 ..) */
  int32_t epochs;
  uint64_t seed0;
  int32_t source;       /* PP_TRAIN_* */
  int32_t snapshots;    /* solver: iterates per load case, k = 0 and geometric up to convergence */
  double lr;            /* Adam learning rate (cosine-decayed to lr_final) */
  double lr_final;      /* 0 = constant lr */
  double max_grad_norm; /* global gradient-norm clip; 0 = none */
} pp_train_opts;

typedef struct {
  int32_t steps;
  int32_t curve_points;
  double loss_first, loss_last; /* relative energy-norm error of the first / last pair */
  double seconds;
  double loss_curve[PP_TRAIN_CURVE]; /* mean loss over equal slices of the steps */
} pp_train_stats;

/* Build a network for a structured plan; weights He-initialised from seed. */
int pp_cnn_create(pp_plan* plan, const pp_cnn_config* config, uint64_t seed, pp_cnn** out);
int pp_cnn_destroy(pp_cnn* net);
int64_t pp_cnn_param_count(const pp_cnn* net);
/* FP32 parameters, per conv layer: weights [cout][3][3][cin] then bias [cout]. */
int pp_cnn_get_params(pp_cnn* net, float* params);
int pp_cnn_set_params(pp_cnn* net, const float* params);
/* The smoothing weight w in use. */
double pp_cnn_smoother_weight(const pp_cnn* net);
/* z = M(r) for host vectors of 3 * node_count doubles: the preconditioner
 * z0 = S r; z1 = z0 + N(r - K z0); z = z1 + S (r - K z1), S the smoother
 * (w B, B the 3x3 block-Jacobi inverse, or with multigrid = 1 one V-cycle),
 * N the network (N(v) = B v + its learned correction). */
int pp_cnn_apply(pp_cnn* net, const double* r, double* z);
/* The network alone: u = N(v) = B v + its learned correction (host vectors
 * of 3 * node_count doubles); for parity tests and for building other
 * preconditioners around it.  No reference counterpart. */
int pp_cnn_forward(pp_cnn* net, const double* v, double* u);
/* Multigrid smoother diagnostics (config multigrid = 1; no reference
 * counterpart).  Levels: returns the level count (0 without multigrid) and
 * fills up to cap entries of the element counts per side and the smoothing
 * weight of each level (0 on the coarsest, which is solved exactly). */
int pp_cnn_mg_levels(const pp_cnn* net, int32_t* nx, int32_t* ny, double* omega, int32_t cap);
/* y = A_l x on level l (below the coarsest), host vectors of
 * 3 (nx_l + 1) (ny_l + 1) doubles; level 0 is the plan's K with its
 * constraint rows. */
int pp_cnn_mg_apply(pp_cnn* net, int32_t level, const double* x, double* y);
/* x = V b: one V-cycle on level 0 (host vectors of 3 * node_count doubles). */
int pp_cnn_mg_cycle(pp_cnn* net, const double* b, double* x);
/* Loss (M(K e) - e)^T K (M(K e) - e) / e^T K e and its gradient in the
 * parameters (grad may be NULL). */
int pp_cnn_loss_grad(pp_cnn* net, const double* e, double* loss, float* grad);
/* The training error field of one seed (host output, 3 * node_count). */
int pp_cnn_gen_error(pp_cnn* net, uint64_t seed, double* e);
/* Adam, one pair per step.  The solver source overwrites the plan's resident
 * right-hand side and solution. */
int pp_cnn_train(pp_cnn* net, const pp_train_opts* opts, pp_train_stats* stats);
/* Attach (or detach with NULL) the preconditioner PP_PRECOND_LEARNED uses. */
int pp_plan_set_learned(pp_plan* plan, pp_cnn* net);

/* ------------------------------------------------------------------------
 * Batched plates (no reference counterpart: the reference solves one plate
 * per solveGPU call).  Many rectangular-grid plates, Q4 (grid mesh) or DKT
 * (each cell split into (n0,n1,n2), (n0,n2,n3)), each with its own material,
 * supports and load, solved to the same tolerance by one persistent launch;
 * the FP64 block-Jacobi PCG of each plate as in pp_solve_resident.  Meshes
 * must carry grid_nx / grid_ny; uniform thickness per plate.
 * ---------------------------------------------------------------------- */

typedef struct pp_batch pp_batch;

typedef struct {
  int32_t workgroups_per_cu;  /* resident workgroups per CU of the persistent launch */
  int32_t tile_shape;         /* 0: 15x15-node tiles, 1: 31x7, 2: 7x31,
                                 3: per plate the shape with the fewest tiles */
  int32_t balance_iterations; /* 1: deal workgroups by measured iterations x tiles,
                                 0: by tiles only */
  int32_t resident_plates;    /* 1: a plate dealt one workgroup per tile keeps its
                                 state in registers and runs the single-reduction
                                 recurrences (kernel limited to 3 workgroups per CU),
                                 0: every plate strided, state in HBM (up to 4 per CU) */
} pp_batch_schedule;

typedef struct {
  int32_t plates, converged, workgroups, max_iterations;
  int64_t total_iterations, total_dof;
  double device_ms;               /* HIP-event time of the persistent launch */
  double worst_true_rel_residual; /* max over plates of |F - K x| / |F| (FP64) */
} pp_batch_stats;

int pp_batch_create(pp_ctx* ctx, int32_t n_plates, const pp_mesh* meshes,
                    const pp_material* materials, const uint8_t* const* constrained_masks,
                    pp_batch** out);
int pp_batch_destroy(pp_batch* batch);
int32_t pp_batch_plate_count(const pp_batch* batch);
/* Consistent nodal pressure load of one plate, as pp_pressure_load
 * (computed on the device; Q4 plates by the 2x2 Gauss rule, DKT plates by the
 * triangles' 3-point rule).  Replaces buildLoadVector + applyBCsToRHS
 * (solver.ts:363-407) for one plate of the batch. */
int pp_batch_pressure_load(pp_batch* batch, int32_t plate, double q0, const double* patches,
                           int32_t n_patches, float* F);
/* Every plate's pressure load at once, on the device, written straight into
 * the batch's right-hand sides (what pp_batch_pressure_load followed by
 * pp_batch_set_rhs does for each plate, without the host round trip).
 * q0: n_plates uniform pressures; plate p's Gaussian patches are
 * patches[4 * patch_offsets[p] .. 4 * patch_offsets[p + 1]) as (A, cx, cy, s);
 * patch_offsets has n_plates + 1 nondecreasing entries from 0.  The loop over
 * plates of solver.ts:363-407, batched. */
int pp_batch_pressure_loads(pp_batch* batch, const double* q0, const double* patches,
                            const int32_t* patch_offsets);
/* F of one plate (3 * node_count floats). */
int pp_batch_set_rhs(pp_batch* batch, int32_t plate, const float* F);
int pp_batch_set_schedule(pp_batch* batch, const pp_batch_schedule* schedule);
/* The schedule in use; wg_per_plate (may be NULL) receives each plate's workgroups. */
int pp_batch_get_schedule(const pp_batch* batch, pp_batch_schedule* schedule, int32_t* wg_per_plate);
/* Autotuner: measures every plate's iteration count once, then times each
 * candidate schedule (workgroups per CU x tile policy x balancing) on the
 * resident loads and keeps the fastest (also left applied). */
int pp_batch_autotune(pp_batch* batch, const pp_solve_opts* opts, pp_batch_schedule* best,
                      pp_batch_stats* best_stats, int32_t* trials);
/* One persistent launch over every plate (opts: tolerance, max_iterations, flags). */
int pp_batch_solve(pp_batch* batch, const pp_solve_opts* opts, pp_batch_stats* stats);
/* Solution and statistics of one plate after pp_batch_solve. */
int pp_batch_plate_result(pp_batch* batch, int32_t plate, double* x64, float* x32,
                          pp_solve_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* PLATEPCG_H */
