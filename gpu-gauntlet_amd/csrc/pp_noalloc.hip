// The FP64 structured solve on caller device buffers with no device
// allocation: every scratch buffer lives in one file-scope device array.
//
// Layout of the array (each part 256-byte aligned):
//   grid barrier words | device statistics | norm partials (kNormParts
//   doubles) | grid-reduction partials (kPartMax doubles) | three DOF
//   vectors (the persistent kernel's two exchange vectors and the residual
//   vector of the true-residual check).
// The host never uses the array's name as an address: the device address
// of this GPU's copy comes from hipGetSymbolAddress, once per device, when
// a context is created on it.
#include <string.h>

#include <mutex>

#include "pp_grid.h"

namespace pp {

namespace {

constexpr size_t kScratchBytes = 16u << 20;  // 16 MiB per GPU
constexpr size_t kPartMax = 8 * 2048;        // partials of up to 2,048 workgroups
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_init_mu;
unsigned char* g_scratch[kMaxDevices];        // device address of this GPU's array
std::mutex g_device_mu[kMaxDevices];          // one call at a time per GPU

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

constexpr size_t kOffGs = 0;
constexpr size_t kOffDs = kOffGs + align256(sizeof(GridSync));
constexpr size_t kOffNorm = kOffDs + align256(sizeof(DevStats));
constexpr size_t kOffPart = kOffNorm + align256(kNormParts * sizeof(double));
constexpr size_t kOffVec = kOffPart + align256(kPartMax * sizeof(double));

// Largest DOF count whose three vectors fit after the fixed parts.
constexpr int64_t kMaxDof = (int64_t)((kScratchBytes - kOffVec) / (3 * sizeof(double)));

}  // namespace

// Called by pp_ctx_create with `device` current.
int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(PP_ERR_INVALID, "device index too large");
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_scratch[device]) return PP_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(PP_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[device] = (unsigned char*)p;
  return PP_OK;
}

}  // namespace pp

using namespace pp;

extern "C" {

int pp_solve_device_noalloc(pp_plan* pl, const double* F, double* x, const pp_solve_opts* o,
                            pp_solve_stats* st) {
  if (!pl || !F || !x || !o || !st) return fail(PP_ERR_INVALID, "NULL argument");
  if (!pl->structured) return fail(PP_ERR_UNSUPPORTED, "pp_solve_device_noalloc needs a structured plan");
  if (o->precision != PP_PRECISION_FP64 || o->precond != PP_PRECOND_BLOCK_JACOBI ||
      (o->flags & (PP_FLAG_SMALL_TILES | PP_FLAG_CLASSIC_PCG)) ||
      (o->path != PP_PATH_AUTO && o->path != PP_PATH_STRUCTURED))
    return fail(PP_ERR_UNSUPPORTED,
                "pp_solve_device_noalloc runs FP64 block-Jacobi PCG on the structured path only");
  if (!(o->tolerance > 0) || o->max_iterations < 0)
    return fail(PP_ERR_INVALID, "bad tolerance/max_iterations");
  if (pl->n_dof > kMaxDof)
    return fail(PP_ERR_INVALID, "pp_solve_device_noalloc: " + std::to_string(pl->n_dof) +
                                    " DOFs exceed the scratch limit of " +
                                    std::to_string(kMaxDof) + " DOFs (16 MiB per GPU)");
  const int dev = pl->ctx->device;
  unsigned char* base = dev >= 0 && dev < kMaxDevices ? g_scratch[dev] : nullptr;
  if (!base) return fail(PP_ERR_INVALID, "scratch array not initialised on this device");
  PP_LOCK(pl->ctx);
  std::lock_guard<std::mutex> dlk(g_device_mu[dev]);
  memset(st, 0, sizeof(*st));
  PP_HIP(hipSetDevice(dev));
  size_t npart = 0;
  int rc = struct_partials_needed(pl, 0, &npart);
  if (rc) return rc;
  if (npart > kPartMax)
    return fail(PP_ERR_INVALID, "pp_solve_device_noalloc: too many workgroups for the scratch partials");
  const size_t nd = (size_t)pl->n_dof;
  double* vec = (double*)(base + kOffVec);
  StructWs ws{F,
              x,
              vec,
              vec + nd,
              vec + 2 * nd,
              (double*)(base + kOffPart),
              kPartMax,
              (double*)(base + kOffNorm),
              (GridSync*)(base + kOffGs),
              (DevStats*)(base + kOffDs)};
  return struct_solve_fp64_ws(pl, o, st, ws, true);
}

}  // extern "C"
