// Mixed precision (FP32 inner PCG, FP64 refinement) on the structured path.
#include "pp_device.h"

namespace pp {

int struct_solve_mixed(pp_plan* pl, const pp_solve_opts* o, pp_solve_stats* st) {
  (void)pl; (void)o; (void)st;
  return fail(PP_ERR_UNSUPPORTED, "mixed precision not built yet");
}

}  // namespace pp
