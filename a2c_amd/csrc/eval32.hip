// Mixed-precision objective + gradient (FP32 bulk, FP64 per-site and global
// reductions).  Placeholder until the table kernels land.
#include "fe_internal.hpp"

namespace a2c {
int build_tables(a2c_fe_ctx* c) { return set_err(c, A2C_ESTATE, "mixed precision not available yet"); }
int launch_eval32(a2c_fe_ctx* c, const double*, double*, double*, bool, Timing*) {
  return set_err(c, A2C_ESTATE, "mixed precision not available yet");
}
}  // namespace a2c
