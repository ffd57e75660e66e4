// speedy_adjoint.hip -- fused forward + VJP of the SPEEDY column physics (gfx950).
// Work in progress: the entry point reports that the adjoint is not built yet.
#include <hip/hip_runtime.h>

#include <string>

#include "speedy_abi.h"

namespace speedy {
int fail(int code, const std::string& msg);
}

extern "C" int speedy_fwd_vjp(const SpeedyFields*, const SpeedyAdjoint*, const SpeedyVertical*, const SpeedyParams*,
                              const SpeedyDate*, int32_t, int32_t, int32_t, uint32_t, int32_t, void*, size_t, void*) {
  return speedy::fail(SPEEDY_ERR_ARG, "speedy_fwd_vjp: the adjoint kernel is not built in this version");
}
