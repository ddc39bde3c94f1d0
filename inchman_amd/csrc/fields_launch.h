// Host-side launch interface of the drift/diffusivity method kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "law.h"

namespace gpgmp {

hipError_t launch_drift_fields(const uint32_t *code, const uint32_t *prog, const LawExtra &ex, float pmw,
                               float pmh, float *out, hipStream_t stream);

}  // namespace gpgmp
