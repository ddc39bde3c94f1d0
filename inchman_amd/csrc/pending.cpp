// Entry points whose kernels are not in this build yet: they report
// GPGMP_E_UNSUPPORTED so that every symbol of include/gpgmp_hip.h is exported.
#include "host.h"

extern "C" {
#define GPGMP_PENDING(sig) \
    int sig { return GPGMP_E_UNSUPPORTED; }
GPGMP_PENDING(gpgmp_min_timestep_device(const float *, const float *, int32_t, float, float *, void *))
}
