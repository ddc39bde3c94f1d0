// k_gmp_step (and the hot-cell kernel) instantiated for MODE_INH_X launches.  The extended
// instances: models with localized reactions, float literals or custom kinetic laws.
#define GPGMP_EXTENDED_LAWS 1
#include "stochastic_kernels.h"

namespace gpgmp {

hipError_t launch_gmp_inh_x_cl(const StepLaunch &a) { return launch_m<MODE_INH_X>(a); }

}  // namespace gpgmp
