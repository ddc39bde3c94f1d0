// k_gmp_step instantiated for MODE_INH_XS launches (X sweeps without an SSA).  The extended
// instances: models with localized reactions, float literals or custom kinetic laws.
#define GPGMP_EXTENDED_LAWS 1
#include "stochastic_kernels.h"

namespace gpgmp {

hipError_t launch_gmp_inh_xs_cl(const StepLaunch &a) { return launch_m<MODE_INH_XS>(a); }

}  // namespace gpgmp
