// k_gmp_step (and the hot-cell kernel) instantiated for MODE_HOM launches.  The extended
// instances: models with localized reactions, float literals or custom kinetic laws.
#define GPGMP_EXTENDED_LAWS 1
#include "stochastic_kernels.h"

namespace gpgmp {

hipError_t launch_gmp_hom_cl(const StepLaunch &a) { return launch_m<MODE_HOM>(a); }

}  // namespace gpgmp
