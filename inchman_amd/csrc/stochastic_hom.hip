// k_gmp_step (and the hot-cell kernel) instantiated for MODE_HOM launches.
#include "stochastic_kernels.h"

namespace gpgmp {

hipError_t launch_gmp_hom(const StepLaunch &a) { return launch_m<MODE_HOM>(a); }

}  // namespace gpgmp
