// k_gmp_step (and the hot-cell kernel) instantiated for MODE_INH_X launches.
#include "stochastic_kernels.h"

namespace gpgmp {

hipError_t launch_gmp_inh_x(const StepLaunch &a) { return launch_m<MODE_INH_X>(a); }

}  // namespace gpgmp
