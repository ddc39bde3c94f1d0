// k_gmp_step instantiated for MODE_INH_XS launches (X sweeps without an SSA).
#include "stochastic_kernels.h"

namespace gpgmp {

hipError_t launch_gmp_inh_xs(const StepLaunch &a) { return launch_m<MODE_INH_XS>(a); }

}  // namespace gpgmp
