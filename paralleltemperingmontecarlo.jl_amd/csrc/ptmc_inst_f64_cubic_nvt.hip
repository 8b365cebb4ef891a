// Sweep kernel instantiations: f64, cubic_nvt.
#include "ptmc_sweep.hpp"
#include "ptmc_kernels.hpp"
PTMC_DEFINE_LOOKUP_F64(lookup_f64_cubic_nvt, BC_CUBIC, ENS_NVT)
