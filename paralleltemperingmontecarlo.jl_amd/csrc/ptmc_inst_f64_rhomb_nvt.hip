// Sweep kernel instantiations: f64, rhomb_nvt.
#include "ptmc_sweep.hpp"
#include "ptmc_kernels.hpp"
PTMC_DEFINE_LOOKUP_F64(lookup_f64_rhomb_nvt, BC_RHOMB, ENS_NVT)
