// Sweep kernel instantiations: f32, rhomb_nvt.
#include "ptmc_sweep.hpp"
#include "ptmc_kernels.hpp"
PTMC_DEFINE_LOOKUP_F32(lookup_f32_rhomb_nvt, BC_RHOMB, ENS_NVT)
