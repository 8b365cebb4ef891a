// Sweep kernel instantiations: f32, rhomb_npt.
#include "ptmc_sweep.hpp"
#include "ptmc_kernels.hpp"
PTMC_DEFINE_LOOKUP_F32(lookup_f32_rhomb_npt, BC_RHOMB, ENS_NPT)
