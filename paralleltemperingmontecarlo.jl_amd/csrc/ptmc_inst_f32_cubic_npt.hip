// Sweep kernel instantiations: f32, cubic_npt.
#include "ptmc_sweep.hpp"
#include "ptmc_kernels.hpp"
PTMC_DEFINE_LOOKUP_F32(lookup_f32_cubic_npt, BC_CUBIC, ENS_NPT)
