// Sweep kernel instantiations: f32, cubic_nvt.
#include "ptmc_sweep.hpp"
#include "ptmc_kernels.hpp"
PTMC_DEFINE_LOOKUP_F32(lookup_f32_cubic_nvt, BC_CUBIC, ENS_NVT)
