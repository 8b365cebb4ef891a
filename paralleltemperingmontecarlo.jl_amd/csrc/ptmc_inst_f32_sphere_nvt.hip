// Sweep kernel instantiations: f32, sphere_nvt.
#include "ptmc_sweep.hpp"
#include "ptmc_kernels.hpp"
PTMC_DEFINE_LOOKUP_F32(lookup_f32_sphere_nvt, BC_SPHERE, ENS_NVT)
