// Sweep kernel instantiations: f64, sphere_nvt.
#include "ptmc_sweep.hpp"
#include "ptmc_kernels.hpp"
PTMC_DEFINE_LOOKUP_F64(lookup_f64_sphere_nvt, BC_SPHERE, ENS_NVT)
