// NNVT kernels of the two-species network potential with binary32
// positions, features and networks (FP64 energy sums).
#define N2R float
#define N2_SWEEP_FN nnp2_sweep_fn_f32
#define N2_INIT_FN nnp2_init_fn_f32
// two workgroups per CU: at most 128 VGPRs
#define N2_SWEEP_ATTR __attribute__((amdgpu_waves_per_eu(4)))
#include "ptmc_nnp2_mc.hpp"
