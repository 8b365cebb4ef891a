// Neural-network potential kernels (one translation unit; see ptmc_nnp.hpp).
#include "ptmc_nnp.hpp"

namespace ptmc {

using SweepFn = void (*)(SweepArgs);

SweepFn nnp_sweep_fn() { return &nnp_sweep_kernel<false>; }
SweepFn nnp_sweep_da_fn() { return &nnp_sweep_kernel<true>; }
SweepFn nnp_init_fn() { return &nnp_init_kernel; }
void (*nnp_cycle_fn())(SweepArgs, int) { return &nnp_cycle_kernel; }

}  // namespace ptmc
