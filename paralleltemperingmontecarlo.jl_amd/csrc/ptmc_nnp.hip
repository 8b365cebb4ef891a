// Neural-network potential kernels (one translation unit; see ptmc_nnp.hpp).
#include "ptmc_nnp.hpp"

namespace ptmc {

using SweepFn = void (*)(SweepArgs);

// np: atoms per slot, NNP_NP (64) or NNP_NP_WIDE (128)
SweepFn nnp_sweep_fn(int np) { return np > NNP_NP ? &nnp_sweep_kernel<false, NNP_NP_WIDE> : &nnp_sweep_kernel<false, NNP_NP>; }
SweepFn nnp_sweep_da_fn() { return &nnp_sweep_kernel<true, NNP_NP>; }
SweepFn nnp_init_fn(int np) { return np > NNP_NP ? &nnp_init_kernel<NNP_NP_WIDE> : &nnp_init_kernel<NNP_NP>; }
void (*nnp_cycle_fn())(SweepArgs, int) { return &nnp_cycle_kernel; }

}  // namespace ptmc
