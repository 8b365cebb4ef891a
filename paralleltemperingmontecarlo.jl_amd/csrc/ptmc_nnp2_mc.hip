// NNVT kernels of the two-species network potential in FP64.
#define N2R double
#define N2_SWEEP_FN nnp2_sweep_fn
#define N2_INIT_FN nnp2_init_fn
#define N2_SWEEP_ATTR 
#include "ptmc_nnp2_mc.hpp"
