// EAM sweep kernel instantiations (both precisions, both containers, NVT).
#include "ptmc_eam.hpp"
#include "ptmc_kernels.hpp"

namespace ptmc {

#define PTMC_EAM_PAIR(KL_) \
    KernelPair { &eam_kernel<Real, BC, KL_, LANES_SMALL>, &eam_init_kernel<Real, BC, KL_, LANES_SMALL> }

template <typename Real, int BC>
static KernelPair eam_kl(int kl)
{
    switch (kl) {
    case 1: return PTMC_EAM_PAIR(1);
    case 2: return PTMC_EAM_PAIR(2);
    case 4: return PTMC_EAM_PAIR(4);
    case 8: return PTMC_EAM_PAIR(8);
    case 16: return PTMC_EAM_PAIR(16);
    default: return {};
    }
}

KernelPair lookup_eam(int precision, int bc, int kl, int lanes)
{
    if (lanes != LANES_SMALL) return {};
    if (precision == 0) return bc == BC_SPHERE ? eam_kl<double, BC_SPHERE>(kl) : eam_kl<double, BC_CUBIC>(kl);
    return bc == BC_SPHERE ? eam_kl<float, BC_SPHERE>(kl) : eam_kl<float, BC_CUBIC>(kl);
}

}  // namespace ptmc
