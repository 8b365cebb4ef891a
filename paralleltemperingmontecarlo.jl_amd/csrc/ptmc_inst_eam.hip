// EAM sweep kernel instantiations (both precisions, both containers, NVT).
#include "ptmc_eam.hpp"
#include "ptmc_kernels.hpp"

namespace ptmc {

template <typename Real, int BC>
static SweepFn eam_kl(int kl)
{
    switch (kl) {
    case 1: return &eam_kernel<Real, BC, 1, LANES_SMALL>;
    case 2: return &eam_kernel<Real, BC, 2, LANES_SMALL>;
    case 4: return &eam_kernel<Real, BC, 4, LANES_SMALL>;
    case 8: return &eam_kernel<Real, BC, 8, LANES_SMALL>;
    case 16: return &eam_kernel<Real, BC, 16, LANES_SMALL>;
    default: return nullptr;
    }
}

SweepFn lookup_eam(int precision, int bc, int kl, int lanes)
{
    if (lanes != LANES_SMALL) return nullptr;
    if (precision == 0) return bc == BC_SPHERE ? eam_kl<double, BC_SPHERE>(kl) : eam_kl<double, BC_CUBIC>(kl);
    return bc == BC_SPHERE ? eam_kl<float, BC_SPHERE>(kl) : eam_kl<float, BC_CUBIC>(kl);
}

}  // namespace ptmc
