// Spread-layout kernels (pair potentials in Horner form): the spherical
// container in NVT, cubic / rectangular boxes in NVT and NPT, every built
// pair form and precision, for 64-, 128- and 256-lane workgroups.
#include "ptmc_kernels.hpp"
#include "ptmc_spread.hpp"

namespace ptmc {

template <typename Real, int PK, int KC, int BC, int ENS>
static KernelPair spread_lanes(int lanes)
{
    switch (lanes) {
    case 64: return {&spread_sweep_kernel<Real, PK, KC, BC, ENS, 64>, &spread_init_kernel<Real, PK, KC, BC, ENS, 64>};
    case 128:
        return {&spread_sweep_kernel<Real, PK, KC, BC, ENS, 128>, &spread_init_kernel<Real, PK, KC, BC, ENS, 128>};
    case 256:
        return {&spread_sweep_kernel<Real, PK, KC, BC, ENS, 256>, &spread_init_kernel<Real, PK, KC, BC, ENS, 256>};
    default: return {};
    }
}

template <typename Real, int PK, int KC>
static KernelPair spread_geometry(int lanes, int bc, int ens)
{
    if (bc == BC_SPHERE && ens == ENS_NVT) return spread_lanes<Real, PK, KC, BC_SPHERE, ENS_NVT>(lanes);
    if (bc == BC_CUBIC && ens == ENS_NVT) return spread_lanes<Real, PK, KC, BC_CUBIC, ENS_NVT>(lanes);
    if (bc == BC_CUBIC && ens == ENS_NPT) return spread_lanes<Real, PK, KC, BC_CUBIC, ENS_NPT>(lanes);
    return {};
}

KernelPair lookup_spread(int precision, int pk, int kc, int lanes, int bc, int ens)
{
    if (precision == 1) {
        if (pk == PK_LJ && kc == 4) return spread_geometry<float, PK_LJ, 4>(lanes, bc, ens);
        if (pk == PK_EVEN_H && kc == 6) return spread_geometry<float, PK_EVEN_H, 6>(lanes, bc, ens);
        if (pk == PK_EVEN_H && kc == 8) return spread_geometry<float, PK_EVEN_H, 8>(lanes, bc, ens);
        if (pk == PK_ODD_H && kc == 8) return spread_geometry<float, PK_ODD_H, 8>(lanes, bc, ens);
        return {};
    }
    if (pk == PK_EVEN_H && kc == 6) return spread_geometry<double, PK_EVEN_H, 6>(lanes, bc, ens);
    if (pk == PK_EVEN_H && kc == 8) return spread_geometry<double, PK_EVEN_H, 8>(lanes, bc, ens);
    if (pk == PK_ODD_H && kc == 8) return spread_geometry<double, PK_ODD_H, 8>(lanes, bc, ens);
    return {};
}

const void *spread_cycle_fn() { return (const void *)&spread_cycle_kernel; }

}  // namespace ptmc
