// Spread-layout kernels (pair potentials in Horner form): the spherical
// container in NVT, cubic / rectangular boxes in NVT and NPT, every built
// pair form and precision, for 64-, 128- and 256-lane workgroups.
#include "ptmc_kernels.hpp"
#include "ptmc_spread.hpp"

namespace ptmc {

template <typename Real, int PK, int KC, int BC, int ENS, int W>
static KernelPair spread_pair(bool mc)
{
    return {mc ? &spread_sweep_kernel<Real, PK, KC, BC, ENS, W, true> : &spread_sweep_kernel<Real, PK, KC, BC, ENS, W, false>,
            &spread_init_kernel<Real, PK, KC, BC, ENS, W>};
}

template <typename Real, int PK, int KC, int BC, int ENS>
static KernelPair spread_lanes(int lanes, bool mc)
{
    switch (lanes) {
    case 64: return spread_pair<Real, PK, KC, BC, ENS, 64>(mc);
    case 128: return spread_pair<Real, PK, KC, BC, ENS, 128>(mc);
    case 256: return spread_pair<Real, PK, KC, BC, ENS, 256>(mc);
    default: return {};
    }
}

template <typename Real, int PK, int KC>
static KernelPair spread_geometry(int lanes, int bc, int ens, bool mc)
{
    if (bc == BC_SPHERE && ens == ENS_NVT) return spread_lanes<Real, PK, KC, BC_SPHERE, ENS_NVT>(lanes, mc);
    if (bc == BC_CUBIC && ens == ENS_NVT) return spread_lanes<Real, PK, KC, BC_CUBIC, ENS_NVT>(lanes, mc);
    if (bc == BC_CUBIC && ens == ENS_NPT) return spread_lanes<Real, PK, KC, BC_CUBIC, ENS_NPT>(lanes, mc);
    return {};
}

KernelPair lookup_spread(int precision, int pk, int kc, int lanes, int bc, int ens, bool multi_cycle)
{
    const bool mc = multi_cycle;
    if (precision == 1) {
        if (pk == PK_LJ && kc == 4) return spread_geometry<float, PK_LJ, 4>(lanes, bc, ens, mc);
        if (pk == PK_EVEN_H && kc == 6) return spread_geometry<float, PK_EVEN_H, 6>(lanes, bc, ens, mc);
        if (pk == PK_EVEN_H && kc == 8) return spread_geometry<float, PK_EVEN_H, 8>(lanes, bc, ens, mc);
        if (pk == PK_ODD_H && kc == 8) return spread_geometry<float, PK_ODD_H, 8>(lanes, bc, ens, mc);
        return {};
    }
    if (pk == PK_EVEN_H && kc == 6) return spread_geometry<double, PK_EVEN_H, 6>(lanes, bc, ens, mc);
    if (pk == PK_EVEN_H && kc == 8) return spread_geometry<double, PK_EVEN_H, 8>(lanes, bc, ens, mc);
    if (pk == PK_ODD_H && kc == 8) return spread_geometry<double, PK_ODD_H, 8>(lanes, bc, ens, mc);
    return {};
}

const void *spread_cycle_fn() { return (const void *)&spread_cycle_kernel; }

}  // namespace ptmc
