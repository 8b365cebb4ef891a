// Spread-layout kernels (spherical container, NVT, pair potentials): every
// built pair form and precision for 64-, 128- and 256-lane workgroups.
#include "ptmc_kernels.hpp"
#include "ptmc_spread.hpp"

namespace ptmc {

template <typename Real, int PK, int KC>
static KernelPair spread_lanes(int lanes)
{
    switch (lanes) {
    case 64: return {&spread_sweep_kernel<Real, PK, KC, 64>, &spread_init_kernel<Real, PK, KC, 64>};
    case 128: return {&spread_sweep_kernel<Real, PK, KC, 128>, &spread_init_kernel<Real, PK, KC, 128>};
    case 256: return {&spread_sweep_kernel<Real, PK, KC, 256>, &spread_init_kernel<Real, PK, KC, 256>};
    default: return {};
    }
}

KernelPair lookup_spread(int precision, int pk, int kc, int lanes)
{
    if (precision == 1) {
        if (pk == PK_LJ && kc == 4) return spread_lanes<float, PK_LJ, 4>(lanes);
        if (pk == PK_EVEN_H && kc == 6) return spread_lanes<float, PK_EVEN_H, 6>(lanes);
        if (pk == PK_EVEN_H && kc == 8) return spread_lanes<float, PK_EVEN_H, 8>(lanes);
        if (pk == PK_ODD_H && kc == 8) return spread_lanes<float, PK_ODD_H, 8>(lanes);
        return {};
    }
    if (pk == PK_EVEN_H && kc == 6) return spread_lanes<double, PK_EVEN_H, 6>(lanes);
    if (pk == PK_EVEN_H && kc == 8) return spread_lanes<double, PK_EVEN_H, 8>(lanes);
    if (pk == PK_ODD_H && kc == 8) return spread_lanes<double, PK_ODD_H, 8>(lanes);
    return {};
}

const void *spread_cycle_fn() { return (const void *)&spread_cycle_kernel; }

}  // namespace ptmc
