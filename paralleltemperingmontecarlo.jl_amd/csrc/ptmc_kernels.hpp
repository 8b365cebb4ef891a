// Kernel lookup for the sweep kernel instantiations.  Each translation unit
// ptmc_inst_<precision>_<bc>.hip instantiates one precision / boundary /
// ensemble family; the engine resolves a pointer at plan time.
#pragma once
#include "ptmc_sweep.hpp"
#include "ptmc_types.hpp"

namespace ptmc {

using SweepFn = void (*)(SweepArgs);

// The cycle kernel of a configuration and its initial-energy kernel.
struct KernelPair {
    SweepFn sweep = nullptr;
    SweepFn init = nullptr;
};

// Workgroup widths the kernels are built for (lanes = threads per workgroup).
constexpr int LANES_SMALL = 256;
constexpr int LANES_HALF = 512;
constexpr int LANES_MID = 768;
constexpr int LANES_LARGE = 1024;

// Returns null kernels if the combination is not built.
KernelPair lookup_f64_sphere_nvt(int pk, int kc, int kl, int lanes);
KernelPair lookup_f64_cubic_nvt(int pk, int kc, int kl, int lanes);
KernelPair lookup_f64_cubic_npt(int pk, int kc, int kl, int lanes);
KernelPair lookup_f32_sphere_nvt(int pk, int kc, int kl, int lanes);
KernelPair lookup_f32_cubic_nvt(int pk, int kc, int kl, int lanes);
KernelPair lookup_f32_cubic_npt(int pk, int kc, int kl, int lanes);
// rhombic prism (rectangular boxes use the cubic kernels)
KernelPair lookup_f64_rhomb_nvt(int pk, int kc, int kl, int lanes);
KernelPair lookup_f64_rhomb_npt(int pk, int kc, int kl, int lanes);
KernelPair lookup_f32_rhomb_nvt(int pk, int kc, int kl, int lanes);
KernelPair lookup_f32_rhomb_npt(int pk, int kc, int kl, int lanes);
// Sutton-Chen EAM, NVT, 256-lane workgroups
KernelPair lookup_eam(int precision, int bc, int kl, int lanes);
// spread layout (spherical container in NVT; cubic / rectangular boxes, bc
// BC_CUBIC, in NVT and NPT): 64-, 128- or 256-lane workgroups, lanes per
// configuration read from each tile; and its per-ladder cycle kernel
// multi_cycle: the variant that runs several cycles between exchanges
KernelPair lookup_spread(int precision, int pk, int kc, int lanes, int bc, int ens, bool multi_cycle);
const void *spread_cycle_fn();

}  // namespace ptmc

namespace ptmc {
// 1024-lane workgroups only where the kernel fits 128 VGPRs without spilling
// (binary32, spherical container, NVT).  768-lane workgroups hold a
// 48-temperature ladder of a periodic box at 8 or 16 lanes per configuration:
// with 256-atom boxes the slab of 48 configurations (147 KB in binary32) fits
// LDS only when the tile has exactly 48 slots.
#define PTMC_PAIR(KL_, W_) \
    KernelPair { &sweep_kernel<Real, PK, KC, BC, ENS, KL_, W_>, &init_kernel<Real, PK, KC, BC, ENS, KL_, W_> }

template <typename Real, int PK, int KC, int BC, int ENS>
KernelPair kl_switch(int kl, int lanes)
{
    if (lanes == 256) {
        switch (kl) {
        case 1: return PTMC_PAIR(1, 256);
        case 2: return PTMC_PAIR(2, 256);
        case 4: return PTMC_PAIR(4, 256);
        case 8: return PTMC_PAIR(8, 256);
        case 16: return PTMC_PAIR(16, 256);
        default: return {};
        }
    }
    if constexpr (sizeof(Real) == 4 && (BC == BC_CUBIC || BC == BC_RHOMB)) {
        if (lanes == 768) {
            switch (kl) {
            case 8: return PTMC_PAIR(8, 768);
            case 16: return PTMC_PAIR(16, 768);
            default: return {};
            }
        }
    }
    if constexpr (sizeof(Real) == 8 && BC == BC_SPHERE && ENS == ENS_NVT) {
        // a 32-temperature FP64 ladder at 16 lanes per configuration: two
        // waves per SIMD hide the latency of the moves' dependent FP64 chains
        if (lanes == 512) {
            switch (kl) {
            case 8: return PTMC_PAIR(8, 512);
            case 16: return PTMC_PAIR(16, 512);
            default: return {};
            }
        }
    }
    if constexpr (sizeof(Real) == 4 && BC == BC_SPHERE && ENS == ENS_NVT) {
        if (lanes == 1024) {
            switch (kl) {
            case 4: return PTMC_PAIR(4, 1024);
            case 8: return PTMC_PAIR(8, 1024);
            case 16: return PTMC_PAIR(16, 1024);
            default: return {};
            }
        }
    }
    return {};
}

// reference pair forms (parity runs): one to four lanes, 256-lane workgroups
template <typename Real, int PK, int KC, int BC, int ENS>
KernelPair kl_switch_small(int kl, int lanes)
{
    if (lanes != 256) return {};
    switch (kl) {
    case 1: return PTMC_PAIR(1, 256);
    case 2: return PTMC_PAIR(2, 256);
    case 4: return PTMC_PAIR(4, 256);
    default: return {};
    }
}
}  // namespace ptmc

#define PTMC_KL_SWITCH(REAL, PK, KC, BC, ENS, kl, lanes) return kl_switch<REAL, PK, KC, BC, ENS>(kl, lanes);
#define PTMC_KL_SWITCH_SMALL(REAL, PK, KC, BC, ENS, kl, lanes) return kl_switch_small<REAL, PK, KC, BC, ENS>(kl, lanes);

#define PTMC_DEFINE_LOOKUP_F64(NAME, BC, ENS)                                                        \
    namespace ptmc {                                                                                 \
    KernelPair NAME(int pk, int kc, int kl, int lanes)                                                  \
    {                                                                                                \
        if (pk == PK_EVEN_H && kc == 6) { PTMC_KL_SWITCH(double, PK_EVEN_H, 6, BC, ENS, kl, lanes) }   \
        if (pk == PK_EVEN_H && kc == 8) { PTMC_KL_SWITCH(double, PK_EVEN_H, 8, BC, ENS, kl, lanes) }   \
        if (pk == PK_ODD_H && kc == 8) { PTMC_KL_SWITCH(double, PK_ODD_H, 8, BC, ENS, kl, lanes) }     \
        if (pk == PK_EVEN_R && kc == 8) { PTMC_KL_SWITCH_SMALL(double, PK_EVEN_R, 8, BC, ENS, kl, lanes) } \
        if (pk == PK_ODD_R && kc == 8) { PTMC_KL_SWITCH_SMALL(double, PK_ODD_R, 8, BC, ENS, kl, lanes) }   \
        if (pk == PK_ELJB && kc == 8) { PTMC_KL_SWITCH(double, PK_ELJB, 8, BC, ENS, kl, lanes) }       \
        if (pk == PK_LUT && kc == 1) { PTMC_KL_SWITCH(double, PK_LUT, 1, BC, ENS, kl, lanes) }         \
        return {};                                                                                   \
    }                                                                                                \
    }

#define PTMC_DEFINE_LOOKUP_F32(NAME, BC, ENS)                                                        \
    namespace ptmc {                                                                                 \
    KernelPair NAME(int pk, int kc, int kl, int lanes)                                                  \
    {                                                                                                \
        if (pk == PK_LJ && kc == 4) { PTMC_KL_SWITCH(float, PK_LJ, 4, BC, ENS, kl, lanes) }            \
        if (pk == PK_EVEN_H && kc == 6) { PTMC_KL_SWITCH(float, PK_EVEN_H, 6, BC, ENS, kl, lanes) }    \
        if (pk == PK_EVEN_H && kc == 8) { PTMC_KL_SWITCH(float, PK_EVEN_H, 8, BC, ENS, kl, lanes) }    \
        if (pk == PK_ODD_H && kc == 8) { PTMC_KL_SWITCH(float, PK_ODD_H, 8, BC, ENS, kl, lanes) }      \
        if (pk == PK_ELJB && kc == 8) { PTMC_KL_SWITCH(float, PK_ELJB, 8, BC, ENS, kl, lanes) }        \
        if (pk == PK_LUT && kc == 1) { PTMC_KL_SWITCH(float, PK_LUT, 1, BC, ENS, kl, lanes) }          \
        return {};                                                                                   \
    }                                                                                                \
    }
