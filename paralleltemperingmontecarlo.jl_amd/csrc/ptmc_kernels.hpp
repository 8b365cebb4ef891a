// Kernel lookup for the sweep kernel instantiations.  Each translation unit
// ptmc_inst_<precision>_<bc>.hip instantiates one precision / boundary /
// ensemble family; the engine resolves a pointer at plan time.
#pragma once
#include "ptmc_types.hpp"

namespace ptmc {

using SweepFn = void (*)(SweepArgs);

// precision: 0 FP64, 1 binary32.  Returns nullptr if not built.
SweepFn lookup_f64_sphere_nvt(int pk, int kc, int kl);
SweepFn lookup_f64_cubic_nvt(int pk, int kc, int kl);
SweepFn lookup_f64_cubic_npt(int pk, int kc, int kl);
SweepFn lookup_f32_sphere_nvt(int pk, int kc, int kl);
SweepFn lookup_f32_cubic_nvt(int pk, int kc, int kl);
SweepFn lookup_f32_cubic_npt(int pk, int kc, int kl);

}  // namespace ptmc

// Instantiation helper used by the ptmc_inst_*.hip files.
#define PTMC_KL_SWITCH(REAL, PK, KC, BC, ENS, kl)                                 \
    switch (kl) {                                                                 \
    case 1: return &sweep_kernel<REAL, PK, KC, BC, ENS, 1>;                       \
    case 2: return &sweep_kernel<REAL, PK, KC, BC, ENS, 2>;                       \
    case 4: return &sweep_kernel<REAL, PK, KC, BC, ENS, 4>;                       \
    case 8: return &sweep_kernel<REAL, PK, KC, BC, ENS, 8>;                       \
    case 16: return &sweep_kernel<REAL, PK, KC, BC, ENS, 16>;                     \
    default: return nullptr;                                                      \
    }

#define PTMC_KL_SWITCH_SMALL(REAL, PK, KC, BC, ENS, kl)                           \
    switch (kl) {                                                                 \
    case 1: return &sweep_kernel<REAL, PK, KC, BC, ENS, 1>;                       \
    case 2: return &sweep_kernel<REAL, PK, KC, BC, ENS, 2>;                       \
    case 4: return &sweep_kernel<REAL, PK, KC, BC, ENS, 4>;                       \
    default: return nullptr;                                                      \
    }

#define PTMC_DEFINE_LOOKUP_F64(NAME, BC, ENS)                                     \
    namespace ptmc {                                                              \
    SweepFn NAME(int pk, int kc, int kl)                                          \
    {                                                                             \
        if (pk == PK_EVEN_H && kc == 6) { PTMC_KL_SWITCH(double, PK_EVEN_H, 6, BC, ENS, kl) } \
        if (pk == PK_EVEN_H && kc == 8) { PTMC_KL_SWITCH(double, PK_EVEN_H, 8, BC, ENS, kl) } \
        if (pk == PK_ODD_H && kc == 8) { PTMC_KL_SWITCH(double, PK_ODD_H, 8, BC, ENS, kl) }   \
        if (pk == PK_EVEN_R && kc == 8) { PTMC_KL_SWITCH_SMALL(double, PK_EVEN_R, 8, BC, ENS, kl) } \
        if (pk == PK_ODD_R && kc == 8) { PTMC_KL_SWITCH_SMALL(double, PK_ODD_R, 8, BC, ENS, kl) }   \
        return nullptr;                                                           \
    }                                                                             \
    }

#define PTMC_DEFINE_LOOKUP_F32(NAME, BC, ENS)                                     \
    namespace ptmc {                                                              \
    SweepFn NAME(int pk, int kc, int kl)                                          \
    {                                                                             \
        if (pk == PK_LJ && kc == 4) { PTMC_KL_SWITCH(float, PK_LJ, 4, BC, ENS, kl) }         \
        if (pk == PK_EVEN_H && kc == 6) { PTMC_KL_SWITCH(float, PK_EVEN_H, 6, BC, ENS, kl) } \
        if (pk == PK_EVEN_H && kc == 8) { PTMC_KL_SWITCH(float, PK_EVEN_H, 8, BC, ENS, kl) } \
        if (pk == PK_ODD_H && kc == 8) { PTMC_KL_SWITCH(float, PK_ODD_H, 8, BC, ENS, kl) }   \
        return nullptr;                                                           \
    }                                                                             \
    }
