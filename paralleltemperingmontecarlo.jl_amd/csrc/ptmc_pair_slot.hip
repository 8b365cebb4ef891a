// Pair-slot kernels (one workgroup per configuration; see ptmc_pair_slot.hpp)
// for the binary32 pair forms: Lennard-Jones and the even extended LJ with 6
// or 8 coefficients.
#include "ptmc_pair_slot.hpp"

namespace ptmc {

using SweepFn = void (*)(SweepArgs);

template <int PK, int KC>
static void pair_slot_pick(bool da, SweepFn *sweep, SweepFn *init)
{
    *sweep = da ? &pair_slot_kernel<true, PK, KC> : &pair_slot_kernel<false, PK, KC>;
    *init = &pair_slot_init_kernel<PK, KC>;
}

// false when the pair form has no pair-slot kernel
bool pair_slot_fns(int pk, int kc, bool da, SweepFn *sweep, SweepFn *init)
{
    if (pk == PK_LJ && kc == 4) pair_slot_pick<PK_LJ, 4>(da, sweep, init);
    else if (pk == PK_EVEN_H && kc == 6) pair_slot_pick<PK_EVEN_H, 6>(da, sweep, init);
    else if (pk == PK_EVEN_H && kc == 8) pair_slot_pick<PK_EVEN_H, 8>(da, sweep, init);
    else return false;
    return true;
}

size_t pair_slot_lds_bytes(bool da) { return pair_slot_lds_floats(da) * sizeof(float); }

}  // namespace ptmc
