// The potential plugin functions of src/EnergyEvaluation.jl as batched FP64
// device work: set_variables + initialise_energy (:1662-1768, :1778-1822)
// and energy_update! for one displaced atom (:1474-1621), for n_conf
// independent configurations of N atoms per call.
//
// The engine's own sweep kernels never call these: a move there updates the
// resident state in place.  They serve callers that drive the plugin
// interface one configuration at a time (the reference's MCState code and
// its tests) and want the same arithmetic on the device: pairs i < j in the
// reference's order with sequential sums, the reference pair form (r^-6
// then repeated division) when the handle asks for it, the EAM terms as
// (1 / sqrt(r^2))^n with pow, the network in FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "ptmc_types.hpp"

namespace ptmc {

constexpr int PLUGIN_MAXN = 4096;  // atoms per configuration
constexpr int PLUGIN_MAXH = 64;    // widest hidden layer of the network

struct PluginJob {
    int n_conf = 0, N = 0;
    int kind = 0;    // PotKind
    int pk = 0;      // PairKind of ELJ / ELJB / LUT
    int bc = 0;      // Boundary
    int cutoff = 0;  // NPT: pairs beyond the box's r_cut skipped, tail correction added
    int div = 1;     // minimum image with the reference's divisions (else reciprocal products)
    PotParams pot{};
    // network potential: FP64 tables as given to ptmc_set_potential_nnp
    int n_rad = 0, n_ang = 0, n_ang_built = 0, h1 = 0, h2 = 0;
    double r_cut = 0.0, hardcore2 = 0.0;
    const double *rad = nullptr;  // n_rad x 4: eta, norm, offset, active
    const double *ang = nullptr;  // n_ang x 6: eta, lambda, zeta, tpz, offset, active
    const double *w = nullptr;    // RuNNer parameter order, [nf, h1, h2, 1]
    // inputs (device)
    const double *x = nullptr;        // n_conf x N x 3 (null: distances from d2_in)
    const double *d2_in = nullptr;    // n_conf x N x N squared distances, or null
    const double *box = nullptr;      // n_conf x 2: length, height (null in a sphere)
    const int32_t *index = nullptr;   // update: the moved atom of each configuration
    const double *trial = nullptr;    // update: n_conf x 3 trial positions
    const double *en_in = nullptr;    // update: current total energies
    const double *vars_in = nullptr;  // update: EAM components (N x 2) / features (N x nf)
    // outputs (device; en_atom and, for EAM and the network, vars_out always present)
    double *en = nullptr, *en_atom = nullptr, *vars_out = nullptr;
    // workspace of plugin_workspace_doubles(job) doubles
    double *ws = nullptr;
};

// features per atom of the network potential
__host__ __device__ inline int plugin_nf(const PluginJob &j) { return j.n_rad + j.n_ang; }
size_t plugin_workspace_doubles(const PluginJob &j);
hipError_t plugin_initialise(const PluginJob &j, hipStream_t s);
hipError_t plugin_update(const PluginJob &j, hipStream_t s);

}  // namespace ptmc
