// Plain structs shared by the host engine and the sweep kernels.
#pragma once
#include <stdint.h>

namespace ptmc {

enum Ensemble : int32_t { ENS_NVT = 0, ENS_NPT = 1 };
// Boundary conditions (src/BoundaryConditions.jl:44-116).  Rectangular boxes
// (L, L, H) run the cubic kernels with the box kind set at run time; the
// rhombic prism has kernels of its own.
enum Boundary : int32_t { BC_SPHERE = 0, BC_CUBIC = 1, BC_RECT = 2, BC_RHOMB = 3 };
enum PotKind : int32_t {
    POT_ELJ_EVEN = 0,  // sum_k c_k r^-(4+2k), src/EnergyEvaluation.jl:138-153, :934-942
    POT_ELJ = 1,       // sum_k c_k r^-(5+k),  src/EnergyEvaluation.jl:891-933
    POT_EAM = 2,       // Sutton-Chen,          src/EnergyEvaluation.jl:1059-1162
    POT_NNP = 3,       // RuNNer network,       src/EnergyEvaluation.jl:1182-1275
    POT_ELJB = 4,      // ELJ in a magnetic field (anisotropic in the angle to z),
                       //                       src/EnergyEvaluation.jl:161-182, :943-958
    POT_LUT = 5,       // tabulated over distance and angle to z, src/EnergyEvaluation.jl:211-295
    POT_NNP2 = 6,      // two-species RuNNer network in the NNVT ensemble,
                       // src/EnergyEvaluation.jl:1287-1433, src/Ensembles.jl:138-150
};
// How a pair energy is evaluated on the device.
enum PairForm : int32_t {
    FORM_HORNER = 0,     // Horner in 1/r^2 (fast path)
    FORM_REFERENCE = 1,  // r^-6 then repeated division, as src/EnergyEvaluation.jl:934-942
    FORM_LJ = 2,         // two-term c6 u^3 + c12 u^6 (binary32 only)
};
enum ExchangeMode : int32_t { EXC_REFERENCE = 0, EXC_EVERY = 1, EXC_EVEN_ODD = 2 };

constexpr int MAX_COEFF = 16;
// longest ladder of the spread layout (its exchange kernel stages a ladder in LDS)
constexpr int SPREAD_MAX_TEMPS = 1024;

struct PotParams {
    int32_t kind;
    int32_t ncoeff;
    double coeff[MAX_COEFF];
    double eam_n, eam_m, eam_ean, eam_ecam;
    double coeff_a[8], coeff_b[8];  // ELJPotentialB: coeff holds c, these a and b
    // LookupTablePotential: table [n_angle][n_dist] in FP64 and binary32
    int32_t lut_ndist, lut_nangle;
    double lut_start, lut_d, lut_dangle, lut_c6;
    const double *lut64;
    const float *lut32;
};

// One PT run (one temperature ladder).
struct RunInfo {
    int32_t n_atoms;
    int32_t n_temps;
    int32_t temp_base;   // first index into the per-temperature arrays
    int32_t slot_base;   // first global slot (a slot holds one configuration)
    int32_t local_base;  // first slot inside its workgroup
    uint32_t run_key;    // Philox counter word 3
    double radius2;      // spherical container radius^2
    double pressure;
    double en_min, delta_en;  // energy histogram (src/Sampling.jl:141-158)
    double v_min, delta_v;    // volume histogram (NPT)
};

// One workgroup of the sweep kernel: slots of several runs with equal N.
struct TileInfo {
    int32_t slot0;      // first global slot
    int32_t n_slots;    // slots used (groups of KL lanes)
    int32_t n_atoms;
    int32_t n_slab;     // atoms per lane = ceil(N / KL)
    int64_t coord_off;  // element offset of the tile's [3][n_slab][WGS] block
    int64_t comp_off;   // EAM: element offset of the tile's [n_slab][2][WGS] phi/rho block
    int32_t kl;         // lanes per configuration (spread layout: read at run time)
    int32_t pad_;
};

// Neural-network potential tables (device memory), prepared on the host:
// active angular functions sorted into groups of equal eta (at most 8 per
// group), network weights in RuNNer order.
struct NnpTables {
    int32_t n_rad, n_ang, nf;      // radial, angular, total feature rows
    int32_t h1, h2;                // hidden widths (input nf, output 1)
    int32_t n_rad_active;          // radial rows that contribute
    int32_t n_groups;              // angular eta groups
    float r_cut;
    double hardcore2;              // squared hard-core distance (0: none)
    const float *rad;              // n_rad_active x 4: eta, norm, row, pad
    const int32_t *grp;            // n_groups x 2: first function, count
    const float *grp_eta;          // n_groups
    const float *afn;              // functions x 4: lambda, zeta, tpz, row
    const int32_t *abuild;         // functions: 1 if built by the full evaluation
    const float *row_off;          // nf: offset added by the full evaluation
    const int32_t *gdesc;          // n_groups x NNP_GD: per-group descriptor of the move update
    const int32_t *wgroups;        // per wave of the move update: count, then up to NNP_WG groups
    const float *w1, *b1, *w2, *b2, *w3;
    float b3;
};

// Two-species network potential with atom swaps (one workgroup per slot,
// FP64 throughout).  Atoms 0..n1-1 of every configuration are species 1.
struct Nnp2Tables {
    int32_t n1, n_swaps;       // species-1 atoms; atom swaps per cycle
    int32_t n_rad, n_ang, ng;  // functions; feature rows 2 n_rad + 3 n_ang
    int32_t h1, h2;            // hidden widths (input ng, output 1)
    int32_t act1, act2, act3;  // activation of each layer: 1 linear, 2 tanh
    double r_cut;
    const double *rad;         // n_rad x 9: eta, norm[4], offset[4]
    const double *ang;         // n_ang x 15: eta, lambda, zeta, tpz[6], offset[6]
    int32_t n_groups;          // angular functions grouped by eta
    const int32_t *grp;        // n_groups x 2: first entry in gfn, count
    const int32_t *gfn;        // n_ang angular function indices, group by group
    const double *p1, *p2;     // networks of species 1 and 2 (RuNNer layout)
    double *feat;              // per slot [NNP_NP][ng] features, atom-major
    double *eatom;             // per slot [NNP_NP] per-atom energies
};

struct SweepArgs {
    const TileInfo *tiles;
    const RunInfo *runs;
    const int32_t *slot_run;   // run of each slot
    int32_t *slot_temp;        // temperature index held by each slot
    double *slot_en;           // total energy (FP64) per slot
    double *slot_box;          // box length per slot
    double *slot_boxh;         // box height per slot (cubic: equal to the length)
    double *slot_drift;        // max |E_running - E_recomputed| per slot
    void *coords;              // per-tile [3][n_slab][WGS] (Real)
    void *comps;               // EAM: per-tile [n_slab][2][WGS] (Real)
    const double *t_beta;      // per temperature
    double *t_displ;           // per temperature x2 (atom, volume)
    const double *t_maxbox;    // per temperature
    const double *t_maxboxh;   // per temperature: max box height
    const double *t_lh;        // per temperature: length / height ratio bound
    double *t_displ_sep;       // per temperature x2: xy and z volume steps (separated moves)
    long long *t_count_volxy;  // per temperature x2
    long long *t_count_volz;   // per temperature x2
    long long *t_count_atom;   // per temperature x2 (total, window)
    long long *t_count_vol;    // per temperature x2
    long long *t_count_exc;    // per temperature x2 (attempted, accepted)
    double *t_ham;             // per temperature x2
    uint32_t *t_hist;          // per temperature x hist_stride
    int32_t *run_error;        // per run
    double *trace_de;          // optional [temp][cycle][move]
    int8_t *trace_acc;
    int32_t *trace_idx;
    double *trace_aux;         // optional: delayed acceptance's surrogate change per traced move
    const int32_t *run_list;   // spread layout: runs of the launch group (cycle kernel blocks)
    int32_t recompute_done;    // spread layout: the recompute due at sweep0 was applied already
    PotParams pot;
    uint64_t seed;
    int64_t sweep0;            // global cycle counter of the first cycle
    int64_t cycle0;            // phase-local cycle index minus one
    int32_t n_cycles;
    int32_t ensemble;
    int32_t exchange_mode;
    int32_t exchange_every;    // exchange attempted after cycles with (sweep + 1) % exchange_every == 0
    int32_t n_adjust;
    int32_t mc_sample;
    int32_t n_bin;
    int32_t hist_stride;
    int32_t sampling;
    int32_t recompute_every;
    int32_t trace_cycles;
    int32_t trace_cycle;       // one-cycle launches: index of this cycle in the trace
    int32_t init_run;          // init kernels: run whose energies are initialised (-1: all)
    int32_t box_kind;          // BC_CUBIC or BC_RECT for the cubic kernels
    int32_t separated;         // NPT: separated xy / z / xyz volume moves
    // neural-network potential (one workgroup per slot)
    NnpTables nnp;
    float *feat;               // per slot [nf][NNP_NP] features
    float *eatom;              // per slot [NNP_NP] per-atom energies
    int32_t *temp_slot;        // per temperature: slot holding it
    Nnp2Tables nnp2;
    // learned surrogate and delayed acceptance
    const float *sur_w;        // surrogate parameters (f32, packed)
    int32_t da;                // 1: delayed-acceptance sweeps
    float sur_rcut;            // descriptor cutoff (bohr)
    long long *da_stats;       // per slot x3: proposals, stage-1 accepted, accepted
    double min_acc, max_acc;
};

}  // namespace ptmc
