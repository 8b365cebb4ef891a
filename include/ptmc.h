/*
 * ptmc.h — C ABI of the MI355X parallel-tempering engine (libptmc.so).
 *
 * Drop-in boundary for the hot path of ParallelTemperingMonteCarlo.jl:
 * single-atom displacement moves with incremental energy updates, Metropolis
 * acceptance and temperature-ladder replica exchange.  A handle holds any
 * number of independent PT runs (one temperature ladder each) that share one
 * potential and one ensemble; all of them advance together on one GPU.
 *
 * Conventions: plain pointers and sizes, caller-owned host buffers that are
 * copied in or out, integer status codes (PTMC_OK = 0), no exceptions across
 * the ABI.  A handle is used from one host thread at a time; different
 * handles are independent.  Arrays indexed by temperature follow the
 * reference's MCState vector order (temperature position 1..R, here 0..R-1).
 */
#ifndef PTMC_H
#define PTMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PTMC_OK 0
#define PTMC_E_ARG 1         /* invalid argument */
#define PTMC_E_STATE 2       /* call not valid in the handle's state */
#define PTMC_E_DEVICE 3      /* HIP runtime error */
#define PTMC_E_SPHERE 4      /* "too many moves out of binding sphere" (src/MCMoves.jl:132) */
#define PTMC_E_UNSUPPORTED 5 /* combination not built */
#define PTMC_E_NO_DEVICE 6   /* no usable GPU */

typedef struct ptmc_engine *ptmc_handle;

/* Engine-wide settings.  Counterparts: MCParams (src/InputParams.jl:35-63),
 * the ensemble (src/Ensembles.jl:42-86) and the boundary condition type
 * (src/BoundaryConditions.jl:44-116). */
typedef struct {
    int32_t precision;         /* 0: FP64 everywhere; 1: binary32 positions and dE, FP64 totals */
    int32_t ensemble;          /* 0: NVT, 1: NPT (uniform volume moves), 2: NPT with separated
                                  volume moves, z 1/6, xy 1/3, xyz 1/2 of them (NPT(n, p, true),
                                  src/MCMoves.jl:301-343; rectangular and rhombic boxes);
                                  3: NNVT, N atom moves and n_swaps atom swaps per cycle
                                  (NNVT, src/Ensembles.jl:138-150, :270-282; with
                                  ptmc_set_potential_nnp2 only) */
    int32_t boundary;          /* 0: SphericalBC, 1: CubicBC, 2: RectangularBC (L, L, H),
                                  3: RhombicBC (60 degree prism, side L, height H),
                                  src/BoundaryConditions.jl:80-116 */
    int32_t exchange_mode;     /* 0: reference (p = 0.1 per cycle, one random adjacent pair,
                                      src/MCRun.jl:203, src/Exchange.jl:135);
                                  1: one random adjacent pair every cycle;
                                  2: all even/odd adjacent pairs, alternating cycles */
    int32_t pair_form;         /* 0: Horner in 1/r^2 (fast); 1: the reference's loop with
                                  repeated division (src/EnergyEvaluation.jl:923-942), FP64 only */
    int32_t n_adjust;          /* MCParams.n_adjust (default 100) */
    int32_t mc_sample;         /* MCParams.mc_sample (default 1) */
    int32_t n_bin;             /* MCParams.n_bin (default 100) */
    double min_acc, max_acc;   /* MCParams.min_acc/max_acc (0.4, 0.6) */
    int32_t recompute_every;   /* full FP64 energy recompute every this many cycles (0 = never) */
    int32_t lanes_per_replica; /* 0: chosen per atom count; else 1, 2, 4, 8 or 16 (32 or 64 in the
                                  spread layout, workgroup_lanes > 0) */
    int32_t cycles_per_launch; /* cycles per kernel launch (0 = 256) */
    int32_t device;            /* HIP device ordinal */
    uint64_t seed;             /* Philox key */
    int32_t workgroup_lanes;   /* 0: every ladder in one workgroup, cycles_per_launch cycles per
                                  launch with the exchange inside the kernel; 64, 128 or 256:
                                  ladders spread over workgroups of this many lanes, one cycle per
                                  launch and a per-ladder exchange kernel (pair potentials in
                                  the Horner form: SphericalBC in NVT, CubicBC or RectangularBC
                                  in NVT or NPT with uniform volume moves); same results */
    int32_t spread_streams;    /* spread layout: the runs are cut into this many independent
                                  groups launched on their own streams, so that one group's
                                  cycle overlaps another's (0 or 1: one group); same results */
    int32_t exchange_every;    /* exchange cadence: replica exchange is attempted at the end of
                                  the cycles with (cycle + 1) % exchange_every == 0 only, as
                                  exchange_mode says (0 or 1: every cycle, the reference's
                                  src/MCRun.jl:203); the spread layout then runs the cycles
                                  between exchanges in one launch */
    int32_t pair_slots;        /* 1: pair potentials (LJ, even extended LJ) one workgroup per
                                  configuration of up to 64 atoms in binary32 with FP64 totals,
                                  spherical container, NVT: the layout of delayed acceptance
                                  with the learned surrogate (ptmc_surrogate_init) */
} ptmc_config;

/* One PT run.  Counterparts: TempGrid (src/InputParams.jl:77-95), the start
 * configuration(s) given to ptmc_run! (src/MCRun.jl:363-376,
 * src/Initialization.jl:87-95 takes start_config[mod1(i, n)]), and the
 * boundary condition / pressure. */
typedef struct {
    int32_t n_atoms;
    int32_t n_temps;
    double radius;             /* SphericalBC radius (bohr) */
    double box_length;         /* CubicBC side / RectangularBC, RhombicBC length (bohr) */
    double pressure;           /* NPT pressure (Hartree / bohr^3) */
    const double *beta;        /* n_temps inverse temperatures (1/Hartree) */
    const double *x0;          /* n_configs * n_atoms * 3 start positions (bohr) */
    int32_t n_configs;         /* start configuration of temperature i is i mod n_configs */
    double max_displ_atom;     /* initial atom step (reference default 0.1, src/MCStates.jl:104) */
    double max_displ_vol;      /* initial volume step (reference default 0.1) */
    int32_t lanes_per_replica; /* 0: engine default */
    uint32_t run_key;          /* Philox counter word for this run; 0xffffffff = run index */
    double box_height;         /* RectangularBC, RhombicBC height (bohr) */
} ptmc_run_desc;

/* Lifecycle. */
int ptmc_create(const ptmc_config *cfg, ptmc_handle *out);
int ptmc_destroy(ptmc_handle h);
int ptmc_last_error(ptmc_handle h, char *buf, size_t len);

/* Potentials — the plugin API of src/EnergyEvaluation.jl:70-76:
 * ELJPotentialEven / ELJPotential (:138-153, :891-908) and
 * EmbeddedAtomPotential (:1059-1073, coefficients as stored there). */
int ptmc_set_potential_elj(ptmc_handle h, int32_t even_powers, int32_t n_coeff, const double *coeff);
/* EAM (NVT, spherical or cubic): E = sum_i ean phi_i - ecam sqrt(rho_i),
 * phi_i = sum_j r_ij^-n, rho_i = sum_j r_ij^-m; the per-atom phi/rho stay
 * resident next to the positions and are updated per move as
 * energy_update! does (src/EnergyEvaluation.jl:1567-1591). */
int ptmc_set_potential_eam(ptmc_handle h, double n, double m, double ean, double ecam);
/* ELJPotentialB, the extended Lennard-Jones potential in a strong magnetic
 * field (src/EnergyEvaluation.jl:161-182, :943-958, tail correction
 * :983-998): e(r, t) = sum_k c_k r^-p_k (1 + a_k cos 2t + b_k cos 4t), p = 6,
 * 8, 9, 10, ..., t the angle of the pair vector to z, 0.1 inside r^2 < 5.30;
 * 1 to 8 coefficients of each.  Every boundary and ensemble of the pair
 * engine. */
int ptmc_set_potential_eljb(ptmc_handle h, int32_t n_coeff, const double *coeff_c, const double *coeff_a,
                            const double *coeff_b);
/* LookupTablePotential (src/EnergyEvaluation.jl:211-295): pair energies
 * tabulated over the angle to z (n_angle rows, d_angle degrees apart from 0;
 * the lookup uses rows 1..91) and the distance (n_dist columns from
 * start_dist, d_dist apart), table[angle][dist] row-major; nearest-entry
 * lookup, c6coeff / r^6 beyond the table, tail correction from c6coeff
 * (:999-1008).  read_lookuptable (:229-263) gives the arguments. */
int ptmc_set_potential_lookup(ptmc_handle h, int32_t n_dist, int32_t n_angle, double start_dist, double d_dist,
                              double d_angle, double c6coeff, const double *table);
/* RuNNer-style network potential (RuNNerPotential, src/EnergyEvaluation.jl:1182-1213;
 * symmetry functions MachineLearningPotential/SymmFunc.jl:24-112): rad holds
 * n_rad rows (eta, G_norm, G_offset, active) of RadialType2, ang n_ang rows
 * (eta, lambda, zeta, tpz, G_offset, active) of AngularType3 (active = type
 * 11 / 111); feature rows are the radial functions then the angular ones.
 * hardcore2 is RuNNerPotential.boundary (squared; 0 = none), q10_compat builds
 * the initial features as total_symm_calc does (SymmFunc.jl:504).  The
 * network is [n_rad + n_ang, h1, h2, 1] with activations [*, 2 (tanh),
 * 2 (tanh), 1 (linear)] and RuNNer parameter order (NeuralNetworkPotential,
 * ForwardPass.jl:23-39).  Binary32 features and network, FP64 totals;
 * NVT in a spherical container, at most 128 atoms (runs of up to 64 atoms
 * share a CU two workgroups at a time; larger ones take a CU each; the
 * delayed-acceptance surrogate covers up to 64) and 128 functions, zeta a
 * power of two up to 16, lambda = +-1, at most 32 groups of functions with
 * equal eta (one function per (lambda, zeta) in a group). */
int ptmc_set_potential_nnp(ptmc_handle h, int32_t n_rad, int32_t n_ang, const double *rad, const double *ang,
                           double r_cut, double hardcore2, int32_t q10_compat, int32_t n_layers,
                           const int32_t *nodes, const int32_t *activations, int32_t n_params,
                           const double *params);

/* Two-species RuNNer network potential in the NNVT ensemble
 * (RuNNerPotential2Atom, src/EnergyEvaluation.jl:1287-1317; moves
 * :1622-1643, swaps :1438-1448, src/MCMoves.jl:363-370; commit
 * src/swap_config.jl:238-277).  Atoms 0..n1-1 of every run are species 1,
 * n1..n1+n2-1 species 2 (every run holds n1 + n2 <= 64 atoms); a cycle
 * draws N + n_swaps moves, each an atom displacement or, for indices
 * beyond N, a swap of one atom of each species, accepted with
 * min(1, exp(-beta dE)).  rad / ang: the tables of ptmc_nnp2_energy;
 * params1 / params2: the networks of species 1 / 2, shape [2 n_rad +
 * 3 n_ang, h1, h2, 1], activations 1 (linear) or 2 (tanh) per layer.
 * Precision 0: FP64 throughout; 1: binary32 positions, distances, features
 * and networks with FP64 energy sums and Metropolis test.  Spherical
 * container.  The volume counters of
 * ptmc_get_counters hold (swaps accepted, swaps attempted) per temperature.
 * No counterpart of the reference for the split of a cycle's draws: the
 * reference's move index, atom and species picks come from Julia's RNG,
 * here from Philox (as for every other move). */
int ptmc_set_potential_nnp2(ptmc_handle h, int32_t n1, int32_t n2, int32_t n_swaps, int32_t n_rad, int32_t n_ang,
                            const double *rad, const double *ang, double r_cut, int32_t n_layers,
                            const int32_t *nodes, const int32_t *activations, int32_t n_params,
                            const double *params1, const double *params2);

/* Runs.  Replaces initialisation (src/Initialization.jl:70-101) and
 * MCState construction (src/MCStates.jl:137-162): energies of the start
 * configurations are computed on the device in FP64. */
int ptmc_add_run(ptmc_handle h, const ptmc_run_desc *desc, int32_t *run_index);
int ptmc_prepare(ptmc_handle h);

/* Cycles.  ptmc_cycles enqueues n_cycles Monte-Carlo cycles of every run
 * (mc_cycle!, src/MCRun.jl:192-253) on the engine's stream and waits for
 * them; device-side errors surface at ptmc_sync.
 * ptmc_start_phase restarts the phase-local cycle index used by n_adjust and
 * mc_sample (equilibration cycles count from 1, src/MCRun.jl:296-298, and so
 * do production cycles, :397); production = 1 turns on sampling_step!. */
int ptmc_start_phase(ptmc_handle h, int32_t production);
int ptmc_cycles(ptmc_handle h, int64_t n_cycles);
int ptmc_sync(ptmc_handle h);

/* Histogram range of a run (initialise_histograms!, src/Sampling.jl:141-250):
 * delta_en = (en_max - en_min) / (n_bin - 1), delta_v = (v_max - v_min) / n_bin.
 * Clears the run's histograms. */
int ptmc_set_histogram(ptmc_handle h, int32_t run, double en_min, double en_max, double v_min,
                       double v_max);
/* reset_counters (src/MCRun.jl:270-276) for every run. */
int ptmc_reset_counters(ptmc_handle h);

/* Read-back, temperature order.  box: n_temps lengths, or n_temps (length,
 * height) pairs for rectangular and rhombic boxes. */
int ptmc_get_energies(ptmc_handle h, int32_t run, double *en);
int ptmc_get_state(ptmc_handle h, int32_t run, double *x, double *en, double *box);
int ptmc_get_counters(ptmc_handle h, int32_t run, double *max_displ, int64_t *count_atom,
                      int64_t *count_vol, int64_t *count_exc, double *ham);
int ptmc_get_histograms(ptmc_handle h, int32_t run, double *hist); /* n_temps * hist_stride */
int ptmc_hist_stride(ptmc_handle h);                                /* n_bin+2 or (n_bin+2)^2 */
int ptmc_get_drift(ptmc_handle h, int32_t run, double *drift);     /* max |E_run - E_FP64| */

/* Write-back of a run's state (checkpoint restart, src/ReadSave.jl:436-481);
 * box as ptmc_get_state returns it. */
int ptmc_set_state(ptmc_handle h, int32_t run, const double *x, const double *box,
                   const double *max_displ, const int64_t *count_atom, const int64_t *count_vol,
                   const int64_t *count_exc, const double *ham);

/* Separated volume moves (NPT(n, p, true)): per temperature the xy and z
 * steps (max_displ[3], [4] of MCState, src/MCStates.jl:104) and the accepted
 * xy / z move counters (count_vol_xy, count_vol_z: total, window;
 * src/swap_config.jl:142-148), n_temps x 2 each, temperature order; the
 * uniform volume moves stay in ptmc_get_counters' count_vol. */
int ptmc_get_counters_sep(ptmc_handle h, int32_t run, double *max_displ_sep, int64_t *count_vol_xy,
                          int64_t *count_vol_z);
int ptmc_set_counters_sep(ptmc_handle h, int32_t run, const double *max_displ_sep, const int64_t *count_vol_xy,
                          const int64_t *count_vol_z);

/* Histogram counts of a run (restart: setresults, src/ReadSave.jl:385-417);
 * n_temps * hist_stride non-negative integral values. */
int ptmc_set_histograms(ptmc_handle h, int32_t run, const double *hist);
/* Position in the random stream: the next cycle draws Philox counter word
 * `cycle` and has phase-local index phase_cycle + 1 (n_adjust / mc_sample
 * cadence).  A run resumed from a checkpoint continues the stream it would
 * have drawn without the interruption. */
int ptmc_seek(ptmc_handle h, int64_t cycle, int64_t phase_cycle);

/* Per-move trace of the next n_cycles cycles (parity tests). */
int ptmc_set_trace(ptmc_handle h, int32_t n_cycles);
int ptmc_get_trace(ptmc_handle h, int32_t run, double *de, int8_t *acc, int32_t *idx);
/* Delayed acceptance: the surrogate energy change E_hat' - E_hat of each
 * traced move ([temperature][cycle][move], NaN where none was evaluated). */
int ptmc_get_trace_aux(ptmc_handle h, int32_t run, double *aux);

/* Introspection for benchmarks: number of workgroups, kernel time (HIP
 * events on the engine's stream) accumulated since the last reset. */
int ptmc_plan_info(ptmc_handle h, int32_t *n_tiles, int32_t *n_launch_groups, int64_t *n_lanes);
int ptmc_run_lanes(ptmc_handle h, int32_t run);
int ptmc_kernel_time(ptmc_handle h, double *total_ms, int64_t *n_launches, int32_t reset);
/* Launch groups (one kernel instantiation each; groups run concurrently):
 * lanes per configuration, workgroup width, workgroups, and the summed
 * duration and count of the group's timed launches (events on the group's
 * stream; every launch in the ladder layout, one cycle's launches in four in
 * the spread layout).  ptmc_kernel_time's n_launches counts every launch.
 * ptmc_run_group returns the group a run was planned into (-1 on error). */
int ptmc_group_info(ptmc_handle h, int32_t group, int32_t *lanes_per_replica, int32_t *workgroup_lanes,
                    int32_t *n_tiles, double *total_ms, int64_t *n_launches);
int ptmc_run_group(ptmc_handle h, int32_t run);

/* Multihistogram heat capacity, multihistogram_NVT (src/multihist_NVT.jl:95-221),
 * FP64 on the device for n_runs independent runs at once.  hist holds the
 * sampled energy histograms (n_temps x (n_bin + 2): underflow, n_bin bins,
 * overflow) of each run; beta_result / t_result the n_result-point output
 * grid (the reference uses TempGrid{10 n_temps}); k_boltzmann the constant of
 * the Cv formula (the reference uses 3.166811429e-6, src/multihist_NVT.jl:18).
 * Outputs: cv (n_runs x n_result), free energies (n_runs x n_temps, may be
 * NULL) and fixed-point iterations used (n_runs, may be NULL). */
int ptmc_multihistogram_nvt(int32_t n_runs, int32_t n_temps, int32_t n_bin, const double *beta,
                            const double *hist, const double *en_min, const double *en_max,
                            int32_t n_result, const double *beta_result, const double *t_result,
                            double k_boltzmann, double conv_threshold, int32_t max_iter, double *cv,
                            double *free_energy, int32_t *iterations, int32_t device);

/* Multihistogram heat capacity at constant pressure, multihistogram_NPT
 * (src/multihist_NPT.jl:126-340), FP64 on the device for n_runs runs at once.
 * hist holds each run's n_temps (n_bin + 2) x (n_bin + 2) histograms, rows
 * energy and columns volume (underflow, n_bin bins, overflow), row-major as
 * ptmc_get_histograms returns them; en_min/en_max/v_min/v_max the histogram
 * ranges and pressure the run's pressure (n_runs each).  The bins weigh
 * exp(-beta (E + pV)); Cp = (<h^2> - <h>^2) / (k T^2) with h = E + pV.
 * Outputs: cp (n_runs x n_result); moments (n_runs x n_result x 3: <E>, <V>,
 * <h>, may be NULL); free energies (n_runs x n_temps, may be NULL);
 * iterations (n_runs, may be NULL). */
int ptmc_multihistogram_npt(int32_t n_runs, int32_t n_temps, int32_t n_bin, const double *beta,
                            const double *hist, const double *en_min, const double *en_max,
                            const double *v_min, const double *v_max, const double *pressure,
                            int32_t n_result, const double *beta_result, const double *t_result,
                            double k_boltzmann, double conv_threshold, int32_t max_iter, double *cp,
                            double *moments, double *free_energy, int32_t *iterations, int32_t device);

/* ptmc_multihistogram_npt for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc); it has no counterpart in the reference.  Same arguments and
 * results, but every array is a device pointer the caller owns, free_energy
 * and iterations may be NULL, the kernel is queued on `stream` (a
 * hipStream_t, NULL for the default stream) and the call returns without
 * waiting.  The per-bin scratch is a device-resident block of
 * PTMC_NOALLOC_NPT_SCRATCH_DOUBLES doubles (16 MiB) per GPU, held until the
 * process ends; calls needing more than 3 n_runs n_bin^2 of it return
 * PTMC_E_UNSUPPORTED.  The call allocates nothing; calls that overlap on one
 * GPU share the block and race. */
#define PTMC_NOALLOC_NPT_SCRATCH_DOUBLES (1 << 21)
int ptmc_multihistogram_npt_noalloc(int32_t n_runs, int32_t n_temps, int32_t n_bin, const double *beta,
                                    const double *hist, const double *en_min, const double *en_max,
                                    const double *v_min, const double *v_max, const double *pressure,
                                    int32_t n_result, const double *beta_result, const double *t_result,
                                    double k_boltzmann, double conv_threshold, int32_t max_iter, double *cp,
                                    double *moments, double *free_energy, int32_t *iterations, int32_t device,
                                    void *stream);

/* Two-species RuNNer potential energies (RuNNerPotential2Atom,
 * src/EnergyEvaluation.jl:1287-1317, initialise_energy :1748-1768, features
 * MachineLearningPotential/SymmFunc.jl:313-547), FP64 on the device for
 * n_conf configurations of n1 atoms of species 1 then n2 of species 2 (at
 * most 64 atoms, spherical container: no periodic images).  rad: n_rad rows
 * (eta, norm[4], offset[4]) of RadialType2a, ang: n_ang rows (eta, lambda,
 * zeta, tpz[6], offset[6]) of AngularType3a (norm / tpz / offset ordered
 * species 1 rows then species 2 rows); features per atom 2 n_rad + 3 n_ang
 * (at most 128) = nodes[0]; params1 / params2 the species' networks in the
 * RuNNer layout (hidden widths up to 64).  Outputs energy (n_conf) and, if
 * not NULL, eatom (n_conf x (n1 + n2)). */
int ptmc_nnp2_energy(int32_t n_conf, int32_t n1, int32_t n2, const double *x, int32_t n_rad, int32_t n_ang,
                     const double *rad, const double *ang, double r_cut, int32_t n_layers, const int32_t *nodes,
                     const int32_t *activations, int32_t n_params, const double *params1, const double *params2,
                     double *energy, double *eatom, int32_t device);

/* The potential plugin functions, batched over n_conf configurations of
 * n_atoms atoms that the caller holds (the engine's runs are not touched),
 * with the handle's potential, boundary (ptmc_config.boundary) and ensemble,
 * in FP64 on the device.  Counterparts: set_variables
 * (src/EnergyEvaluation.jl:1778-1822), initialise_energy (:1662-1747) and
 * energy_update! (:1474-1621) for ELJPotentialEven / ELJPotential,
 * ELJPotentialB, LookupTablePotential, EmbeddedAtomPotential and
 * RuNNerPotential (spherical container).  The arithmetic is the
 * reference's: pairs i < j in order with sequential sums, the reference pair
 * form (r^-6 then repeated division) and minimum image with divisions when
 * ptmc_config.pair_form = 1 (Horner in 1/r^2 and reciprocal products
 * otherwise), EAM terms (1 / sqrt(r^2))^n by pow, the network in FP64.
 *
 *   x      n_conf x n_atoms x 3 positions (bohr); may be NULL for
 *          initialise_energy of a distance-only potential (ELJ, EAM) when d2
 *          is given
 *   d2     NULL, or n_conf x n_atoms x n_atoms squared distances (the
 *          reference's dist2_mat, get_distance2_mat src/Configurations.jl:
 *          135-162); only for ELJ and EAM.  NULL: built from x as
 *          get_distance2_mat does (pair i < j as distance2(x_i, x_j), mirrored)
 *   box    n_conf x 2 (box length, box height) in a periodic box (the height
 *          of a cubic box is not read); NULL in a spherical container.  NPT
 *          (ensemble 1, 2) applies the box's cut-off r_cut (get_r_cut,
 *          src/Ensembles.jl:118-155) and tail correction (lrc, :968-1008;
 *          none for ELJPotential with odd powers, PTMC_E_UNSUPPORTED)
 *   vars   the potential variables, ptmc_potential_vars values per
 *          configuration: EAM component_vector (n_atoms x 2: phi, rho),
 *          network g_matrix (atom-major n_atoms x (n_rad + n_ang), i.e. the
 *          reference's column-major features x atoms); none for pair
 *          potentials.
 *
 * ptmc_initialise_energy writes en_tot (n_conf) and, if not NULL, en_atom
 * (n_conf x n_atoms: the dimer rows of dimer_energy_config, the network's
 * per-atom energies, the EAM terms ean phi_i - ecam sqrt(rho_i)) and vars.
 * ptmc_energy_update moves atom index[c] of configuration c to trial[c]
 * (n_conf x 3) and writes new_en = the reference's new_en (dimer: the row
 * difference + en_tot[c]; EAM: calc_energies_from_components of the new
 * components; network: the new features' energy, or 100.0 when a partner is
 * inside the hard-core distance, src/EnergyEvaluation.jl:1601-1606), and,
 * for EAM and the network, new_vars and new_en_atom (a configuration that
 * failed the hard-core test gets new_vars = vars and keeps its new_en_atom
 * row as passed in).  Pair potentials keep no variables: en_tot is the
 * current energy, vars / new_vars / new_en_atom are not used. */
int ptmc_potential_vars(ptmc_handle h, int32_t n_atoms, int64_t *n_vars);
int ptmc_initialise_energy(ptmc_handle h, int32_t n_conf, int32_t n_atoms, const double *x, const double *d2,
                           const double *box, double *en_tot, double *en_atom, double *vars);
int ptmc_energy_update(ptmc_handle h, int32_t n_conf, int32_t n_atoms, const double *x, const double *d2,
                       const double *box, const int32_t *index, const double *trial, const double *en_tot,
                       const double *vars, double *new_en, double *new_en_atom, double *new_vars);

/* Learned surrogate and delayed-acceptance Metropolis for the network
 * potential, or for a pair potential in pair slots (ptmc_config.pair_slots).
 * No counterpart in the reference: these entry points serve the engine's own
 * acceleration (Christen & Fox 2005).  The surrogate is a per-atom MLP on 64
 * Gaussian radial descriptors between 2 bohr and the cutoff (the network
 * potential's; 11.338 bohr for pair potentials) (64 -> 64 -> 64 SiLU -> 1),
 * trained on the device on (configuration, exact energy) samples the run
 * collects.  With delayed acceptance on, a move is first tested on the
 * surrogate energy change and only a surviving move is evaluated with the
 * exact potential and accepted with min(1, exp(-beta (dE - dE_hat))), so
 * the sampled distribution is the exact one.  Parameters are 8385 values:
 * W1 [64][64], b1 [64], W2 [64][64], b2 [64], w3 [64], b3. */
int ptmc_surrogate_init(ptmc_handle h, int32_t sample_capacity, uint64_t seed);
int ptmc_surrogate_collect(ptmc_handle h, int32_t every_cycles);          /* 0 = stop */
int ptmc_surrogate_samples(ptmc_handle h, int64_t *n_samples);
int ptmc_surrogate_train(ptmc_handle h, int32_t epochs, int32_t batch, double learning_rate,
                         double valid_fraction, double *rmse_train, double *rmse_valid);
int ptmc_surrogate_enable(ptmc_handle h, int32_t delayed_acceptance);
int ptmc_surrogate_get_params(ptmc_handle h, double *params);
int ptmc_surrogate_set_params(ptmc_handle h, const double *params);
/* Surrogate energies of n configurations (n x n_atoms x 3), and the training
 * loss mean_c ((E_hat_c - E_c) / n_atoms)^2 with its gradient. */
int ptmc_surrogate_predict(ptmc_handle h, int32_t n, int32_t n_atoms, const double *x, double *e_hat);
int ptmc_surrogate_grad(ptmc_handle h, int32_t n, int32_t n_atoms, const double *x, const double *e,
                        double *loss, double *grad);
/* Delayed-acceptance counts of a run: proposals, surrogate-stage accepts
 * (exact evaluations), final accepts. */
int ptmc_surrogate_stats(ptmc_handle h, int32_t run, int64_t *proposals, int64_t *stage1, int64_t *accepted);

/* The RuNNer forward pass with the argument list of the reference's ccall
 * (MachineLearningPotential/ForwardPass.jl:60-84, :121-145; the prebuilt
 * librunnerjulia.so symbol `forward`), so ForwardPass.lib_path() can point at
 * this library: input is column-major n_in x batch, params in RuNNer order
 * (per layer: weights input-major, then biases), activations 1 (linear) or
 * 2 (tanh) per layer (the input layer's code is not used).  Writes the
 * per-atom energies to eatom[batch] and returns eatom[batch - 1]; returns NaN
 * for an invalid description.  Computed on the GPU in FP64. */
double forward(const double *input, const int32_t *n_in, const int32_t *n_layers, const int32_t *nodes,
               const int32_t *batch, const int32_t *n_params, const double *params, const int32_t *activations,
               double *eatom);

#ifdef __cplusplus
}
#endif
#endif
