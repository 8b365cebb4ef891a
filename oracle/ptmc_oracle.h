/*
 * ORACLE — test infrastructure only (see orc_math.h).  CPU FP64 restatement
 * of the reference's parallel-tempering sweep for pair and EAM potentials.
 *
 * One call advances one ladder (one PT run) by n_sweeps Monte-Carlo cycles
 * exactly as src/MCRun.jl:192-253 does: N (NVT) or N+1 (NPT) moves per
 * replica, then at most one exchange attempt, then step-size adaptation,
 * then sampling.  All random numbers come from the Philox stream defined in
 * orc_math.h.
 */
#ifndef PTMC_ORACLE_H
#define PTMC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_NVT = 0, ORC_NPT = 1 };
enum { ORC_SPHERE = 0, ORC_CUBIC = 1, ORC_RECT = 2, ORC_RHOMB = 3 };
enum { ORC_POT_ELJ_EVEN = 0, ORC_POT_ELJ = 1, ORC_POT_EAM = 2, ORC_POT_NNP = 3, ORC_POT_ELJB = 4, ORC_POT_LUT = 5, ORC_POT_NNP2 = 6 };
/* exchange_mode: 0 = reference (p = 0.1 per cycle, one random adjacent pair),
 *                1 = one random adjacent pair every cycle,
 *                2 = all even (cycle even) / odd (cycle odd) pairs every cycle */
enum { ORC_EXC_REFERENCE = 0, ORC_EXC_EVERY = 1, ORC_EXC_EVEN_ODD = 2 };

typedef struct {
    int32_t n_atoms, n_temps;
    int32_t ensemble, bc;
    int32_t pair_form;    /* 0: reference form (r^-6 then repeated division), 1: Horner in 1/r^2 */
    int32_t dist_cache;   /* 1: cached r^2 matrix as the reference; 0: recompute from positions */
    int32_t lanes;        /* summation tree width of the device engine (1 = sequential) */
    int32_t fp32;         /* 1: positions and trial moves in binary32, energies in FP64 */
    int32_t exchange_mode;
    int32_t n_adjust, mc_sample, n_bin;
    int32_t sampling;     /* 1: production cycles (sampling_step! runs) */
    int32_t math_lib;     /* 1: use libm exp/log/pow (reference form), 0: deterministic exp/log */
    double  min_acc, max_acc;
    double  radius2;      /* spherical container radius^2 */
    double  pressure;
    uint64_t seed;
    uint32_t run;
    int64_t sweep0;       /* global cycle counter of the first cycle (RNG counter) */
    int64_t cycle0;       /* phase-local index of the first cycle minus one */
    double  en_min, delta_en_hist, v_min, delta_v_hist;
    int32_t separated;    /* NPT: separated xy / z / xyz volume moves (src/MCMoves.jl:301-343) */
    /* planted mistake in the atom-move Metropolis decision, for tests that
     * show a decision check catches it: 0 none, 1 beta doubled, 2 exponent
     * sign flipped, 3 uniform taken as 1 - u */
    int32_t plant_decision;
    /* exchange cadence: attempts only at the end of cycles with
     * (sweep + 1) % exchange_every == 0 (0 or 1: every cycle) */
    int32_t exchange_every;
} orc_params;

/* RuNNer-style neural-network potential (src/EnergyEvaluation.jl:1182-1275,
 * MachineLearningPotential/).  Symmetry functions:
 *   rad: n_rad rows of (eta, g_norm, g_offset, active)      RadialType2, type 11
 *   ang: n_ang rows of (eta, lambda, zeta, tpz, g_offset, active)  AngularType3, type 111
 * with tpz = 2^(1-zeta) g_norm.  The feature matrix G is stored atom-major,
 * G[a * (n_rad + n_ang) + g], i.e. the column-major (n_rad+n_ang) x N matrix
 * of the reference. */
typedef struct {
    int32_t n_rad, n_ang;
    const double *rad;
    const double *ang;
    double  r_cut;
    double  hardcore2;    /* RuNNerPotential.boundary (squared); 0 = none */
    int32_t q10_compat;   /* 1: initial G as total_symm_calc builds it (SymmFunc.jl:504) */
    int32_t n_layers;
    const int32_t *nodes;
    const int32_t *act;   /* per layer: 1 linear, 2 tanh (layer 1 = input) */
    int32_t n_params;
    const double *params;
} orc_nnp;

/* Two-species network potential (RuNNerPotential2Atom).  rad: n_rad rows
 * (eta, norm[4], offset[4]) of RadialType2a (norm / offset order: species 1
 * rows 1, 2, species 2 rows 1, 2); ang: n_ang rows (eta, lambda, zeta,
 * tpz[6], offset[6]) of AngularType3a.  params1 / params2: the networks of
 * species 1 / 2 (same layout). */
typedef struct {
    int32_t n_rad, n_ang;
    const double *rad;
    const double *ang;
    double  r_cut;
    int32_t n_layers;
    const int32_t *nodes;
    const int32_t *act;
    int32_t n_params;
    const double *params1, *params2;
} orc_nnp2;

typedef struct {
    int32_t kind;
    int32_t ncoeff;
    double  coeff[16];
    double  eam_n, eam_m, eam_ean, eam_ecam;
    const orc_nnp *nnp;
    /* ELJPotentialB (src/EnergyEvaluation.jl:161-182): coeff holds c, these a and b */
    double  coeff_a[8], coeff_b[8];
    /* LookupTablePotential (src/EnergyEvaluation.jl:211-295): table[angle][dist] */
    int32_t lut_ndist, lut_nangle;
    double  lut_start, lut_d, lut_dangle, lut_c6;
    const double *lut;
    /* RuNNerPotential2Atom in the NNVT ensemble (src/Ensembles.jl:138-150):
     * atoms 0..nnvt_n1-1 are species 1, the rest species 2; a cycle holds
     * N atom moves and nnvt_swaps atom swaps (src/Ensembles.jl:270-282) */
    const orc_nnp2 *nnp2;
    int32_t nnvt_n1, nnvt_swaps;
} orc_pot;

typedef struct {
    double  *x;            /* R*N*3 positions, temperature-position order */
    double  *box;          /* R box lengths (cubic) */
    double  *max_box;      /* R max box lengths (NPT) */
    double  *r_cut;        /* R squared cut-off (NPT) */
    double  *en_tot;       /* R */
    double  *beta;         /* R */
    double  *max_displ;    /* R*2 : atom, volume */
    int64_t *count_atom;   /* R*2 : total, window */
    int64_t *count_vol;    /* R*2 */
    int64_t *count_exc;    /* R*2 : attempted, accepted */
    double  *ham;          /* R*2 */
    double  *hist;         /* R*(n_bin+2) NVT energy histograms, or R*(n_bin+2)^2 NPT (E,V) */
    double  *d2;           /* R*N*N cached squared distances (dist_cache = 1) */
    double  *comp;         /* R*N*2 EAM phi/rho components */
    double  *g;            /* R*N*(n_rad+n_ang) NNP feature matrices */
    double  *eat;          /* R*N NNP per-atom energies */
    int32_t *error;        /* out: 0 ok, 1 too many moves out of binding sphere */
    /* optional per-move trace, indexed [r][cycle][move] */
    double  *trace_de;     /* new_en - en_tot as used by the Metropolis test */
    int8_t  *trace_acc;    /* 1 accepted */
    int32_t *trace_idx;    /* moved atom (N = volume move) */
    const int8_t *replay_acc; /* if non-null: take these accept decisions (fp32 replay) */
    /* rectangular / rhombic boxes (NULL otherwise) */
    double  *box_h;        /* R box heights */
    double  *max_box_h;    /* R max box heights (src/MCStates.jl:75-86) */
    double  *lh_ratio;     /* R length / height ratio bounds (src/MCStates.jl:106) */
    double  *max_displ_sep;/* R*2 : xy, z volume steps (separated moves) */
    int64_t *count_vol_xy; /* R*2 */
    int64_t *count_vol_z;  /* R*2 */
    /* optional per-move trace of the oracle's own Metropolis test in FP64,
     * also in replay mode: prob = min(1, e^{-beta dE}) (NPT volume moves:
     * with the P dV and N ln(V'/V) terms) and the uniform it is compared with */
    double  *trace_prob;
    double  *trace_u;
} orc_state;

/* Build the r^2 matrices / EAM components and total energies from positions
 * (src/MCStates.jl:137-162 -> src/EnergyEvaluation.jl:1662-1734,1778-1811). */
int orc_init_states(const orc_params *p, const orc_pot *pot, orc_state *s);

/* Advance n_sweeps cycles. */
int orc_run_cycles(const orc_params *p, const orc_pot *pot, orc_state *s, int64_t n_sweeps);

/* Helpers exported for direct known-answer tests. */
double orc_dimer_energy(const orc_pot *pot, double r2, int pair_form);
/* ELJPotentialB pair energy at r^2 and tan of the angle to z (src/EnergyEvaluation.jl:943-958) */
double orc_dimer_energy_b(const orc_pot *pot, double r2, double tan_z);
/* LookupTablePotential pair energy (src/EnergyEvaluation.jl:273-295) */
double orc_dimer_energy_lut(const orc_pot *pot, double r2, double tan_z);
/* squared distance and tan of its angle to z (get_tan, src/Configurations.jl:172-218) */
double orc_pair_geom(const double *a, const double *b, int bc, double l, double h, int form, double *tan_z);
double orc_distance2(const double *a, const double *b, int bc, double box, int cached_form);
double orc_energy_config(const orc_params *p, const orc_pot *pot, const double *x,
                         double box, double r_cut, double *en_atom);
double orc_lrc(int n_atoms, double r_cut, const orc_pot *pot);
/* boxes with a height (rectangular, rhombic; cubic and sphere ignore h) */
double orc_distance2_box(const double *a, const double *b, int bc, double l, double h, int form);
double orc_energy_config_box(const orc_params *p, const orc_pot *pot, const double *x, double l, double h,
                             double r_cut, double *en_atom);
double orc_r_cut(int bc, double l, double h);
double orc_volume(int bc, double l, double h);
double orc_lrc_box(int n_atoms, double r_cut, const orc_pot *pot, int bc, double l, double h);
double orc_hist_volume(int bc, double l, double h);
void   orc_wrap(int bc, double l, double h, int form, double *t);
double orc_exp_det(double x);
double orc_log_det(double x);
void   orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* The potential plugin functions of src/EnergyEvaluation.jl for one
 * configuration (r^2 matrices N x N row-major, EAM components N x 2). */
void   orc_distance2_mat(const orc_params *p, const double *x, double l, double h, double *d2);
double orc_energy_config_d2(const orc_params *p, const orc_pot *pot, const double *d2, double r_cut, double l,
                            double h, double *en_atom);
void   orc_eam_components(const orc_params *p, const orc_pot *pot, const double *d2, double *comp);
double orc_eam_energy(const orc_pot *pot, const double *comp, int n_atoms);
double orc_energy_update(const orc_params *p, const orc_pot *pot, const double *x, const double *d2, double l,
                         double h, int i, const double *trial, double en_tot, const double *comp, const double *g,
                         double *new_comp, double *g_new, double *eat_new);

/* NNP pieces (ptmc_nnp.c). */
/* the RuNNer forward pass with the argument list of the reference's ccall
 * (MachineLearningPotential/ForwardPass.jl:60-84); returns eatom[batch-1] */
double orc_forward(const double *input, const int32_t *n_in, const int32_t *n_layers, const int32_t *nodes,
                   const int32_t *batch, const int32_t *n_params, const double *params, const int32_t *act,
                   double *eatom);
double orc_cutoff(double r, double r_cut);
/* total_symm_calc (SymmFunc.jl:494-510): full feature matrix from positions */
int    orc_nnp_full_g(const orc_nnp *nn, const double *x, int n_atoms, double *g);
/* total_thr_symm! (DeltaMatrix.jl:933-975): features after moving atom i to trial */
int    orc_nnp_move_g(const orc_nnp *nn, const double *x, int n_atoms, int i, const double *trial,
                      const double *g, double *g_new);
/* sum of the per-atom network energies of a feature matrix */
double orc_nnp_energy(const orc_nnp *nn, const double *g, int n_atoms, double *eatom);

/* features (2 n_rad + 3 n_ang per atom, atom-major) of atoms 0..n1-1
 * (species 1) and n1..n1+n2-1 (species 2) in a spherical container */
int    orc_nnp2_full_g(const orc_nnp2 *nn, const double *x, int n1, int n2, double *g);
double orc_nnp2_energy(const orc_nnp2 *nn, const double *g, int n1, int n2, double *eatom);
/* calc_delta_matrix (DeltaMatrix.jl:219-262): features after moving atom i to trial */
int    orc_nnp2_move_g(const orc_nnp2 *nn, const double *x, int n1, int n2, int i, const double *trial,
                       const double *g, double *g_new);
/* calc_swap_matrix (DeltaMatrix.jl:583-624): features after atoms i1 (species 1)
 * and i2 (species 2) trade places */
int    orc_nnp2_swap_g(const orc_nnp2 *nn, const double *x, int n1, int n2, int i1, int i2, const double *g,
                       double *g_new);

#ifdef __cplusplus
}
#endif
#endif
