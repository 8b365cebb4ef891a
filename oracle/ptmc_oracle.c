/*
 * ORACLE — test infrastructure only.  Only tests/, __graft_entry__.smoke()
 * and bench.py's CPU-baseline leg load this library, and only as the checker
 * or the timed CPU baseline; the shipped engine never links it.
 *
 * FP64 CPU restatement of the reference's PT sweep for pair (ELJ) and EAM
 * potentials.  Each function cites the reference lines it follows.  Two
 * switches select the arithmetic:
 *   - reference form (pair_form = 0, dist_cache = 1, lanes = 1, math_lib = 1)
 *     restates the Julia code operation by operation;
 *   - engine form (pair_form = 1, dist_cache = 0, lanes = k, math_lib = 0)
 *     restates what the device engine computes (Horner pair energy, distances
 *     recomputed from positions, k-wide butterfly sums, deterministic exp/log)
 *     so device results can be compared bit for bit.
 */
#include "ptmc_oracle.h"
#include "orc_math.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXN 1024

double orc_exp_det(double x) { return orc_exp(x); }
double orc_log_det(double x) { return orc_log(x); }

void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    orc_u32x4 c;
    memcpy(c.v, ctr, sizeof c.v);
    orc_u32x4 r = orc_philox4x32_10(c, key[0], key[1]);
    memcpy(out, r.v, sizeof r.v);
}

static double xexp(const orc_params *p, double x) { return p->math_lib ? exp(x) : orc_exp(x); }
static double xlog(const orc_params *p, double x) { return p->math_lib ? log(x) : orc_log(x); }

/* ---------------------------------------------------------------- pairs */

/* ELJ pair energies, src/EnergyEvaluation.jl:923-942.
 * pair_form 0 restates the Julia loops exactly; pair_form 1 is the engine's
 * Horner evaluation in u = 1/r^2 (even) or v = 1/r (all powers). */
double orc_dimer_energy(const orc_pot *pot, double r2, int pair_form)
{
    const int K = pot->ncoeff;
    if (pot->kind == ORC_POT_ELJ_EVEN) {
        if (pair_form == 0) {
            double r6inv = 1.0 / (r2 * r2 * r2);
            double sum1 = 0.0;
            for (int i = 0; i < K; ++i) {
                sum1 += pot->coeff[i] * r6inv;
                r6inv /= r2;
            }
            return sum1;
        }
        double u = 1.0 / r2;
        double q = pot->coeff[K - 1];
        for (int i = K - 2; i >= 0; --i) q = fma(q, u, pot->coeff[i]);
        double u3 = u * u * u;
        return q * u3;
    }
    if (pot->kind == ORC_POT_ELJ) {
        if (pair_form == 0) {
            double r = sqrt(r2);
            double r6inv = 1.0 / (r2 * r2 * r2);
            double sum1 = 0.0;
            for (int i = 0; i < K; ++i) {
                sum1 += pot->coeff[i] * r6inv;
                r6inv /= r;
            }
            return sum1;
        }
        double v = 1.0 / sqrt(r2);
        double q = pot->coeff[K - 1];
        for (int i = K - 2; i >= 0; --i) q = fma(q, v, pot->coeff[i]);
        double u = 1.0 / r2;
        double u3 = u * u * u;
        return q * u3;
    }
    return 0.0;
}

/* Long-range correction, src/EnergyEvaluation.jl:968-982 (even powers only). */
double orc_lrc(int n_atoms, double r_cut, const orc_pot *pot)
{
    if (r_cut <= 50.0) return 0.0;
    double rcs = sqrt(r_cut);
    double rc3 = r_cut * rcs;
    double e = 0.0;
    for (int i = 1; i <= pot->ncoeff; ++i) {
        e += pot->coeff[i - 1] / rc3 / (double)(2 * i + 1);
        rc3 *= r_cut;
    }
    double n2 = (double)((int64_t)n_atoms * n_atoms);
    e *= M_PI * n2 / 4.0 / (rcs * rcs * rcs);
    return e;
}

/* Squared distance, src/Configurations.jl:101-109.  bc cubic applies the
 * minimum image to b: b' = b + L round((a-b)/L).  form 0 divides by L as the
 * reference does; form 1 multiplies by 1/L as the engine does. */
double orc_distance2(const double *a, const double *b, int bc, double box, int form)
{
    double bb[3] = {b[0], b[1], b[2]};
    if (bc == ORC_CUBIC) {
        double inv = 1.0 / box;
        for (int c = 0; c < 3; ++c) {
            double q = form == 0 ? (a[c] - b[c]) / box : (a[c] - b[c]) * inv;
            bb[c] = b[c] + box * rint(q);
        }
    }
    double dx = a[0] - bb[0], dy = a[1] - bb[1], dz = a[2] - bb[2];
    return dx * dx + dy * dy + dz * dz;
}

/* EAM invrexp, src/EnergyEvaluation.jl:1088-1095. */
static void eam_terms(const orc_params *p, const orc_pot *pot, double r2, double *phi, double *rho)
{
    if (r2 == 0.0) { *phi = 0.0; *rho = 0.0; return; }
    if (p->math_lib) {
        double rt = 1.0 / sqrt(r2);
        *phi = pow(rt, pot->eam_n);
        *rho = pow(rt, pot->eam_m);
    } else {
        double hl = -0.5 * orc_log(r2);
        *phi = orc_exp(pot->eam_n * hl);
        *rho = orc_exp(pot->eam_m * hl);
    }
}

/* k-wide xor-butterfly sum of per-lane partials (engine summation order). */
static double tree_sum(double *v, int k)
{
    double w[64];
    for (int m = 1; m < k; m <<= 1) {
        for (int q = 0; q < k; ++q) w[q] = v[q] + v[q ^ m];
        for (int q = 0; q < k; ++q) v[q] = w[q];
    }
    return v[0];
}

static inline int dist_form(const orc_params *p) { return p->pair_form == 0 ? 0 : 1; }

/* Full configuration energy, src/EnergyEvaluation.jl:378-420: sum over i<j
 * (NPT: pairs inside r_cut, plus lrc).  en_atom optional.  With lanes k > 1
 * rows are dealt to lanes round-robin and combined by the butterfly. */
double orc_energy_config(const orc_params *p, const orc_pot *pot, const double *x,
                         double box, double r_cut, double *en_atom)
{
    const int N = p->n_atoms;
    const int k = p->lanes > 0 ? p->lanes : 1;
    double part[64] = {0};
    if (en_atom) for (int i = 0; i < N; ++i) en_atom[i] = 0.0;
    for (int i = 0; i < N; ++i) {
        int q = i % k;
        for (int j = i + 1; j < N; ++j) {
            double r2 = orc_distance2(&x[3 * i], &x[3 * j], p->bc, box, dist_form(p));
            if (p->ensemble == ORC_NPT && !(r2 <= r_cut)) continue;
            double e = orc_dimer_energy(pot, r2, p->pair_form);
            if (en_atom) { en_atom[i] += e; en_atom[j] += e; }
            part[q] += e;
        }
    }
    double tot = tree_sum(part, k);
    if (p->ensemble == ORC_NPT) tot += orc_lrc(N, r_cut, pot);
    return tot;
}

/* EAM components and energy, src/EnergyEvaluation.jl:1105-1162, 1800-1811.
 * k = 1: the sequential sum of the reference; k > 1: atoms dealt to lanes
 * round-robin and combined by the butterfly (engine initialisation). */
static double eam_energy_from_comp(const orc_pot *pot, const double *comp, int N, int k)
{
    double part[64] = {0};
    for (int i = 0; i < N; ++i) part[i % k] += pot->eam_ean * comp[2 * i] - pot->eam_ecam * sqrt(comp[2 * i + 1]);
    return tree_sum(part, k);
}

static void eam_build(const orc_params *p, const orc_pot *pot, const double *x, double box,
                      double *comp)
{
    const int N = p->n_atoms;
    for (int i = 0; i < N; ++i) {
        double a = 0.0, b = 0.0;
        for (int j = 0; j < N; ++j) {
            double r2 = i == j ? 0.0 : orc_distance2(&x[3 * i], &x[3 * j], p->bc, box, dist_form(p));
            double f, g;
            eam_terms(p, pot, r2, &f, &g);
            a += f; b += g;
        }
        comp[2 * i] = a; comp[2 * i + 1] = b;
    }
}

int orc_init_states(const orc_params *p, const orc_pot *pot, orc_state *s)
{
    const int N = p->n_atoms, R = p->n_temps;
    if (N > MAXN || N < 2) return -1;
    for (int r = 0; r < R; ++r) {
        double *x = &s->x[(size_t)r * N * 3];
        double box = s->box ? s->box[r] : 0.0;
        if (p->ensemble == ORC_NPT) s->r_cut[r] = box * box / 4.0;   /* src/Ensembles.jl:118-120 */
        if (s->d2) {
            double *d2 = &s->d2[(size_t)r * N * N];
            for (int i = 0; i < N; ++i) {
                d2[i * N + i] = 0.0;
                for (int j = i + 1; j < N; ++j)
                    d2[i * N + j] = d2[j * N + i] =
                        orc_distance2(&x[3 * i], &x[3 * j], p->bc, box, dist_form(p));
            }
        }
        if (pot->kind == ORC_POT_NNP) {
            /* set_variables + initialise_energy, src/EnergyEvaluation.jl:1812-1822, :1735-1747 */
            const int nf = pot->nnp->n_rad + pot->nnp->n_ang;
            if (orc_nnp_full_g(pot->nnp, x, N, &s->g[(size_t)r * N * nf])) return -1;
            s->en_tot[r] = orc_nnp_energy(pot->nnp, &s->g[(size_t)r * N * nf], N, &s->eat[(size_t)r * N]);
        } else if (pot->kind == ORC_POT_EAM) {
            eam_build(p, pot, x, box, &s->comp[(size_t)r * N * 2]);
            s->en_tot[r] = eam_energy_from_comp(pot, &s->comp[(size_t)r * N * 2], N,
                                                p->math_lib ? 1 : (p->lanes > 0 ? p->lanes : 1));
        } else {
            s->en_tot[r] = orc_energy_config(p, pot, x, box,
                                             p->ensemble == ORC_NPT ? s->r_cut[r] : 0.0, NULL);
        }
    }
    return 0;
}

/* ------------------------------------------------------------- the move */

typedef struct {
    uint32_t index;
    double ux, uy, uz, umet;
} draws;

static draws draw_move(const orc_params *p, uint32_t nsteps, int r, uint32_t sweep,
                       uint32_t m, int attempt)
{
    draws d;
    if (!p->fp32) {
        orc_u32x4 A = orc_draw(p->seed, p->run, ORC_T_MOVE, 2u * attempt, r, sweep, m);
        orc_u32x4 B = orc_draw(p->seed, p->run, ORC_T_MOVE, 2u * attempt + 1u, r, sweep, m);
        d.ux = orc_u53(A.v[0], A.v[1]);
        d.uy = orc_u53(A.v[2], A.v[3]);
        d.uz = orc_u53(B.v[0], B.v[1]);
        d.umet = orc_u53(B.v[2], B.v[3]);
        uint32_t w = (A.v[1] & 0x7ffu) | ((A.v[3] & 0x7ffu) << 11) | ((B.v[1] & 0x3ffu) << 22);
        d.index = orc_below(w, nsteps);
    } else {
        orc_u32x4 C = orc_draw(p->seed, p->run, ORC_T_MOVE, (uint32_t)attempt, r, sweep, m);
        d.index = orc_below(C.v[0], nsteps);
        d.ux = orc_u24(C.v[1]);
        d.uy = orc_u24(C.v[2]);
        d.uz = orc_u24(C.v[3]);
        uint32_t w = ((C.v[1] & 0xffu) << 16) | ((C.v[2] & 0xffu) << 8) | (C.v[3] & 0xffu);
        d.umet = (double)((float)w * 0x1.0p-24f);
    }
    return d;
}

/* trial position, src/MCMoves.jl:33-64: pos + (rand()-0.5)*delta per axis,
 * cubic: wrapped back with round(trial/L). */
static void make_trial(const orc_params *p, const double *pos, double delta, double box,
                       const draws *d, double *t)
{
    double u[3] = {d->ux, d->uy, d->uz};
    if (!p->fp32) {
        for (int c = 0; c < 3; ++c) t[c] = pos[c] + (u[c] - 0.5) * delta;
        if (p->bc == ORC_CUBIC) {
            double inv = 1.0 / box;
            for (int c = 0; c < 3; ++c) {
                double q = p->pair_form == 0 ? t[c] / box : t[c] * inv;
                t[c] = t[c] - box * rint(q);
            }
        }
    } else {
        float fd = (float)delta;
        for (int c = 0; c < 3; ++c) {
            float v = (float)pos[c] + ((float)u[c] - 0.5f) * fd;
            t[c] = (double)v;
        }
        if (p->bc == ORC_CUBIC) {
            float L = (float)box, inv = (float)(1.0 / box);
            for (int c = 0; c < 3; ++c) {
                float v = (float)t[c];
                float q = rintf(v * inv);
                v = v - L * q;
                t[c] = (double)v;
            }
        }
    }
}

/* check_boundary, src/BoundaryConditions.jl:130 */
static int outside(const orc_params *p, const double *t)
{
    if (p->fp32) {
        float a = (float)t[0], b = (float)t[1], c = (float)t[2];
        float s = a * a + b * b;
        s = s + c * c;
        return s > (float)p->radius2;
    }
    return t[0] * t[0] + t[1] * t[1] + t[2] * t[2] > p->radius2;
}

/* recentre!, src/Configurations.jl:79-85 (sequential mean) */
static void recentre(const orc_params *p, double *x)
{
    const int N = p->n_atoms;
    double c[3] = {0, 0, 0};
    for (int i = 0; i < N; ++i) for (int k = 0; k < 3; ++k) c[k] += x[3 * i + k];
    for (int k = 0; k < 3; ++k) c[k] /= (double)N;
    for (int i = 0; i < N; ++i)
        for (int k = 0; k < 3; ++k) {
            double v = x[3 * i + k] - c[k];
            x[3 * i + k] = p->fp32 ? (double)(float)v : v;
        }
}

typedef struct {
    double new_en;
    double new_row[MAXN];
    double new_comp[2 * MAXN];
    double *g_new;    /* NNP: N * n_features */
    double *eat_new;  /* NNP: N */
} trial_out;

/* Energy of an atom move: src/EnergyEvaluation.jl:804-821 (dimer) and
 * :1567-1591 (EAM); new_dist2_vec as src/MCMoves.jl:138-143. */
static void atom_move_energy(const orc_params *p, const orc_pot *pot, const orc_state *s, int r,
                             int i, const double *trial, double en_tot, trial_out *o)
{
    const int N = p->n_atoms;
    const int k = p->lanes > 0 ? p->lanes : 1;
    const double *x = &s->x[(size_t)r * N * 3];
    const double box = s->box ? s->box[r] : 0.0;
    const double *d2 = s->d2 ? &s->d2[(size_t)r * N * N] : NULL;
    const int cutoff = p->ensemble == ORC_NPT;
    const double rc = cutoff ? s->r_cut[r] : 0.0;

    for (int j = 0; j < N; ++j)
        o->new_row[j] = j == i ? 0.0 : orc_distance2(trial, &x[3 * j], p->bc, box, dist_form(p));

    if (pot->kind == ORC_POT_NNP) {
        /* energy_update!(RuNNerPotential), src/EnergyEvaluation.jl:1593-1621: the
         * hard-core test gives the sentinel energy 100 */
        for (int j = 0; j < N; ++j)
            if (j != i && o->new_row[j] < pot->nnp->hardcore2) {
                o->new_en = 100.0;
                return;
            }
        const int nf = pot->nnp->n_rad + pot->nnp->n_ang;
        orc_nnp_move_g(pot->nnp, x, N, i, trial, &s->g[(size_t)r * N * nf], o->g_new);
        o->new_en = orc_nnp_energy(pot->nnp, o->g_new, N, o->eat_new);
        return;
    }

    if (pot->kind == ORC_POT_EAM) {
        const double *comp = &s->comp[(size_t)r * N * 2];
        if (p->math_lib) {
            /* reference: copy, then add j terms to rows j and i in order */
            memcpy(o->new_comp, comp, sizeof(double) * 2 * N);
            for (int j = 0; j < N; ++j) {
                double old = j == i ? 0.0 : (d2 ? d2[i * N + j]
                                                : orc_distance2(&x[3 * i], &x[3 * j], p->bc, box, dist_form(p)));
                double fn, gn, fo, go;
                eam_terms(p, pot, o->new_row[j], &fn, &gn);
                eam_terms(p, pot, old, &fo, &go);
                double t1 = fn - fo, t2 = gn - go;
                o->new_comp[2 * j] += t1;
                o->new_comp[2 * i] += t1;
                o->new_comp[2 * j + 1] += t2;
                o->new_comp[2 * i + 1] += t2;
            }
            o->new_en = eam_energy_from_comp(pot, o->new_comp, N, 1);
        } else {
            /* engine: per-lane partial sums of the j != i terms, butterfly,
             * then the moved atom's own term last */
            double pt[64] = {0}, pf[64] = {0}, pg[64] = {0};
            memcpy(o->new_comp, comp, sizeof(double) * 2 * N);
            for (int j = 0; j < N; ++j) {
                if (j == i) continue;
                int q = j % k;
                double old = d2 ? d2[i * N + j]
                                : orc_distance2(&x[3 * i], &x[3 * j], p->bc, box, dist_form(p));
                double fn, gn, fo, go;
                eam_terms(p, pot, o->new_row[j], &fn, &gn);
                eam_terms(p, pot, old, &fo, &go);
                double t1 = fn - fo, t2 = gn - go;
                double nf = comp[2 * j] + t1, ng = comp[2 * j + 1] + t2;
                o->new_comp[2 * j] = nf;
                o->new_comp[2 * j + 1] = ng;
                pt[q] += pot->eam_ean * nf - pot->eam_ecam * sqrt(ng);
                pf[q] += t1;
                pg[q] += t2;
            }
            double st = tree_sum(pt, k), sf = tree_sum(pf, k), sg = tree_sum(pg, k);
            double nfi = comp[2 * i] + sf, ngi = comp[2 * i + 1] + sg;
            o->new_comp[2 * i] = nfi;
            o->new_comp[2 * i + 1] = ngi;
            o->new_en = st + (pot->eam_ean * nfi - pot->eam_ecam * sqrt(ngi));
        }
        return;
    }

    double pn[64] = {0}, po[64] = {0};
    for (int j = 0; j < N; ++j) {
        if (j == i) continue;
        int q = j % k;
        double rn = o->new_row[j];
        double ro = d2 ? d2[i * N + j]
                       : orc_distance2(&x[3 * i], &x[3 * j], p->bc, box, dist_form(p));
        if (!cutoff || rn <= rc) pn[q] += orc_dimer_energy(pot, rn, p->pair_form);
        if (!cutoff || ro <= rc) po[q] += orc_dimer_energy(pot, ro, p->pair_form);
    }
    double sn = tree_sum(pn, k), so = tree_sum(po, k);
    double delta_en = sn - so;
    o->new_en = delta_en + en_tot;
}

static void swap_states(const orc_params *p, const orc_pot *pot, orc_state *s, int a, int b)
{
    const int N = p->n_atoms;
#define SWAPN(ptr, n)                                                                      \
    if (ptr) {                                                                             \
        for (size_t t_ = 0; t_ < (size_t)(n); ++t_) {                                      \
            double tmp_ = (ptr)[(size_t)a * (n) + t_];                                     \
            (ptr)[(size_t)a * (n) + t_] = (ptr)[(size_t)b * (n) + t_];                     \
            (ptr)[(size_t)b * (n) + t_] = tmp_;                                            \
        }                                                                                  \
    }
    SWAPN(s->x, (size_t)N * 3);
    SWAPN(s->d2, (size_t)N * N);
    SWAPN(s->en_tot, 1);
    if (p->bc == ORC_CUBIC) SWAPN(s->box, 1);
    if (p->ensemble == ORC_NPT) SWAPN(s->r_cut, 1);
    if (pot->kind == ORC_POT_EAM) SWAPN(s->comp, (size_t)N * 2);
    if (pot->kind == ORC_POT_NNP) {
        SWAPN(s->g, (size_t)N * (pot->nnp->n_rad + pot->nnp->n_ang));
        SWAPN(s->eat, (size_t)N);
    }
#undef SWAPN
}

/* exc_acceptance src/Exchange.jl:103-108, parallel_tempering_exchange!
 * :132-186 (NPT uses E + P V). */
static void try_exchange(const orc_params *p, const orc_pot *pot, orc_state *s, int n, uint32_t sweep)
{
    orc_u32x4 X = orc_draw(p->seed, p->run, ORC_T_EXCH, 0, 0, sweep, 1u + (uint32_t)n);
    double u = orc_u53(X.v[0], X.v[1]);
    s->count_exc[2 * n] += 1;
    s->count_exc[2 * (n + 1)] += 1;
    double e1 = s->en_tot[n], e2 = s->en_tot[n + 1];
    if (p->ensemble == ORC_NPT) {
        double v1 = s->box[n] * s->box[n] * s->box[n];
        double v2 = s->box[n + 1] * s->box[n + 1] * s->box[n + 1];
        e1 = e1 + p->pressure * v1;
        e2 = e2 + p->pressure * v2;
    }
    double de = e1 - e2;
    double db = s->beta[n] - s->beta[n + 1];
    double acc = xexp(p, db * de);
    if (acc > 1.0) acc = 1.0;
    if (acc > u) {
        s->count_exc[2 * n + 1] += 1;
        s->count_exc[2 * (n + 1) + 1] += 1;
        swap_states(p, pot, s, n, n + 1);
    }
}

/* find_hist_index, src/Sampling.jl:53-66 (returns 1-based index) */
static int hist_index(double hi, int n_bin)
{
    if (hi < 1.0) return 1;
    if (hi > (double)n_bin) return n_bin + 2;
    return (int)floor(hi) + 1;
}

int orc_run_cycles(const orc_params *p, const orc_pot *pot, orc_state *s, int64_t n_sweeps)
{
    const int N = p->n_atoms, R = p->n_temps;
    const uint32_t nsteps = (uint32_t)(p->ensemble == ORC_NPT ? N + 1 : N);
    trial_out *o = (trial_out *)calloc(1, sizeof(trial_out));
    double *vol_x = (double *)malloc(sizeof(double) * 3 * N);
    if (!o || !vol_x) return -2;
    if (pot->kind == ORC_POT_NNP) {
        if (p->ensemble != ORC_NVT) return -3;
        o->g_new = (double *)malloc(sizeof(double) * N * (pot->nnp->n_rad + pot->nnp->n_ang));
        o->eat_new = (double *)malloc(sizeof(double) * N);
    }
    if (s->error) *s->error = 0;

    for (int64_t c = 0; c < n_sweeps; ++c) {
        const uint32_t sweep = (uint32_t)(p->sweep0 + c);
        const int64_t cyc = p->cycle0 + c + 1;

        /* mc_step!: every replica makes nsteps moves, src/MCRun.jl:169-182 */
        for (int r = 0; r < R; ++r) {
            double *x = &s->x[(size_t)r * N * 3];
            double box = s->box ? s->box[r] : 0.0;
            for (uint32_t m = 0; m < nsteps; ++m) {
                draws d0 = draw_move(p, nsteps, r, sweep, m, 0);
                const int i = (int)d0.index;
                double new_en;
                int accept;
                size_t tix = ((size_t)r * n_sweeps + c) * nsteps + m;
                if (i < N) {
                    /* atom move, src/MCMoves.jl:115-145 / :100-113 */
                    double t[3];
                    make_trial(p, &x[3 * i], s->max_displ[2 * r], box, &d0, t);
                    if (p->bc == ORC_SPHERE) {
                        int count = 0, attempt = 0;
                        while (outside(p, t)) {
                            count += 1;
                            if (count == 50) {
                                recentre(p, x);
                            } else {
                                attempt += 1;
                                draws dd = draw_move(p, nsteps, r, sweep, m, attempt);
                                make_trial(p, &x[3 * i], s->max_displ[2 * r], box, &dd, t);
                                if (count == 100) {
                                    if (s->error) *s->error = 1;
                                    free(o->g_new); free(o->eat_new); free(o); free(vol_x);
                                    return 1;
                                }
                            }
                        }
                    }
                    atom_move_energy(p, pot, s, r, i, t, s->en_tot[r], o);
                    new_en = o->new_en;
                    double dE = new_en - s->en_tot[r];
                    if (p->fp32 && s->replay_acc) {
                        accept = s->replay_acc[tix];
                    } else {
                        double prob = xexp(p, -dE * s->beta[r]);
                        if (prob > 1.0) prob = 1.0;
                        accept = prob >= d0.umet;
                    }
                    if (s->trace_de) s->trace_de[tix] = dE;
                    if (s->trace_acc) s->trace_acc[tix] = (int8_t)accept;
                    if (s->trace_idx) s->trace_idx[tix] = i;
                    if (accept) {
                        /* swap_atom_config!, src/swap_config.jl:40-51 */
                        for (int k = 0; k < 3; ++k) x[3 * i + k] = t[k];
                        if (s->d2) {
                            double *d2 = &s->d2[(size_t)r * N * N];
                            for (int j = 0; j < N; ++j) d2[i * N + j] = d2[j * N + i] = o->new_row[j];
                        }
                        if (pot->kind == ORC_POT_EAM)
                            memcpy(&s->comp[(size_t)r * N * 2], o->new_comp, sizeof(double) * 2 * N);
                        if (pot->kind == ORC_POT_NNP && new_en != 100.0) {
                            /* swap_vars!(NNPVariables), src/swap_config.jl:213-221 */
                            const size_t nf = (size_t)(pot->nnp->n_rad + pot->nnp->n_ang);
                            memcpy(&s->g[(size_t)r * N * nf], o->g_new, sizeof(double) * N * nf);
                            memcpy(&s->eat[(size_t)r * N], o->eat_new, sizeof(double) * N);
                        }
                        s->en_tot[r] = new_en;
                        s->count_atom[2 * r] += 1;
                        s->count_atom[2 * r + 1] += 1;
                    }
                } else {
                    /* uniform volume move, src/MCMoves.jl:169-177, 277-287 */
                    double scale;
                    double y = (d0.ux - 0.5) * s->max_displ[2 * r + 1];
                    if (p->math_lib) scale = pow(exp(y), 1.0 / 3.0);
                    else scale = orc_exp(y / 3.0);
                    if (box >= s->max_box[r] && scale > 1.0) scale = 1.0;
                    double nbox = box * scale;
                    double nrc = nbox * nbox / 4.0;
                    double s2 = scale * scale;
                    /* energy of the scaled configuration: scaled r^2 matrix,
                     * src/EnergyEvaluation.jl:397-420 */
                    double part[64] = {0};
                    const int k = p->lanes > 0 ? p->lanes : 1;
                    for (int a = 0; a < N; ++a) {
                        for (int b = a + 1; b < N; ++b) {
                            double r2 = s->d2 ? s->d2[(size_t)r * N * N + a * N + b]
                                              : orc_distance2(&x[3 * a], &x[3 * b], p->bc, box, dist_form(p));
                            r2 = r2 * s2;
                            if (r2 <= nrc) part[a % k] += orc_dimer_energy(pot, r2, p->pair_form);
                        }
                    }
                    new_en = tree_sum(part, k) + orc_lrc(N, nrc, pot);
                    double dE = new_en - s->en_tot[r];
                    double v_new = nbox * nbox * nbox, v_old = box * box * box;
                    double dh = dE + p->pressure * (v_new - v_old);
                    double arg = -dh * s->beta[r] + (double)N * xlog(p, v_new / v_old);
                    double prob = xexp(p, arg);
                    if (prob > 1.0) prob = 1.0;
                    if (p->fp32 && s->replay_acc) accept = s->replay_acc[tix];
                    else accept = prob >= d0.umet;
                    if (s->trace_de) s->trace_de[tix] = dE;
                    if (s->trace_acc) s->trace_acc[tix] = (int8_t)accept;
                    if (s->trace_idx) s->trace_idx[tix] = i;
                    if (accept) {
                        /* swap_config_v!, src/swap_config.jl:113-140 */
                        for (int a = 0; a < 3 * N; ++a) {
                            double v = x[a] * scale;
                            x[a] = p->fp32 ? (double)(float)v : v;
                        }
                        if (s->d2)
                            for (int a = 0; a < N * N; ++a) s->d2[(size_t)r * N * N + a] *= s2;
                        s->box[r] = nbox;
                        box = nbox;
                        s->r_cut[r] = nrc;
                        s->en_tot[r] = new_en;
                        s->count_vol[2 * r] += 1;
                        s->count_vol[2 * r + 1] += 1;
                    }
                }
            }
        }

        /* exchange, src/MCRun.jl:203-205 */
        if (R > 1) {
            if (p->exchange_mode == ORC_EXC_EVEN_ODD) {
                for (int n = (int)(sweep & 1u); n < R - 1; n += 2) try_exchange(p, pot, s, n, sweep);
            } else {
                orc_u32x4 X = orc_draw(p->seed, p->run, ORC_T_EXCH, 0, 0, sweep, 0);
                double u = orc_u53(X.v[0], X.v[1]);
                if (p->exchange_mode == ORC_EXC_EVERY || u < 0.1) {
                    int n = (int)orc_below(X.v[2], (uint32_t)(R - 1));
                    try_exchange(p, pot, s, n, sweep);
                }
            }
        }

        /* update_max_stepsize!, src/Exchange.jl:198-263 */
        if (cyc % p->n_adjust == 0) {
            for (int r = 0; r < R; ++r) {
                double acc = (double)s->count_atom[2 * r + 1] / (double)((int64_t)p->n_adjust * N);
                if (acc < p->min_acc) s->max_displ[2 * r] *= 0.9;
                else if (acc > p->max_acc) s->max_displ[2 * r] *= 1.1;
                s->count_atom[2 * r + 1] = 0;
                if (p->ensemble == ORC_NPT) {
                    double av = (double)s->count_vol[2 * r + 1] / (double)((int64_t)p->n_adjust * 1);
                    if (av < p->min_acc) s->max_displ[2 * r + 1] *= 0.9;
                    else if (av > p->max_acc) s->max_displ[2 * r + 1] *= 1.1;
                    s->count_vol[2 * r + 1] = 0;
                }
            }
        }

        /* sampling_step!, src/Sampling.jl:330-371 */
        if (p->sampling && cyc % p->mc_sample == 0) {
            const int nb2 = p->n_bin + 2;
            for (int r = 0; r < R; ++r) {
                double e = s->en_tot[r];
                if (p->ensemble == ORC_NPT) {
                    double v = s->box[r] * s->box[r] * s->box[r];
                    double h = e + p->pressure * v;
                    s->ham[2 * r] += h;
                    s->ham[2 * r + 1] += h * h;
                    if (s->hist) {
                        double ie = floor((e - p->en_min) / p->delta_en_hist) + 1.0;
                        double iv = floor((v - p->v_min) / p->delta_v_hist) + 1.0;
                        int he = hist_index(ie, p->n_bin), hv = hist_index(iv, p->n_bin);
                        s->hist[(size_t)r * nb2 * nb2 + (size_t)(he - 1) * nb2 + (hv - 1)] += 1.0;
                    }
                } else {
                    s->ham[2 * r] += e;
                    s->ham[2 * r + 1] += e * e;
                    if (s->hist) {
                        double hi = (e - p->en_min) / p->delta_en_hist + 1.0;
                        int h = hist_index(hi, p->n_bin);
                        s->hist[(size_t)r * nb2 + (h - 1)] += 1.0;
                    }
                }
            }
        }
    }
    free(o->g_new);
    free(o->eat_new);
    free(o);
    free(vol_x);
    return 0;
}
