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
 *     recomputed from positions, k-wide butterfly sums, deterministic exp/log,
 *     and for FP64 pair-potential atom moves the Metropolis test in threshold
 *     form dE <= -ln(u) / beta, equivalent to min(1, e^{-beta dE}) >= u)
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

/* ELJPotentialB, src/EnergyEvaluation.jl:943-958: c_k r^-p_k (1 + a_k cos 2t +
 * b_k cos 4t) with p = 6, 8, 9, 10, ... and t the angle to z (cos 2t from
 * tan t); a constant 0.1 inside r^2 < 5.30. */
double orc_dimer_energy_b(const orc_pot *pot, double r2, double tan_z)
{
    if (!(r2 >= 5.30)) return 0.1;
    double r6inv = 1.0 / (r2 * r2 * r2);
    const double t2 = 2.0 / (tan_z * tan_z + 1.0) - 1.0;
    const double t4 = 2.0 * (t2 * t2) - 1.0;
    double sum1 = pot->coeff[0] * r6inv * (1.0 + pot->coeff_a[0] * t2 + pot->coeff_b[0] * t4);
    r6inv /= r2;
    for (int i = 1; i < pot->ncoeff; ++i) {
        sum1 += pot->coeff[i] * r6inv * (1.0 + pot->coeff_a[i] * t2 + pot->coeff_b[i] * t4);
        r6inv /= sqrt(r2);
    }
    return sum1;
}

/* LookupTablePotential, src/EnergyEvaluation.jl:273-295: the table entry at
 * the nearest angle (1 degree rows from 0: index 1 below tan 0.5 degrees,
 * 91 above tan 89.5 degrees) and the nearest distance, the first column
 * inside the first half step, c6 / r^6 beyond the table. */
double orc_dimer_energy_lut(const orc_pot *pot, double r2, double tan_z)
{
    const double at = fabs(tan_z);
    int ai = 1;
    if (at > 0.00872687153 && at <= 114.592845357)
        ai = (int)rint(atan(at) / pot->lut_dangle * 180.0 / M_PI + 1.0);
    else if (at > 114.592845357)
        ai = 91;
    const double lo = pot->lut_start + 0.5 * pot->lut_d;
    const double hi = pot->lut_start + (double)pot->lut_ndist * pot->lut_d;
    if (r2 <= lo * lo) return pot->lut[(size_t)(ai - 1) * pot->lut_ndist];
    if (r2 <= hi * hi) {
        const int di = (int)rint((sqrt(r2) - pot->lut_start) / pot->lut_d);
        return pot->lut[(size_t)(ai - 1) * pot->lut_ndist + (di - 1)];
    }
    return pot->lut_c6 / (r2 * r2 * r2);
}

/* the energy of the pair (a, b): positions a, b in the box (l, h), squared
 * distance times m2 (a uniformly scaled trial keeps the angles) */
static double pair_energy_ab(const orc_params *p, const orc_pot *pot, const double *a, const double *b, double l,
                             double h, double m2, int form, double rc, int cutoff)
{
    double tz = 0.0;
    const int ang = pot->kind == ORC_POT_ELJB || pot->kind == ORC_POT_LUT;
    double r2 = orc_pair_geom(a, b, p->bc, l, h, form, ang ? &tz : NULL);
    r2 = r2 * m2;
    if (cutoff && !(r2 <= rc)) return 0.0;
    if (pot->kind == ORC_POT_LUT) return orc_dimer_energy_lut(pot, r2, tz);
    return pot->kind == ORC_POT_ELJB ? orc_dimer_energy_b(pot, r2, tz) : orc_dimer_energy(pot, r2, p->pair_form);
}

/* Long-range correction, src/EnergyEvaluation.jl:968-982 (even powers only). */
double orc_lrc(int n_atoms, double r_cut, const orc_pot *pot)
{
    if (pot->kind == ORC_POT_LUT) {
        /* lrc(LookupTablePotential), src/EnergyEvaluation.jl:999-1008 (1.0 for r_cut <= 10) */
        if (r_cut <= 10.0) return 1.0;
        double rcs = sqrt(r_cut);
        double rc3 = r_cut * rcs;
        double e = pot->lut_c6 / rc3 / 3.0;
        e *= M_PI * (double)((int64_t)n_atoms * n_atoms) / 4.0 / (rcs * rcs * rcs);
        return e;
    }
    if (pot->kind == ORC_POT_ELJB) {
        /* lrc(ELJPotentialB), src/EnergyEvaluation.jl:983-998: fixed coefficients */
        static const double cb[4] = {-0.1279111890228638, -1.328138539967966, 12.260941135261255,
                                     41.12212408251662};
        if (r_cut <= 16.0) return 0.0;
        double rcs = sqrt(r_cut);
        double rc3 = r_cut * rcs;
        double e = 0.0;
        for (int i = 1; i <= 4; ++i) {
            e += cb[i - 1] / rc3 / (double)(2 * i + 1);
            rc3 *= r_cut;
        }
        double n2 = (double)((int64_t)n_atoms * n_atoms);
        e *= M_PI * n2 / 4.0 / (rcs * rcs * rcs);
        return e;
    }
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

/* 3^0.5 and 1 / 3^0.5 as the reference's rhombic formulas evaluate them */
#define ORC_SQRT3 1.7320508075688772
#define ORC_INV_SQRT3 (1.0 / 1.7320508075688772)

/* Squared distance, src/Configurations.jl:101-126.  Periodic boxes move b to
 * its image nearest a.  Cubic: b' = b + L round((a-b)/L) per axis;
 * rectangular: the same with the height H on z; rhombic (60 degree prism):
 * y by the row spacing sqrt(3)/2 L, x along the skewed axis, z by H.  form 0
 * divides as the reference does; form 1 multiplies by reciprocals computed
 * once per box, as the engine does. */
double orc_pair_geom(const double *a, const double *b, int bc, double l, double h, int form, double *tan_z)
{
    double bb[3] = {b[0], b[1], b[2]};
    if (bc == ORC_CUBIC || bc == ORC_RECT) {
        const double box[3] = {l, l, bc == ORC_RECT ? h : l};
        for (int c = 0; c < 3; ++c) {
            double inv = 1.0 / box[c];
            double q = form == 0 ? (a[c] - b[c]) / box[c] : (a[c] - b[c]) * inv;
            bb[c] = b[c] + box[c] * rint(q);
        }
    } else if (bc == ORC_RHOMB) {
        const double ly = ORC_SQRT3 / 2.0 * l;
        if (form == 0) {
            bb[1] = b[1] + ly * rint((a[1] - b[1]) / ly);
            bb[0] = b[0] - b[1] / ORC_SQRT3 + l * rint(((a[0] - b[0]) - ORC_INV_SQRT3 * (a[1] - b[1])) / l) +
                    ORC_INV_SQRT3 * bb[1];
            bb[2] = b[2] + h * rint((a[2] - b[2]) / h);
        } else {
            const double inv_ly = 1.0 / ly, inv_l = 1.0 / l, inv_h = 1.0 / h;
            bb[1] = b[1] + ly * rint((a[1] - b[1]) * inv_ly);
            bb[0] = b[0] - b[1] * ORC_INV_SQRT3 + l * rint(((a[0] - b[0]) - ORC_INV_SQRT3 * (a[1] - b[1])) * inv_l) +
                    ORC_INV_SQRT3 * bb[1];
            bb[2] = b[2] + h * rint((a[2] - b[2]) * inv_h);
        }
    }
    double dx = a[0] - bb[0], dy = a[1] - bb[1], dz = a[2] - bb[2];
    const double dxy2 = dx * dx + dy * dy;
    if (tan_z) *tan_z = fabs(sqrt(dxy2) / dz);
    return dxy2 + dz * dz;
}

double orc_distance2_box(const double *a, const double *b, int bc, double l, double h, int form)
{
    return orc_pair_geom(a, b, bc, l, h, form, NULL);
}

double orc_distance2(const double *a, const double *b, int bc, double box, int form)
{
    return orc_distance2_box(a, b, bc, box, box, form);
}

/* get_r_cut, src/Ensembles.jl:118-155 (squared cut-off) */
double orc_r_cut(int bc, double l, double h)
{
    if (bc == ORC_RECT) return fmin(l * l / 4.0, h * h / 4.0);
    if (bc == ORC_RHOMB) return fmin(l * l * 3.0 / 16.0, h * h / 4.0);
    return l * l / 4.0;
}

/* get_volume, src/Configurations.jl:246-256 */
double orc_volume(int bc, double l, double h)
{
    if (bc == ORC_RECT) return l * l * h;
    if (bc == ORC_RHOMB) return l * l * h * ORC_SQRT3 / 2.0;
    return l * l * l;
}

/* the tail correction of a box, src/EnergyEvaluation.jl:420, :447-448, :469-473 */
double orc_lrc_box(int n_atoms, double r_cut, const orc_pot *pot, int bc, double l, double h)
{
    const double e = orc_lrc(n_atoms, r_cut, pot);
    if (bc == ORC_RHOMB) return e * 3.0 / 4.0 * l / h;
    if (bc == ORC_RECT) return l < h ? e * l / h : e * (h * h) / (l * l);
    return e;
}

/* the volume find_hist_index bins, src/Sampling.jl:72-125: the rhombic form
 * there reads L^2 * 3 H^0.5 / 2 (Julia's 3x^0.5 is 3 * x^0.5), which this
 * restates as written */
double orc_hist_volume(int bc, double l, double h)
{
    if (bc == ORC_RECT) return l * l * h;
    if (bc == ORC_RHOMB) return l * l * (3.0 * sqrt(h)) / 2.0;
    return l * l * l;
}

/* wrap of a trial position into the box, src/MCMoves.jl:51-98 (form as in
 * orc_distance2_box) */
void orc_wrap(int bc, double l, double h, int form, double *t)
{
    if (bc == ORC_CUBIC || bc == ORC_RECT) {
        const double box[3] = {l, l, bc == ORC_RECT ? h : l};
        for (int c = 0; c < 3; ++c) {
            double inv = 1.0 / box[c];
            double q = form == 0 ? t[c] / box[c] : t[c] * inv;
            t[c] = t[c] - box[c] * rint(q);
        }
    } else if (bc == ORC_RHOMB) {
        const double w = l * ORC_SQRT3 / 2.0, qo = l * ORC_SQRT3 / 4.0;
        double k1, k2, k3;
        if (form == 0) {
            k2 = rint((t[1] - qo) / w);
            k1 = rint((t[0] - t[1] / ORC_SQRT3 - l / 2.0) / l);
            k3 = rint((t[2] - h / 2.0) / h);
        } else {
            const double inv_w = 1.0 / w, inv_l = 1.0 / l, inv_h = 1.0 / h;
            k2 = rint((t[1] - qo) * inv_w);
            k1 = rint((t[0] - t[1] * ORC_INV_SQRT3 - l / 2.0) * inv_l);
            k3 = rint((t[2] - h / 2.0) * inv_h);
        }
        const double d0 = l * k1 + l / 2.0 * k2, d1 = w * k2, d2 = h * k3;
        t[0] = t[0] - d0;
        t[1] = t[1] - d1;
        t[2] = t[2] - d2;
    }
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

/* The engine's FP64 pair-potential sweeps decide atom moves by the threshold
 * form of src/Exchange.jl:53-57: accept iff dE <= -ln(u) / beta, which equals
 * min(1, e^{-beta dE}) >= u up to the rounding of the two sides. */
static inline int metro_threshold(const orc_params *p, const orc_pot *pot)
{
    return !p->math_lib && !p->fp32 &&
           (pot->kind == ORC_POT_ELJ_EVEN || pot->kind == ORC_POT_ELJ || pot->kind == ORC_POT_ELJB ||
            pot->kind == ORC_POT_LUT);
}

/* Full configuration energy, src/EnergyEvaluation.jl:378-474: sum over i<j
 * (NPT: pairs inside r_cut, plus the box's lrc).  en_atom optional.  With
 * lanes k > 1 rows are dealt to lanes round-robin and combined by the
 * butterfly. */
double orc_energy_config_box(const orc_params *p, const orc_pot *pot, const double *x, double l, double h,
                             double r_cut, double *en_atom)
{
    const int N = p->n_atoms;
    const int k = p->lanes > 0 ? p->lanes : 1;
    double part[64] = {0};
    if (en_atom) for (int i = 0; i < N; ++i) en_atom[i] = 0.0;
    for (int i = 0; i < N; ++i) {
        int q = i % k;
        for (int j = i + 1; j < N; ++j) {
            double tz = 0.0;
            double r2 = orc_pair_geom(&x[3 * i], &x[3 * j], p->bc, l, h, dist_form(p),
                                      (pot->kind == ORC_POT_ELJB || pot->kind == ORC_POT_LUT) ? &tz : NULL);
            if (p->ensemble == ORC_NPT && !(r2 <= r_cut)) continue;
            double e = pot->kind == ORC_POT_ELJB  ? orc_dimer_energy_b(pot, r2, tz)
                       : pot->kind == ORC_POT_LUT ? orc_dimer_energy_lut(pot, r2, tz)
                                                  : orc_dimer_energy(pot, r2, p->pair_form);
            if (en_atom) { en_atom[i] += e; en_atom[j] += e; }
            part[q] += e;
        }
    }
    double tot = tree_sum(part, k);
    if (p->ensemble == ORC_NPT) tot += orc_lrc_box(N, r_cut, pot, p->bc, l, h);
    return tot;
}

double orc_energy_config(const orc_params *p, const orc_pot *pot, const double *x,
                         double box, double r_cut, double *en_atom)
{
    return orc_energy_config_box(p, pot, x, box, box, r_cut, en_atom);
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
        double h = s->box_h ? s->box_h[r] : box;
        if (p->ensemble == ORC_NPT) s->r_cut[r] = orc_r_cut(p->bc, box, h);   /* src/Ensembles.jl:118-155 */
        if (s->d2) {
            double *d2 = &s->d2[(size_t)r * N * N];
            for (int i = 0; i < N; ++i) {
                d2[i * N + i] = 0.0;
                for (int j = i + 1; j < N; ++j)
                    d2[i * N + j] = d2[j * N + i] =
                        orc_distance2_box(&x[3 * i], &x[3 * j], p->bc, box, h, dist_form(p));
            }
        }
        if (pot->kind == ORC_POT_NNP) {
            /* set_variables + initialise_energy, src/EnergyEvaluation.jl:1812-1822, :1735-1747 */
            const int nf = pot->nnp->n_rad + pot->nnp->n_ang;
            if (orc_nnp_full_g(pot->nnp, x, N, &s->g[(size_t)r * N * nf])) return -1;
            s->en_tot[r] = orc_nnp_energy(pot->nnp, &s->g[(size_t)r * N * nf], N, &s->eat[(size_t)r * N]);
        } else if (pot->kind == ORC_POT_NNP2) {
            /* set_variables + initialise_energy (RuNNerPotential2Atom),
             * src/EnergyEvaluation.jl:1824-1856, :1748-1768 */
            const int n1 = pot->nnvt_n1, ng = 2 * pot->nnp2->n_rad + 3 * pot->nnp2->n_ang;
            if (orc_nnp2_full_g(pot->nnp2, x, n1, N - n1, &s->g[(size_t)r * N * ng])) return -1;
            s->en_tot[r] = orc_nnp2_energy(pot->nnp2, &s->g[(size_t)r * N * ng], n1, N - n1, &s->eat[(size_t)r * N]);
        } else if (pot->kind == ORC_POT_EAM) {
            eam_build(p, pot, x, box, &s->comp[(size_t)r * N * 2]);
            s->en_tot[r] = eam_energy_from_comp(pot, &s->comp[(size_t)r * N * 2], N,
                                                p->math_lib ? 1 : (p->lanes > 0 ? p->lanes : 1));
        } else {
            s->en_tot[r] = orc_energy_config_box(p, pot, x, box, h,
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

/* trial position, src/MCMoves.jl:33-98: pos + (rand()-0.5)*delta per axis,
 * periodic boxes wrap it back (orc_wrap; binary32: the same steps in
 * binary32 with the box constants rounded from FP64). */
static void make_trial(const orc_params *p, const double *pos, double delta, double box, double h,
                       const draws *d, double *t)
{
    double u[3] = {d->ux, d->uy, d->uz};
    if (!p->fp32) {
        for (int c = 0; c < 3; ++c) t[c] = pos[c] + (u[c] - 0.5) * delta;
        if (p->bc != ORC_SPHERE) orc_wrap(p->bc, box, h, p->pair_form == 0 ? 0 : 1, t);
    } else {
        float fd = (float)delta;
        float v[3];
        for (int c = 0; c < 3; ++c) v[c] = (float)pos[c] + ((float)u[c] - 0.5f) * fd;
        if (p->bc == ORC_CUBIC || p->bc == ORC_RECT) {
            const double bx[3] = {box, box, p->bc == ORC_RECT ? h : box};
            for (int c = 0; c < 3; ++c) {
                float L = (float)bx[c], inv = (float)(1.0 / bx[c]);
                float q = rintf(v[c] * inv);
                v[c] = v[c] - L * q;
            }
        } else if (p->bc == ORC_RHOMB) {
            const double w = box * ORC_SQRT3 / 2.0;
            const float L = (float)box, H = (float)h, W = (float)w, qo = (float)(box * ORC_SQRT3 / 4.0);
            const float inv_w = (float)(1.0 / w), inv_l = (float)(1.0 / box), inv_h = (float)(1.0 / h);
            const float hl = (float)(box / 2.0), hh = (float)(h / 2.0), s3 = (float)ORC_INV_SQRT3;
            float k2 = rintf((v[1] - qo) * inv_w);
            float k1 = rintf((v[0] - v[1] * s3 - hl) * inv_l);
            float k3 = rintf((v[2] - hh) * inv_h);
            float d0 = L * k1 + hl * k2;
            v[0] = v[0] - d0;
            v[1] = v[1] - W * k2;
            v[2] = v[2] - H * k3;
        }
        for (int c = 0; c < 3; ++c) t[c] = (double)v[c];
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
    const double h = s->box_h ? s->box_h[r] : box;
    const double *d2 = s->d2 ? &s->d2[(size_t)r * N * N] : NULL;
    const int cutoff = p->ensemble == ORC_NPT;
    const double rc = cutoff ? s->r_cut[r] : 0.0;

    for (int j = 0; j < N; ++j)
        o->new_row[j] = j == i ? 0.0 : orc_distance2_box(trial, &x[3 * j], p->bc, box, h, dist_form(p));

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

    if (pot->kind == ORC_POT_NNP2) {
        /* energy_update!(RuNNerPotential2Atom), src/EnergyEvaluation.jl:1622-1643 */
        const int n1 = pot->nnvt_n1, ng = 2 * pot->nnp2->n_rad + 3 * pot->nnp2->n_ang;
        orc_nnp2_move_g(pot->nnp2, x, n1, N - n1, i, trial, &s->g[(size_t)r * N * ng], o->g_new);
        o->new_en = orc_nnp2_energy(pot->nnp2, o->g_new, n1, N - n1, o->eat_new);
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
    if (pot->kind == ORC_POT_ELJB || pot->kind == ORC_POT_LUT) {
        /* src/EnergyEvaluation.jl:1508-1566: new and old tan rows with the r^2 rows */
        for (int j = 0; j < N; ++j) {
            if (j == i) continue;
            int q = j % k;
            pn[q] += pair_energy_ab(p, pot, trial, &x[3 * j], box, h, 1.0, dist_form(p), rc, cutoff);
            po[q] += pair_energy_ab(p, pot, &x[3 * i], &x[3 * j], box, h, 1.0, dist_form(p), rc, cutoff);
        }
        o->new_en = (tree_sum(pn, k) - tree_sum(po, k)) + en_tot;
        return;
    }
    for (int j = 0; j < N; ++j) {
        if (j == i) continue;
        int q = j % k;
        double rn = o->new_row[j];
        double ro = d2 ? d2[i * N + j]
                       : orc_distance2_box(&x[3 * i], &x[3 * j], p->bc, box, h, dist_form(p));
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
    if (p->bc != ORC_SPHERE) SWAPN(s->box, 1);
    if (p->bc == ORC_RECT || p->bc == ORC_RHOMB) SWAPN(s->box_h, 1);
    if (p->ensemble == ORC_NPT) SWAPN(s->r_cut, 1);
    if (pot->kind == ORC_POT_EAM) SWAPN(s->comp, (size_t)N * 2);
    if (pot->kind == ORC_POT_NNP) {
        SWAPN(s->g, (size_t)N * (pot->nnp->n_rad + pot->nnp->n_ang));
        SWAPN(s->eat, (size_t)N);
    }
    if (pot->kind == ORC_POT_NNP2) {
        SWAPN(s->g, (size_t)N * (2 * pot->nnp2->n_rad + 3 * pot->nnp2->n_ang));
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
        double v1 = orc_volume(p->bc, s->box[n], s->box_h ? s->box_h[n] : s->box[n]);
        double v2 = orc_volume(p->bc, s->box[n + 1], s->box_h ? s->box_h[n + 1] : s->box[n + 1]);
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
    const int nnvt = pot->kind == ORC_POT_NNP2;
    const uint32_t nsteps = (uint32_t)(p->ensemble == ORC_NPT ? N + 1 : (nnvt ? N + pot->nnvt_swaps : N));
    trial_out *o = (trial_out *)calloc(1, sizeof(trial_out));
    double *vol_x = (double *)malloc(sizeof(double) * 3 * N);
    double *vol_d2 = (p->separated && s->d2) ? (double *)malloc(sizeof(double) * N * N) : NULL;
    if (!o || !vol_x || (p->separated && s->d2 && !vol_d2)) return -2;
    if (p->separated && (p->ensemble != ORC_NPT || (p->bc != ORC_RECT && p->bc != ORC_RHOMB) || !s->box_h ||
                         !s->max_box_h || !s->lh_ratio || !s->max_displ_sep || !s->count_vol_xy ||
                         !s->count_vol_z))
        return -4;
    if (pot->kind == ORC_POT_NNP) {
        if (p->ensemble != ORC_NVT) return -3;
        o->g_new = (double *)malloc(sizeof(double) * N * (pot->nnp->n_rad + pot->nnp->n_ang));
        o->eat_new = (double *)malloc(sizeof(double) * N);
    }
    if (nnvt) {
        const int n1 = pot->nnvt_n1;
        if (p->ensemble != ORC_NVT || n1 < 0 || n1 > N || pot->nnvt_swaps < 0 ||
            (pot->nnvt_swaps > 0 && (n1 < 1 || n1 == N)))
            return -3;
        o->g_new = (double *)malloc(sizeof(double) * N * (2 * pot->nnp2->n_rad + 3 * pot->nnp2->n_ang));
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
            double h = s->box_h ? s->box_h[r] : box;
            for (uint32_t m = 0; m < nsteps; ++m) {
                draws d0 = draw_move(p, nsteps, r, sweep, m, 0);
                const int i = (int)d0.index;
                double new_en;
                int accept;
                size_t tix = ((size_t)r * n_sweeps + c) * nsteps + m;
                if (i < N) {
                    /* atom move, src/MCMoves.jl:115-145 / :100-113 */
                    double t[3];
                    make_trial(p, &x[3 * i], s->max_displ[2 * r], box, h, &d0, t);
                    if (p->bc == ORC_SPHERE) {
                        int count = 0, attempt = 0;
                        while (outside(p, t)) {
                            count += 1;
                            if (count == 50) {
                                recentre(p, x);
                            } else {
                                attempt += 1;
                                draws dd = draw_move(p, nsteps, r, sweep, m, attempt);
                                make_trial(p, &x[3 * i], s->max_displ[2 * r], box, h, &dd, t);
                                if (count == 100) {
                                    if (s->error) *s->error = 1;
                                    free(o->g_new); free(o->eat_new); free(o); free(vol_x); free(vol_d2);
                                    return 1;
                                }
                            }
                        }
                    }
                    atom_move_energy(p, pot, s, r, i, t, s->en_tot[r], o);
                    new_en = o->new_en;
                    double dE = new_en - s->en_tot[r];
                    {
                        /* src/Exchange.jl:53-57: the oracle's own FP64 decision,
                         * traced also when the device's decision is replayed */
                        double prob = xexp(p, -dE * s->beta[r]);
                        if (prob > 1.0) prob = 1.0;
                        accept = prob >= d0.umet;
                        if (metro_threshold(p, pot)) accept = dE <= -orc_log(d0.umet) / s->beta[r];
                        if (s->trace_prob) s->trace_prob[tix] = prob;
                        if (s->trace_u) s->trace_u[tix] = d0.umet;
                        if (p->plant_decision) {
                            double bp = s->beta[r], up = d0.umet;
                            if (p->plant_decision == 1) bp = 2.0 * bp;
                            if (p->plant_decision == 2) bp = -bp;
                            if (p->plant_decision == 3) up = 1.0 - up;
                            double pp = xexp(p, -dE * bp);
                            accept = (pp > 1.0 ? 1.0 : pp) >= up;
                        }
                        if (p->fp32 && s->replay_acc) accept = s->replay_acc[tix];
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
                        if (nnvt) {
                            const size_t ng = (size_t)(2 * pot->nnp2->n_rad + 3 * pot->nnp2->n_ang);
                            memcpy(&s->g[(size_t)r * N * ng], o->g_new, sizeof(double) * N * ng);
                            memcpy(&s->eat[(size_t)r * N], o->eat_new, sizeof(double) * N);
                        }
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
                } else if (nnvt) {
                    /* atom swap: swap_atoms (src/MCMoves.jl:363-370) picks one
                     * atom of each species, swap_energy_update
                     * (src/EnergyEvaluation.jl:1438-1448), Metropolis on the
                     * energy change (src/Exchange.jl:82-83), swap_move_config!
                     * (src/swap_config.jl:238-277) on acceptance */
                    const int n1 = pot->nnvt_n1, ng = 2 * pot->nnp2->n_rad + 3 * pot->nnp2->n_ang;
                    orc_u32x4 S = orc_draw(p->seed, p->run, ORC_T_MOVE, 254u, r, sweep, m);
                    const int i1 = (int)orc_below(S.v[0], (uint32_t)n1);
                    const int i2 = n1 + (int)orc_below(S.v[1], (uint32_t)(N - n1));
                    orc_nnp2_swap_g(pot->nnp2, x, n1, N - n1, i1, i2, &s->g[(size_t)r * N * ng], o->g_new);
                    new_en = orc_nnp2_energy(pot->nnp2, o->g_new, n1, N - n1, o->eat_new);
                    const double dE = new_en - s->en_tot[r];
                    {
                        /* src/Exchange.jl:53-57: the oracle's own FP64 decision,
                         * traced also when the device's decision is replayed */
                        double prob = xexp(p, -dE * s->beta[r]);
                        if (prob > 1.0) prob = 1.0;
                        accept = prob >= d0.umet;
                        if (metro_threshold(p, pot)) accept = dE <= -orc_log(d0.umet) / s->beta[r];
                        if (s->trace_prob) s->trace_prob[tix] = prob;
                        if (s->trace_u) s->trace_u[tix] = d0.umet;
                        if (p->fp32 && s->replay_acc) accept = s->replay_acc[tix];
                    }
                    if (s->trace_de) s->trace_de[tix] = dE;
                    if (s->trace_acc) s->trace_acc[tix] = (int8_t)accept;
                    if (s->trace_idx) s->trace_idx[tix] = N + i1 * N + i2;
                    s->count_vol[2 * r + 1] += 1;
                    if (accept) {
                        for (int k = 0; k < 3; ++k) {
                            const double t_ = x[3 * i1 + k];
                            x[3 * i1 + k] = x[3 * i2 + k];
                            x[3 * i2 + k] = t_;
                        }
                        if (s->d2) {
                            double *d2 = &s->d2[(size_t)r * N * N];
                            for (int pass = 0; pass < 2; ++pass)  /* rows, then columns */
                                for (int j = 0; j < N; ++j) {
                                    double *u = pass ? &d2[(size_t)j * N + i1] : &d2[(size_t)i1 * N + j];
                                    double *v = pass ? &d2[(size_t)j * N + i2] : &d2[(size_t)i2 * N + j];
                                    const double t_ = *u;
                                    *u = *v;
                                    *v = t_;
                                }
                            d2[(size_t)i1 * N + i1] = d2[(size_t)i2 * N + i2] = 0.0;
                        }
                        memcpy(&s->g[(size_t)r * N * ng], o->g_new, sizeof(double) * N * ng);
                        memcpy(&s->eat[(size_t)r * N], o->eat_new, sizeof(double) * N);
                        s->en_tot[r] = new_en;
                        s->count_vol[2 * r] += 1;
                    }
                } else {
                    /* volume move: uniform, src/MCMoves.jl:169-177, :277-287, or
                     * separated (z 1/6, xy 1/3, xyz 1/2 of the moves),
                     * src/MCMoves.jl:235-343 */
                    int kind = 0; /* 0: xyz, 1: xy, 2: z (xy_or_z) */
                    double sx, sz;
                    if (!p->separated) {
                        double y = (d0.ux - 0.5) * s->max_displ[2 * r + 1];
                        double scale = p->math_lib ? pow(exp(y), 1.0 / 3.0) : orc_exp(y / 3.0);
                        if (box >= s->max_box[r] && scale > 1.0) scale = 1.0;
                        sx = sz = scale;
                    } else {
                        const int ra = 1 + (int)(d0.uy * 6.0);
                        const double lh = box / h, lr = s->lh_ratio[r];
                        if (ra == 1) {
                            kind = 2;
                            double y = (d0.ux - 0.5) * s->max_displ_sep[2 * r + 1];
                            double scale = p->math_lib ? exp(y) : orc_exp(y);
                            if (lh <= lr * 1.1 && scale > 1.0) scale = 1.0 / scale;
                            else if (lh >= lr * 0.909 && scale < 1.0) scale = 1.0 / scale;
                            if (h >= s->max_box_h[r] && scale > 1.0) scale = 1.0 / scale;
                            sx = 1.0;
                            sz = scale;
                        } else if (ra <= 3) {
                            kind = 1;
                            double y = (d0.ux - 0.5) * s->max_displ_sep[2 * r];
                            double scale = p->math_lib ? pow(exp(y), 0.5) : orc_exp(y / 2.0);
                            if (lh >= lr * 1.1 && scale > 1.0) scale = 1.0 / scale;
                            else if (lh <= lr * 0.909 && scale < 1.0) scale = 1.0 / scale;
                            if (box >= s->max_box[r] && scale > 1.0) scale = 1.0 / scale;
                            sx = scale;
                            sz = 1.0;
                        } else {
                            double y = (d0.ux - 0.5) * s->max_displ[2 * r + 1];
                            double scale = p->math_lib ? pow(exp(y), 1.0 / 3.0) : orc_exp(y / 3.0);
                            if (box >= s->max_box[r] && scale > 1.0) scale = 1.0;
                            sx = sz = scale;
                        }
                    }
                    double nbox = box * sx;
                    double nh = h * sz;
                    double nrc = orc_r_cut(p->bc, nbox, nh);
                    /* energy of the trial configuration, src/EnergyEvaluation.jl:397-474:
                     * uniform moves scale the r^2 matrix, separated moves
                     * recompute it from the scaled positions in the new box */
                    double part[64] = {0};
                    const int k = p->lanes > 0 ? p->lanes : 1;
                    if (!p->separated) {
                        const double s2 = sx * sx;
                        for (int a = 0; a < N; ++a) {
                            for (int b = a + 1; b < N; ++b) {
                                if (pot->kind == ORC_POT_ELJB || pot->kind == ORC_POT_LUT) {
                                    part[a % k] += pair_energy_ab(p, pot, &x[3 * a], &x[3 * b], box, h, s2,
                                                                  dist_form(p), nrc, 1);
                                    continue;
                                }
                                double r2 = s->d2 ? s->d2[(size_t)r * N * N + a * N + b]
                                                  : orc_distance2_box(&x[3 * a], &x[3 * b], p->bc, box, h,
                                                                      dist_form(p));
                                r2 = r2 * s2;
                                if (r2 <= nrc) part[a % k] += orc_dimer_energy(pot, r2, p->pair_form);
                            }
                        }
                    } else {
                        for (int a = 0; a < N; ++a) {
                            vol_x[3 * a] = x[3 * a] * sx;
                            vol_x[3 * a + 1] = x[3 * a + 1] * sx;
                            vol_x[3 * a + 2] = x[3 * a + 2] * sz;
                        }
                        for (int a = 0; a < N; ++a) {
                            if (vol_d2) vol_d2[a * N + a] = 0.0;
                            for (int b = a + 1; b < N; ++b) {
                                double tz = 0.0;
                                const int ang = pot->kind == ORC_POT_ELJB || pot->kind == ORC_POT_LUT;
                                double r2 = orc_pair_geom(&vol_x[3 * a], &vol_x[3 * b], p->bc, nbox, nh,
                                                          dist_form(p), ang ? &tz : NULL);
                                if (vol_d2) vol_d2[a * N + b] = vol_d2[b * N + a] = r2;
                                if (r2 <= nrc)
                                    part[a % k] += pot->kind == ORC_POT_ELJB  ? orc_dimer_energy_b(pot, r2, tz)
                                                   : pot->kind == ORC_POT_LUT ? orc_dimer_energy_lut(pot, r2, tz)
                                                                              : orc_dimer_energy(pot, r2, p->pair_form);
                            }
                        }
                    }
                    new_en = tree_sum(part, k) + orc_lrc_box(N, nrc, pot, p->bc, nbox, nh);
                    double dE = new_en - s->en_tot[r];
                    double v_new = orc_volume(p->bc, nbox, nh), v_old = orc_volume(p->bc, box, h);
                    double dh = dE + p->pressure * (v_new - v_old);
                    double arg = -dh * s->beta[r] + (double)N * xlog(p, v_new / v_old);
                    double prob = xexp(p, arg);
                    if (prob > 1.0) prob = 1.0;
                    accept = prob >= d0.umet;
                    if (s->trace_prob) s->trace_prob[tix] = prob;
                    if (s->trace_u) s->trace_u[tix] = d0.umet;
                    if (p->fp32 && s->replay_acc) accept = s->replay_acc[tix];
                    if (s->trace_de) s->trace_de[tix] = dE;
                    if (s->trace_acc) s->trace_acc[tix] = (int8_t)accept;
                    if (s->trace_idx) s->trace_idx[tix] = i;
                    if (accept) {
                        /* swap_config_v!, src/swap_config.jl:113-140 */
                        for (int a = 0; a < N; ++a) {
                            const double f[3] = {sx, sx, sz};
                            for (int c = 0; c < 3; ++c) {
                                double v = x[3 * a + c] * f[c];
                                x[3 * a + c] = p->fp32 ? (double)(float)v : v;
                            }
                        }
                        if (s->d2) {
                            if (!p->separated) {
                                const double s2 = sx * sx;
                                for (int a = 0; a < N * N; ++a) s->d2[(size_t)r * N * N + a] *= s2;
                            } else {
                                memcpy(&s->d2[(size_t)r * N * N], vol_d2, sizeof(double) * N * N);
                            }
                        }
                        s->box[r] = nbox;
                        box = nbox;
                        if (s->box_h) s->box_h[r] = nh;
                        h = nh;
                        s->r_cut[r] = nrc;
                        s->en_tot[r] = new_en;
                        int64_t *cnt = kind == 0 ? s->count_vol : (kind == 1 ? s->count_vol_xy : s->count_vol_z);
                        cnt[2 * r] += 1;
                        cnt[2 * r + 1] += 1;
                    }
                }
            }
        }

        /* exchange, src/MCRun.jl:203-205 (attempted every exchange_every cycles) */
        if (R > 1 && (p->exchange_every <= 1 || (sweep + 1) % (uint32_t)p->exchange_every == 0)) {
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
                if (p->ensemble == ORC_NPT && !p->separated) {
                    double av = (double)s->count_vol[2 * r + 1] / (double)((int64_t)p->n_adjust * 1);
                    if (av < p->min_acc) s->max_displ[2 * r + 1] *= 0.9;
                    else if (av > p->max_acc) s->max_displ[2 * r + 1] *= 1.1;
                    s->count_vol[2 * r + 1] = 0;
                } else if (p->ensemble == ORC_NPT) {
                    /* separated: expected shares 1/2, 1/3, 1/6 of the volume moves */
                    double av = (double)s->count_vol[2 * r + 1] / ((double)p->n_adjust / 2.0);
                    if (av < p->min_acc) s->max_displ[2 * r + 1] *= 0.9;
                    else if (av > p->max_acc) s->max_displ[2 * r + 1] *= 1.1;
                    s->count_vol[2 * r + 1] = 0;
                    av = (double)s->count_vol_xy[2 * r + 1] / ((double)p->n_adjust / 3.0);
                    if (av < p->min_acc) s->max_displ_sep[2 * r] *= 0.9;
                    else if (av > p->max_acc) s->max_displ_sep[2 * r] *= 1.1;
                    s->count_vol_xy[2 * r + 1] = 0;
                    av = (double)s->count_vol_z[2 * r + 1] / ((double)p->n_adjust / 6.0);
                    if (av < p->min_acc) s->max_displ_sep[2 * r + 1] *= 0.9;
                    else if (av > p->max_acc) s->max_displ_sep[2 * r + 1] *= 1.1;
                    s->count_vol_z[2 * r + 1] = 0;
                }
            }
        }

        /* sampling_step!, src/Sampling.jl:330-371 */
        if (p->sampling && cyc % p->mc_sample == 0) {
            const int nb2 = p->n_bin + 2;
            for (int r = 0; r < R; ++r) {
                double e = s->en_tot[r];
                if (p->ensemble == ORC_NPT) {
                    const double bh = s->box_h ? s->box_h[r] : s->box[r];
                    double v = orc_volume(p->bc, s->box[r], bh);
                    double h = e + p->pressure * v;
                    s->ham[2 * r] += h;
                    s->ham[2 * r + 1] += h * h;
                    if (s->hist) {
                        double vh = orc_hist_volume(p->bc, s->box[r], bh);
                        double ie = floor((e - p->en_min) / p->delta_en_hist) + 1.0;
                        double iv = floor((vh - p->v_min) / p->delta_v_hist) + 1.0;
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
    free(vol_d2);
    return 0;
}

/* ------------------------------------------- the potential plugin functions */

/* get_distance2_mat (src/Configurations.jl:135-162): each pair i < j once,
 * distance2(config[i], config[j]), mirrored; zero diagonal. */
void orc_distance2_mat(const orc_params *p, const double *x, double l, double h, double *d2)
{
    const int N = p->n_atoms;
    for (int i = 0; i < N; ++i) {
        d2[(size_t)i * N + i] = 0.0;
        for (int j = i + 1; j < N; ++j)
            d2[(size_t)i * N + j] = d2[(size_t)j * N + i] =
                orc_distance2_box(&x[3 * i], &x[3 * j], p->bc, l, h, dist_form(p));
    }
}

/* dimer_energy_config from a squared-distance matrix (src/EnergyEvaluation.jl:
 * 378-474), ELJ potentials: pairs i < j in order, the pair energy added to
 * both atoms' rows and to the total; NPT skips pairs beyond r_cut and adds
 * the box's tail correction. */
double orc_energy_config_d2(const orc_params *p, const orc_pot *pot, const double *d2, double r_cut, double l,
                            double h, double *en_atom)
{
    const int N = p->n_atoms, cut = p->ensemble == ORC_NPT;
    double tot = 0.0;
    if (en_atom) for (int i = 0; i < N; ++i) en_atom[i] = 0.0;
    for (int i = 0; i < N; ++i)
        for (int j = i + 1; j < N; ++j) {
            const double r2 = d2[(size_t)i * N + j];
            if (cut && !(r2 <= r_cut)) continue;
            const double e = orc_dimer_energy(pot, r2, p->pair_form);
            if (en_atom) { en_atom[i] += e; en_atom[j] += e; }
            tot += e;
        }
    if (cut) tot += orc_lrc_box(N, r_cut, pot, p->bc, l, h);
    return tot;
}

/* set_variables(EmbeddedAtomPotential) (src/EnergyEvaluation.jl:1800-1811):
 * calc_components (:1105-1110) over each row of the squared-distance matrix,
 * diagonal included (invrexp(0) = (0, 0)). */
void orc_eam_components(const orc_params *p, const orc_pot *pot, const double *d2, double *comp)
{
    const int N = p->n_atoms;
    for (int i = 0; i < N; ++i) {
        double a = 0.0, b = 0.0;
        for (int j = 0; j < N; ++j) {
            double f, g;
            eam_terms(p, pot, d2[(size_t)i * N + j], &f, &g);
            a += f;
            b += g;
        }
        comp[2 * i] = a;
        comp[2 * i + 1] = b;
    }
}

/* calc_energies_from_components (src/EnergyEvaluation.jl:1156-1162) */
double orc_eam_energy(const orc_pot *pot, const double *comp, int N)
{
    return eam_energy_from_comp(pot, comp, N, 1);
}

/* energy_update! (src/EnergyEvaluation.jl:1474-1643) of one configuration x
 * (box l, h) whose atom i moves to trial: the move energy of the sweep above
 * for a one-replica state.  d2: the configuration's squared-distance matrix
 * (NULL: distances from x); comp: EAM components (N x 2); g: network
 * features (atom-major).  Writes new_comp (EAM), g_new and eat_new (network;
 * g_new is a copy of g when the hard-core test gives 100).  Returns the new
 * total energy. */
double orc_energy_update(const orc_params *p, const orc_pot *pot, const double *x, const double *d2, double l,
                         double h, int i, const double *trial, double en_tot, const double *comp, const double *g,
                         double *new_comp, double *g_new, double *eat_new)
{
    const int N = p->n_atoms;
    if (N < 2 || N > MAXN || i < 0 || i >= N) return NAN;
    orc_state s;
    memset(&s, 0, sizeof s);
    double box = l, boxh = h, rc = p->ensemble == ORC_NPT ? orc_r_cut(p->bc, l, h) : 0.0;
    s.x = (double *)x;
    s.box = &box;
    s.box_h = &boxh;
    s.r_cut = &rc;
    s.d2 = (double *)d2;
    s.comp = (double *)comp;
    s.g = (double *)g;
    trial_out *o = (trial_out *)malloc(sizeof *o);
    if (!o) return NAN;
    o->g_new = g_new;
    o->eat_new = eat_new;
    if (pot->kind == ORC_POT_NNP && g && g_new)
        memcpy(g_new, g, sizeof(double) * (size_t)N * (pot->nnp->n_rad + pot->nnp->n_ang));
    atom_move_energy(p, pot, &s, 0, i, trial, en_tot, o);
    if (pot->kind == ORC_POT_EAM && new_comp) memcpy(new_comp, o->new_comp, sizeof(double) * 2 * N);
    const double e = o->new_en;
    free(o);
    return e;
}
