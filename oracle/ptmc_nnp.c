/*
 * ORACLE — test infrastructure only (see orc_math.h).  FP64 restatement of
 * the reference's single-species RuNNer neural-network potential:
 *   cutoff            MachineLearningPotential/Cutoff.jl:79-90
 *   angles            MachineLearningPotential/Cutoff.jl:31-68
 *   G2 / G4 values    MachineLearningPotential/SymmFunc.jl:184-312
 *   full G matrix     MachineLearningPotential/SymmFunc.jl:494-510
 *   moved-atom update MachineLearningPotential/DeltaMatrix.jl:639-975
 *   forward pass      the RuNNer Fortran `forward` behind
 *                     MachineLearningPotential/ForwardPass.jl:60-84, restated
 *                     from the network's definition (input-major weight
 *                     blocks, each followed by its layer's biases), pinned by
 *                     test/test_runner_forward.jl:73 and the reference's
 *                     Cu55 energies in test/testing_data/save.data.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ptmc_oracle.h"

double orc_forward(const double *input, const int32_t *n_in, const int32_t *n_layers, const int32_t *nodes,
                   const int32_t *batch, const int32_t *n_params, const double *params, const int32_t *act,
                   double *eatom)
{
    const int L = *n_layers, B = *batch;
    if (L < 2 || *n_in != nodes[0] || B < 1) return NAN;
    int need = 0, wmax = 0;
    for (int l = 1; l < L; ++l) {
        need += nodes[l - 1] * nodes[l] + nodes[l];
        if (nodes[l] > wmax) wmax = nodes[l];
    }
    if (nodes[0] > wmax) wmax = nodes[0];
    if (need != *n_params || nodes[L - 1] != 1) return NAN;
    double *a = (double *)malloc(sizeof(double) * wmax), *z = (double *)malloc(sizeof(double) * wmax);
    for (int b = 0; b < B; ++b) {
        memcpy(a, input + (size_t)b * nodes[0], sizeof(double) * nodes[0]);
        const double *p = params;
        for (int l = 1; l < L; ++l) {
            const int ni = nodes[l - 1], no = nodes[l];
            const double *w = p, *bias = p + (size_t)ni * no;
            for (int j = 0; j < no; ++j) {
                double s = 0.0;
                for (int i = 0; i < ni; ++i) s += a[i] * w[(size_t)i * no + j];
                s += bias[j];
                z[j] = act[l] == 2 ? tanh(s) : s;
            }
            memcpy(a, z, sizeof(double) * no);
            p = bias + no;
        }
        eatom[b] = a[0];
    }
    free(a);
    free(z);
    return eatom[B - 1];
}

double orc_cutoff(double r, double r_cut)
{
    return r < r_cut ? 0.5 * (cos(M_PI * (r / r_cut)) + 1.0) : 0.0;
}

static double d2(const double *a, const double *b)
{
    const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return dx * dx + dy * dy + dz * dz;
}

/* thetacalc(xy, xz, |xy|, |xz|) = xy . xz / (|xy| |xz|) */
static double thetacalc(const double *xy, const double *xz, double dxy, double dxz)
{
    return (xy[0] * xz[0] + xy[1] * xz[1] + xy[2] * xz[2]) / (dxy * dxz);
}

/* all_angular_measure(a, b, c, r2ab, r2ac, r2bc): cosines at a, b and c */
static void angles(const double *a, const double *b, const double *c, double r2ab, double r2ac, double r2bc,
                   double th[3])
{
    double ab[3], ac[3], bc[3], mab[3], mac[3], mbc[3];
    for (int k = 0; k < 3; ++k) {
        ab[k] = a[k] - b[k];
        ac[k] = a[k] - c[k];
        bc[k] = b[k] - c[k];
        mab[k] = -ab[k];
        mac[k] = -ac[k];
        mbc[k] = -bc[k];
    }
    const double dab = sqrt(r2ab), dac = sqrt(r2ac), dbc = sqrt(r2bc);
    th[0] = thetacalc(ab, ac, dab, dac);
    th[1] = thetacalc(mab, bc, dab, dbc);
    th[2] = thetacalc(mac, mbc, dac, dbc);
}

#define RAD(g, k) nn->rad[(g) * 4 + (k)]
#define ANG(g, k) nn->ang[(g) * 6 + (k)]

int orc_nnp_full_g(const orc_nnp *nn, const double *x, int N, double *g)
{
    const int nf = nn->n_rad + nn->n_ang;
    double *r2 = (double *)malloc(sizeof(double) * N * N), *f = (double *)malloc(sizeof(double) * N * N);
    if (!r2 || !f) return -1;
    for (int a = 0; a < N; ++a)
        for (int b = 0; b < N; ++b) {
            r2[a * N + b] = d2(&x[3 * a], &x[3 * b]);
            f[a * N + b] = orc_cutoff(sqrt(r2[a * N + b]), nn->r_cut);
        }
    memset(g, 0, sizeof(double) * nf * N);
    double *row = (double *)malloc(sizeof(double) * N);
    /* radial rows, calc_symm_vals!(RadialType2) SymmFunc.jl:259-278 */
    for (int q = 0; q < nn->n_rad; ++q) {
        if (RAD(q, 3) == 0.0) continue;  /* inactive type: zeros, no offset */
        const double eta = RAD(q, 0), norm = RAD(q, 1), off = RAD(q, 2);
        for (int a = 0; a < N; ++a) row[a] = 0.0;
        for (int a = 0; a < N; ++a)
            for (int b = a + 1; b < N; ++b) {
                const double fc = f[a * N + b];
                const double v = (fc != 0.0 && fc != 1.0) ? fc * exp(-eta * r2[a * N + b]) * norm : 0.0;
                row[a] += v;
                row[b] += v;
            }
        for (int a = 0; a < N; ++a) g[a * nf + q] = row[a] + off;
    }
    /* angular rows, SymmFunc.jl:280-312; total_symm_calc runs g_index over
     * (Nrad+1):Nang, so with q10_compat the last n_rad rows stay zero */
    const int n_ang_built = nn->q10_compat ? (nn->n_ang - nn->n_rad > 0 ? nn->n_ang - nn->n_rad : 0) : nn->n_ang;
    for (int q = 0; q < n_ang_built; ++q) {
        if (ANG(q, 5) == 0.0) continue;
        const double eta = ANG(q, 0), lam = ANG(q, 1), zeta = ANG(q, 2), tpz = ANG(q, 3), off = ANG(q, 4);
        for (int a = 0; a < N; ++a) row[a] = 0.0;
        for (int a = 0; a < N; ++a)
            for (int b = a + 1; b < N; ++b)
                for (int c = b + 1; c < N; ++c) {
                    double th[3];
                    const double rab = r2[a * N + b], rac = r2[a * N + c], rbc = r2[b * N + c];
                    angles(&x[3 * a], &x[3 * b], &x[3 * c], rab, rac, rbc, th);
                    const double e = exp(-eta * (rab + rac + rbc)) * f[a * N + b] * f[a * N + c] * f[b * N + c];
                    row[a] += e * pow(1.0 + lam * th[0], zeta) * tpz;
                    row[b] += e * pow(1.0 + lam * th[1], zeta) * tpz;
                    row[c] += e * pow(1.0 + lam * th[2], zeta) * tpz;
                }
        for (int a = 0; a < N; ++a) g[a * nf + nn->n_rad + q] = row[a] + off;
    }
    free(row);
    free(r2);
    free(f);
    return 0;
}

int orc_nnp_move_g(const orc_nnp *nn, const double *x, int N, int i, const double *trial, const double *g,
                   double *g_new)
{
    const int nf = nn->n_rad + nn->n_ang;
    memcpy(g_new, g, sizeof(double) * nf * N);
    double *rn = (double *)malloc(sizeof(double) * N), *fn = (double *)malloc(sizeof(double) * N);
    double *ro = (double *)malloc(sizeof(double) * N), *fo = (double *)malloc(sizeof(double) * N);
    for (int j = 0; j < N; ++j) {
        rn[j] = j == i ? 0.0 : d2(trial, &x[3 * j]);
        ro[j] = j == i ? 0.0 : d2(&x[3 * i], &x[3 * j]);
        fn[j] = orc_cutoff(sqrt(rn[j]), nn->r_cut);
        fo[j] = orc_cutoff(sqrt(ro[j]), nn->r_cut);
    }
    /* radial_symmetry_calculation!, DeltaMatrix.jl:794-822 */
    for (int q = 0; q < nn->n_rad; ++q) {
        if (RAD(q, 3) == 0.0) continue;
        const double eta = RAD(q, 0), norm = RAD(q, 1);
        for (int j = 0; j < N; ++j) {
            if (j == i) continue;
            double gi = g_new[i * nf + q], gj = g_new[j * nf + q];
            const double vn = exp(-eta * rn[j]) * (fn[j] * norm);
            gi += vn;
            gj += vn;
            const double vo = exp(-eta * ro[j]) * (-fo[j] * norm);
            gi += vo;
            gj += vo;
            g_new[i * nf + q] = gi;
            g_new[j * nf + q] = gj;
        }
    }
    /* angular_symmetry_calculation!, DeltaMatrix.jl:827-872 */
    for (int q = 0; q < nn->n_ang; ++q) {
        if (ANG(q, 5) == 0.0) continue;
        const double eta = ANG(q, 0), lam = ANG(q, 1), zeta = ANG(q, 2), tpz = ANG(q, 3);
        const int col = nn->n_rad + q;
        for (int j = 0; j < N; ++j) {
            if (j == i) continue;
            for (int k = j + 1; k < N; ++k) {
                if (k == i) continue;
                const double rjk = d2(&x[3 * j], &x[3 * k]);
                const double fjk = orc_cutoff(sqrt(rjk), nn->r_cut);
                double tn[3], to[3];
                angles(trial, &x[3 * j], &x[3 * k], rn[j], rn[k], rjk, tn);
                angles(&x[3 * i], &x[3 * j], &x[3 * k], ro[j], ro[k], rjk, to);
                const double en = exp(-eta * (rn[j] + rn[k] + rjk)) * fn[j] * fn[k] * fjk;
                const double eo = exp(-eta * (ro[j] + ro[k] + rjk)) * fo[j] * fo[k] * fjk;
                const int idx[3] = {i, j, k};
                for (int a = 0; a < 3; ++a) {
                    double v = g_new[idx[a] * nf + col];
                    v += en * pow(1.0 + lam * tn[a], zeta) * tpz;
                    v -= eo * pow(1.0 + lam * to[a], zeta) * tpz;
                    g_new[idx[a] * nf + col] = v;
                }
            }
        }
    }
    free(rn);
    free(fn);
    free(ro);
    free(fo);
    return 0;
}

double orc_nnp_energy(const orc_nnp *nn, const double *g, int N, double *eatom)
{
    const int32_t nin = nn->n_rad + nn->n_ang, batch = N;
    orc_forward(g, &nin, &nn->n_layers, nn->nodes, &batch, &nn->n_params, nn->params, nn->act, eatom);
    double e = 0.0;
    for (int a = 0; a < N; ++a) e += eatom[a];
    return e;
}

/* ---- two species (RuNNerPotential2Atom, src/EnergyEvaluation.jl:1276-1305,
 * :1748-1768, :1822-1856; MachineLearningPotential/SymmFunc.jl:54-175,
 * :313-547).  Atoms 0..n1-1 are species 1, n1..N-1 species 2.  Each radial
 * function fills two rows (neighbours of species 1, of species 2), each
 * angular function three (neighbour pairs 11, 12, 22); rows are scaled per
 * centre species: radial g * norm + offset (4 pairs), angular g * tpz +
 * offset (6 pairs).  The loops and the order in which each row accumulates
 * are the reference's: pair / triple blocks 11(1), 22(2), 12 for radial and
 * 111, 222, 112, 122 for angular. */
#define RAD2(q, k) nn->rad[(q) * 9 + (k)]
#define ANG2(q, k) nn->ang[(q) * 15 + (k)]

int orc_nnp2_full_g(const orc_nnp2 *nn, const double *x, int n1, int n2, double *g)
{
    const int N = n1 + n2, Ng = 2 * nn->n_rad + 3 * nn->n_ang;
    double *r2 = (double *)malloc(sizeof(double) * N * N), *f = (double *)malloc(sizeof(double) * N * N);
    if (!r2 || !f) return -1;
    for (int a = 0; a < N; ++a)
        for (int b = 0; b < N; ++b) {
            r2[a * N + b] = d2(&x[3 * a], &x[3 * b]);
            f[a * N + b] = orc_cutoff(sqrt(r2[a * N + b]), nn->r_cut);
        }
    memset(g, 0, sizeof(double) * Ng * N);
#define G(row, atom) g[(size_t)(atom) * Ng + (row)]
    for (int q = 0; q < nn->n_rad; ++q) {
        const int r0 = 2 * q;
        const double eta = RAD2(q, 0);
        /* calc_symm_vals! radial (SymmFunc.jl:319-361) */
#define RV(a, b) ((f[(a) * N + (b)] != 0.0 && f[(a) * N + (b)] != 1.0) ? f[(a) * N + (b)] * exp(-eta * r2[(a) * N + (b)]) : 0.0)
        for (int a = 0; a < n1; ++a)
            for (int b = a + 1; b < n1; ++b) {
                const double v = RV(a, b);
                G(r0, a) += v;
                G(r0, b) += v;
            }
        for (int a = n1; a < N; ++a)
            for (int b = a + 1; b < N; ++b) {
                const double v = RV(a, b);
                G(r0 + 1, a) += v;
                G(r0 + 1, b) += v;
            }
        for (int a = 0; a < n1; ++a)
            for (int b = n1; b < N; ++b) {
                const double v = RV(a, b);
                G(r0 + 1, a) += v;
                G(r0, b) += v;
            }
#undef RV
        for (int a = 0; a < N; ++a) {
            const int s = a < n1 ? 0 : 2;
            G(r0, a) = G(r0, a) * RAD2(q, 1 + s) + RAD2(q, 5 + s);
            G(r0 + 1, a) = G(r0 + 1, a) * RAD2(q, 2 + s) + RAD2(q, 6 + s);
        }
    }
    for (int q = 0; q < nn->n_ang; ++q) {
        const int r0 = 2 * nn->n_rad + 3 * q;
        const double eta = ANG2(q, 0), lam = ANG2(q, 1), zeta = ANG2(q, 2);
        /* calc_symm_vals! angular (SymmFunc.jl:363-468): the block of the
         * triple (a < b < c) decides the row of each centre */
        for (int blk = 0; blk < 4; ++blk) {
            int a0, a1, ra, rb, rc_;
            /* 111: rows 0,0,0; 222: 2,2,2; 112 (a, b in 1, c in 2): 1,1,0; 122: 2,1,1 */
            if (blk == 0) { a0 = 0; a1 = n1; ra = 0; rb = 0; rc_ = 0; }
            else if (blk == 1) { a0 = n1; a1 = N; ra = 2; rb = 2; rc_ = 2; }
            else if (blk == 2) { a0 = 0; a1 = n1; ra = 1; rb = 1; rc_ = 0; }
            else { a0 = 0; a1 = n1; ra = 2; rb = 1; rc_ = 1; }
            for (int a = a0; a < a1; ++a) {
                int blo, bhi;
                if (blk == 0) { blo = a + 1; bhi = n1; }
                else if (blk == 1) { blo = a + 1; bhi = N; }
                else if (blk == 2) { blo = a + 1; bhi = n1; }
                else { blo = n1; bhi = N; }
                for (int b = blo; b < bhi; ++b) {
                    int clo, chi;
                    if (blk == 0) { clo = b + 1; chi = n1; }
                    else if (blk == 1) { clo = b + 1; chi = N; }
                    else if (blk == 2) { clo = n1; chi = N; }
                    else { clo = b + 1; chi = N; }
                    for (int c = clo; c < chi; ++c) {
                        double th[3];
                        const double rab = r2[a * N + b], rac = r2[a * N + c], rbc = r2[b * N + c];
                        angles(&x[3 * a], &x[3 * b], &x[3 * c], rab, rac, rbc, th);
                        const double e = exp(-eta * (rab + rac + rbc)) * f[a * N + b] * f[a * N + c] * f[b * N + c];
                        G(r0 + ra, a) += e * pow(1.0 + lam * th[0], zeta);
                        G(r0 + rb, b) += e * pow(1.0 + lam * th[1], zeta);
                        G(r0 + rc_, c) += e * pow(1.0 + lam * th[2], zeta);
                    }
                }
            }
        }
        for (int a = 0; a < N; ++a) {
            const int s = a < n1 ? 0 : 3;
            for (int k = 0; k < 3; ++k) G(r0 + k, a) = G(r0 + k, a) * ANG2(q, 3 + s + k) + ANG2(q, 9 + s + k);
        }
    }
#undef G
    free(r2);
    free(f);
    return 0;
}

double orc_nnp2_energy(const orc_nnp2 *nn, const double *g, int n1, int n2, double *eatom)
{
    const int Ng = 2 * nn->n_rad + 3 * nn->n_ang;
    if (nn->nodes[0] != Ng) return NAN;
    double tot = 0.0;
    if (n1 > 0) orc_forward(g, &nn->nodes[0], &nn->n_layers, nn->nodes, &n1, &nn->n_params, nn->params1, nn->act, eatom);
    if (n2 > 0)
        orc_forward(g + (size_t)n1 * Ng, &nn->nodes[0], &nn->n_layers, nn->nodes, &n2, &nn->n_params, nn->params2,
                    nn->act, eatom + n1);
    for (int a = 0; a < n1 + n2; ++a) tot += eatom[a];
    return tot;
}

/* ---------------------------------------------------------------------------
 * Two-species incremental updates.  Both return g_new = g + delta with the
 * delta built in the reference's loop order, scaled per centre species after
 * the loops, as the reference does; r^2 and the cutoffs come from the
 * positions (the reference's cached dist2_mat / f_matrix hold the same
 * values: a squared difference does not depend on the operand order). */

/* exponential_part(eta, r2, f) (SymmFunc.jl:187) */
static double expo1(double eta, double r2, double f) { return exp(-eta * r2) * f; }

/* symmfunc_calc of three given cosines (SymmFunc.jl:184-202) */
static void sym3(const double th[3], double r2ij, double r2ik, double r2jk, double fij, double fik, double fjk,
                 double eta, double lam, double zeta, double out[3])
{
    const double e = exp(-eta * (r2ij + r2ik + r2jk)) * fij * fik * fjk;
    for (int c = 0; c < 3; ++c) out[c] = e * pow(1.0 + lam * th[c], zeta);
}

/* calc_one_symm_val(position1, position2, position3, ...) (SymmFunc.jl:233-241) */
static void one3(const double *a, const double *b, const double *c, double r2ab, double r2ac, double r2bc,
                 double fab, double fac, double fbc, double eta, double lam, double zeta, double out[3])
{
    double th[3];
    angles(a, b, c, r2ab, r2ac, r2bc, th);
    sym3(th, r2ab, r2ac, r2bc, fab, fac, fbc, eta, lam, zeta, out);
}

/* calc_one_symm_val(r2, f, eta) (SymmFunc.jl:229-231): zero at f = 0 or 1 */
static double one1(double r2, double f, double eta) { return (f != 0.0 && f != 1.0) ? f * exp(-eta * r2) : 0.0; }

static int pair_tables(const orc_nnp2 *nn, const double *x, int N, double **r2, double **f)
{
    *r2 = (double *)malloc(sizeof(double) * N * N);
    *f = (double *)malloc(sizeof(double) * N * N);
    if (!*r2 || !*f) return -1;
    for (int a = 0; a < N; ++a)
        for (int b = 0; b < N; ++b) {
            (*r2)[a * N + b] = d2(&x[3 * a], &x[3 * b]);
            (*f)[a * N + b] = orc_cutoff(sqrt((*r2)[a * N + b]), nn->r_cut);
        }
    return 0;
}

/* the per-species scaling of a delta block (DeltaMatrix.jl:110-116, :208-215) */
static void scale_rows(double *dg, int Ng, int n1, int N, int r0, int nrow, const double *s1, const double *s2)
{
    for (int a = 0; a < N; ++a)
        for (int k = 0; k < nrow; ++k) dg[(size_t)a * Ng + r0 + k] *= a < n1 ? s1[k] : s2[k];
}

/* get_new_state_vars!(trial_pos, atomindex, ...) (src/EnergyEvaluation.jl:1351-1386)
 * with calc_delta_matrix (DeltaMatrix.jl:219-262): features after moving atom
 * i to trial, through calc_delta_symm_val! (radial :69-117, angular :119-217). */
int orc_nnp2_move_g(const orc_nnp2 *nn, const double *x, int n1, int n2, int i, const double *trial,
                    const double *g, double *g_new)
{
    const int N = n1 + n2, Ng = 2 * nn->n_rad + 3 * nn->n_ang;
    double *r2, *f;
    if (pair_tables(nn, x, N, &r2, &f)) return -1;
    double *rn = (double *)malloc(sizeof(double) * N), *fn = (double *)malloc(sizeof(double) * N);
    double *dg = (double *)calloc((size_t)Ng * N, sizeof(double));
    if (!rn || !fn || !dg) return -1;
    for (int j = 0; j < N; ++j) {
        rn[j] = j == i ? 0.0 : d2(trial, &x[3 * j]);
        fn[j] = orc_cutoff(sqrt(rn[j]), nn->r_cut);
    }
#define DG(row, atom) dg[(size_t)(atom) * Ng + (row)]
    for (int q = 0; q < nn->n_rad; ++q) {
        const int r0 = 2 * q, ind = i < n1 ? 0 : 1;
        const double eta = RAD2(q, 0);
        for (int j = 0; j < n1; ++j) {
            if (j == i) continue;
            const double gn = expo1(eta, rn[j], fn[j]) - expo1(eta, r2[i * N + j], f[i * N + j]);
            DG(r0, i) += gn;
            DG(r0 + ind, j) = gn;
        }
        for (int j = n1; j < N; ++j) {
            if (j == i) continue;
            const double gn = expo1(eta, rn[j], fn[j]) - expo1(eta, r2[i * N + j], f[i * N + j]);
            DG(r0 + 1, i) += gn;
            DG(r0 + ind, j) += gn;
        }
        const double s1[2] = {RAD2(q, 1), RAD2(q, 2)}, s2[2] = {RAD2(q, 3), RAD2(q, 4)};
        scale_rows(dg, Ng, n1, N, r0, 2, s1, s2);
    }
    for (int q = 0; q < nn->n_ang; ++q) {
        const int r0 = 2 * nn->n_rad + 3 * q;
        const double eta = ANG2(q, 0), lam = ANG2(q, 1), zeta = ANG2(q, 2);
        int iv[4] = {0, 1, 1, 0};
        if (i >= n1)
            for (int k = 0; k < 4; ++k) iv[k] += 1;
        /* the three blocks of partner pairs: (11), (22), (12) */
        for (int blk = 0; blk < 3; ++blk) {
            const int jlo = blk == 1 ? n1 : 0, jhi = blk == 1 ? N : n1;
            const int ci = blk == 0 ? 0 : (blk == 1 ? 2 : 1);
            const int rj = blk == 0 ? iv[0] : (blk == 1 ? iv[1] : iv[2]);
            const int rk = blk == 0 ? iv[0] : (blk == 1 ? iv[1] : iv[3]);
            for (int j = jlo; j < jhi; ++j) {
                if (j == i) continue;
                const int klo = blk == 2 ? n1 : j + 1, khi = blk == 0 ? n1 : N;
                for (int k = klo; k < khi; ++k) {
                    if (k == i) continue;
                    /* new_angular_symm_vals (DeltaMatrix.jl:30-61) */
                    double tn[3], to[3], vn[3], vo[3];
                    angles(trial, &x[3 * j], &x[3 * k], rn[j], rn[k], r2[j * N + k], tn);
                    angles(&x[3 * i], &x[3 * j], &x[3 * k], r2[i * N + j], r2[i * N + k], r2[j * N + k], to);
                    for (int c = 0; c < 3; ++c) {
                        vo[c] = pow(1.0 + lam * to[c], zeta);
                        vn[c] = pow(1.0 + lam * tn[c], zeta);
                    }
                    const double en = exp(-eta * (rn[j] + rn[k] + r2[j * N + k])) * fn[j] * fn[k] * f[j * N + k];
                    const double eo = exp(-eta * (r2[i * N + j] + r2[i * N + k] + r2[j * N + k])) * f[i * N + j] *
                                      f[i * N + k] * f[j * N + k];
                    DG(r0 + ci, i) += vn[0] * en - vo[0] * eo;
                    DG(r0 + rj, j) += vn[1] * en - vo[1] * eo;
                    DG(r0 + rk, k) += vn[2] * en - vo[2] * eo;
                }
            }
        }
        const double s1[3] = {ANG2(q, 3), ANG2(q, 4), ANG2(q, 5)}, s2[3] = {ANG2(q, 6), ANG2(q, 7), ANG2(q, 8)};
        scale_rows(dg, Ng, n1, N, r0, 3, s1, s2);
    }
#undef DG
    for (size_t e = 0; e < (size_t)Ng * N; ++e) g_new[e] = g[e] + dg[e];
    free(r2);
    free(f);
    free(rn);
    free(fn);
    free(dg);
    return 0;
}

/* get_new_state_vars!(indices, ...) (src/EnergyEvaluation.jl:1387-1411) with
 * calc_swap_matrix (DeltaMatrix.jl:583-624): features after atoms i1
 * (species 1) and i2 (species 2) trade places, through calc_swap_symm_val
 * (radial :322-371, angular :373-578).  Column i1 then describes a species-1
 * atom at i2's old position and column i2 a species-2 atom at i1's. */
int orc_nnp2_swap_g(const orc_nnp2 *nn, const double *x, int n1, int n2, int i1, int i2, const double *g,
                    double *g_new)
{
    const int N = n1 + n2, Ng = 2 * nn->n_rad + 3 * nn->n_ang;
    if (i1 < 0 || i1 >= n1 || i2 < n1 || i2 >= N) return -2;
    double *r2, *f;
    if (pair_tables(nn, x, N, &r2, &f)) return -1;
    double *dg = (double *)calloc((size_t)Ng * N, sizeof(double));
    if (!dg) return -1;
#define DG(row, atom) dg[(size_t)(atom) * Ng + (row)]
#define R2(a, b) r2[(a) * N + (b)]
#define FC(a, b) f[(a) * N + (b)]
#define P(a) (&x[3 * (a)])
    for (int q = 0; q < nn->n_rad; ++q) {
        const int r0 = 2 * q;
        const double eta = RAD2(q, 0);
        for (int idx = 0; idx < n1; ++idx) {
            if (idx == i1) continue;
            const double v1 = one1(R2(i1, idx), FC(i1, idx), eta), v2 = one1(R2(i2, idx), FC(i2, idx), eta);
            DG(r0, idx) += v2 - v1;
            DG(r0 + 1, idx) += v1 - v2;
            DG(r0, i1) += v2 - v1;
            DG(r0, i2) += v1 - v2;
        }
        for (int idx = n1; idx < N; ++idx) {
            if (idx == i2) continue;
            const double v1 = one1(R2(i1, idx), FC(i1, idx), eta), v2 = one1(R2(i2, idx), FC(i2, idx), eta);
            DG(r0, idx) += v2 - v1;
            DG(r0 + 1, idx) += v1 - v2;
            DG(r0 + 1, i1) += v2 - v1;
            DG(r0 + 1, i2) += v1 - v2;
        }
        const double s1[2] = {RAD2(q, 1), RAD2(q, 2)}, s2[2] = {RAD2(q, 3), RAD2(q, 4)};
        scale_rows(dg, Ng, n1, N, r0, 2, s1, s2);
    }
    for (int q = 0; q < nn->n_ang; ++q) {
        const int r0 = 2 * nn->n_rad + 3 * q;
        const double eta = ANG2(q, 0), lam = ANG2(q, 1), zeta = ANG2(q, 2);
        double a[3], b[3];
        /* both partners species 1 */
        for (int j = 0; j < n1; ++j) {
            if (j == i1) continue;
            for (int k = j + 1; k < n1; ++k) {
                if (k == i1) continue;
                one3(P(i1), P(j), P(k), R2(i1, j), R2(i1, k), R2(j, k), FC(i1, j), FC(i1, k), FC(j, k), eta, lam,
                     zeta, a);
                one3(P(i2), P(j), P(k), R2(i2, j), R2(i2, k), R2(j, k), FC(i2, j), FC(i2, k), FC(j, k), eta, lam,
                     zeta, b);
                DG(r0, i1) += b[0] - a[0];
                DG(r0, i2) += a[0] - b[0];
                DG(r0, j) += b[1] - a[1];
                DG(r0, k) += b[2] - a[2];
                DG(r0 + 1, j) += a[1] - b[1];
                DG(r0 + 1, k) += a[2] - b[2];
            }
        }
        /* both partners species 2 */
        for (int j = n1; j < N; ++j) {
            if (j == i2) continue;
            for (int k = j + 1; k < N; ++k) {
                if (k == i2) continue;
                one3(P(i1), P(j), P(k), R2(i1, j), R2(i1, k), R2(j, k), FC(i1, j), FC(i1, k), FC(j, k), eta, lam,
                     zeta, a);
                one3(P(i2), P(j), P(k), R2(i2, j), R2(i2, k), R2(j, k), FC(i2, j), FC(i2, k), FC(j, k), eta, lam,
                     zeta, b);
                DG(r0 + 2, i1) += b[0] - a[0];
                DG(r0 + 2, i2) += a[0] - b[0];
                DG(r0 + 1, j) += b[1] - a[1];
                DG(r0 + 1, k) += b[2] - a[2];
                DG(r0 + 2, j) += a[1] - b[1];
                DG(r0 + 2, k) += a[2] - b[2];
            }
        }
        /* one partner of each species */
        for (int j = 0; j < n1; ++j) {
            if (j == i1) continue;
            for (int k = n1; k < N; ++k) {
                if (k == i2) continue;
                one3(P(i1), P(j), P(k), R2(i1, j), R2(i1, k), R2(j, k), FC(i1, j), FC(i1, k), FC(j, k), eta, lam,
                     zeta, a);
                one3(P(i2), P(j), P(k), R2(i2, j), R2(i2, k), R2(j, k), FC(i2, j), FC(i2, k), FC(j, k), eta, lam,
                     zeta, b);
                DG(r0 + 1, i1) += b[0] - a[0];
                DG(r0 + 1, i2) += a[0] - b[0];
                DG(r0 + 1, j) += b[1] - a[1];
                DG(r0, k) += b[2] - a[2];
                DG(r0 + 2, j) += a[1] - b[1];
                DG(r0 + 1, k) += a[2] - b[2];
            }
        }
        /* triangles holding both swapped atoms: the third of species 1, then 2 */
        for (int k = 0; k < n1; ++k) {
            if (k == i1) continue;
            one3(P(i1), P(i2), P(k), R2(i1, i2), R2(i1, k), R2(i2, k), FC(i1, i2), FC(i1, k), FC(i2, k), eta, lam,
                 zeta, a);
            DG(r0 + 1, i1) -= a[0];
            DG(r0, i2) += a[0];
            DG(r0, i2) -= a[1];
            DG(r0 + 1, i1) += a[1];
        }
        for (int k = n1; k < N; ++k) {
            if (k == i2) continue;
            one3(P(i1), P(i2), P(k), R2(i1, i2), R2(i1, k), R2(i2, k), FC(i1, i2), FC(i1, k), FC(i2, k), eta, lam,
                 zeta, a);
            DG(r0 + 2, i1) -= a[0];
            DG(r0 + 1, i2) += a[0];
            DG(r0 + 1, i2) -= a[1];
            DG(r0 + 2, i1) += a[1];
        }
        const double s1[3] = {ANG2(q, 3), ANG2(q, 4), ANG2(q, 5)}, s2[3] = {ANG2(q, 6), ANG2(q, 7), ANG2(q, 8)};
        scale_rows(dg, Ng, n1, N, r0, 3, s1, s2);
    }
#undef DG
#undef R2
#undef FC
#undef P
    for (size_t e = 0; e < (size_t)Ng * N; ++e) g_new[e] = g[e] + dg[e];
    free(r2);
    free(f);
    free(dg);
    return 0;
}
