// Batched FP64 kernels of the potential plugin functions (ptmc_plugin.hpp).
//
// Work layout: the squared-distance matrix of every configuration is built
// once per call (one thread per matrix element), then the potential's
// kernels run one thread per (configuration, atom) or per (configuration,
// feature, atom), and every total is summed by one thread per configuration
// in the reference's order.  The order of every sum follows the reference:
//   dimer_energy_config       pairs i < j, row i then j (src/EnergyEvaluation.jl:378-396)
//   dimer_energy_update!      sum over j != i of the new row minus the old (:313-336, :804-821)
//   calc_components           row sums with the diagonal (:1105-1148)
//   calc_energies_from_components  atoms in order (:1156-1162)
//   total_symm_calc           per atom: partners / partner triples in the
//                             reference's pair and triple order
//                             (MachineLearningPotential/SymmFunc.jl:259-312, :494-510)
//   total_thr_symm!           per atom: the moved atom's pairs / pair couples in order
//                             (MachineLearningPotential/DeltaMatrix.jl:794-872, :933-975)
// so pair potentials reproduce the reference's arithmetic operation by
// operation; the EAM and network results differ from it only by the device's
// pow / exp / cos / tanh against the host library's.
#include "ptmc_plugin.hpp"
#include "ptmc_sweep.hpp"

namespace ptmc {
namespace {

constexpr int PT = 256;  // threads per workgroup

struct WS {
    double *d2;    // squared distances
    double *m1;    // network: cut-off function; ELJB / LUT: xy extent^2
    double *m2;    // ELJB / LUT: z extent
    double *pe;    // pair energies
    double *rows;  // update: per (configuration, partner) values, 4 rows of N
    double *flag;  // network update: 1 = the hard-core test failed
};

__host__ __device__ inline int plugin_mats(int kind)
{
    if (kind == POT_EAM) return 1;
    if (kind == POT_ELJB || kind == POT_LUT) return 4;
    return 2;  // ELJ: d2, pair energies; network: d2, cut-off function
}

__host__ __device__ inline WS ws_of(const PluginJob &j)
{
    const size_t CM = (size_t)j.n_conf * j.N * j.N;
    WS w{};
    w.d2 = j.ws;
    if (j.kind == POT_NNP) {
        w.m1 = j.ws + CM;
    } else if (j.kind == POT_ELJB || j.kind == POT_LUT) {
        w.m1 = j.ws + CM;
        w.m2 = j.ws + 2 * CM;
        w.pe = j.ws + 3 * CM;
    } else if (j.kind != POT_EAM) {
        w.pe = j.ws + CM;
    }
    w.rows = j.ws + (size_t)plugin_mats(j.kind) * CM;
    w.flag = w.rows + (size_t)4 * j.n_conf * j.N;
    return w;
}

__device__ __forceinline__ BoxC<double> box_of(const PluginJob &j, int c, double &l, double &h)
{
    l = j.box ? j.box[2 * c] : 1.0;
    h = j.bc == BC_CUBIC ? l : (j.box ? j.box[2 * c + 1] : 1.0);
    return make_box<double>(l, h);
}

template <bool DIV>
__device__ double geom_t(int bc, const double *a, const double *b, const BoxC<double> &B, double &dxy2, double &dz)
{
    if (bc == BC_CUBIC || bc == BC_RECT)
        return box_d2c<double, BC_CUBIC, DIV>(a[0], a[1], a[2], b[0], b[1], b[2], B, dxy2, dz);
    if (bc == BC_RHOMB) return box_d2c<double, BC_RHOMB, DIV>(a[0], a[1], a[2], b[0], b[1], b[2], B, dxy2, dz);
    return box_d2c<double, BC_SPHERE, DIV>(a[0], a[1], a[2], b[0], b[1], b[2], B, dxy2, dz);
}

// distance2(a, b, bc) (src/Configurations.jl:95-126): b moved to its image
// nearest a; the xy / z extents of a - b for the angle to z (get_tan, :177-209)
__device__ __forceinline__ double geom(const PluginJob &j, const double *a, const double *b, const BoxC<double> &B,
                                       double &dxy2, double &dz)
{
    return j.div ? geom_t<true>(j.bc, a, b, B, dxy2, dz) : geom_t<false>(j.bc, a, b, B, dxy2, dz);
}

struct Coef {
    double c[8], ca[8], cb[8];
};

__device__ __forceinline__ Coef coef_of(const PluginJob &j)
{
    Coef C;
    for (int k = 0; k < 8; ++k) {
        const bool in = k < j.pot.ncoeff;
        C.c[k] = in ? j.pot.coeff[k] : 0.0;
        C.ca[k] = in && j.kind == POT_ELJB ? j.pot.coeff_a[k] : 0.0;
        C.cb[k] = in && j.kind == POT_ELJB ? j.pot.coeff_b[k] : 0.0;
    }
    return C;
}

// dimer_energy(pot, r2[, angle]) (src/EnergyEvaluation.jl:273-295, :923-958)
__device__ double pair_e(const PluginJob &j, const Coef &C, double r2, double dxy2, double dz)
{
    switch (j.pk) {
    case PK_EVEN_R: return pair_energy<double, PK_EVEN_R, 8>(r2, C.c);
    case PK_ODD_R: return pair_energy<double, PK_ODD_R, 8>(r2, C.c);
    case PK_ODD_H: return pair_energy<double, PK_ODD_H, 8>(r2, C.c);
    case PK_ELJB: return pair_energy_b<double, 8>(r2, dxy2, dz, C.c, C.ca, C.cb);
    case PK_LUT: return pair_energy_lut<double>(r2, dxy2, dz, j.pot);
    default: return pair_energy<double, PK_EVEN_H, 8>(r2, C.c);
    }
}

__device__ __forceinline__ double rcut_of(const PluginJob &j, double l, double h)
{
    return box_rcut(j.bc == BC_RECT ? BC_RECT : (j.bc == BC_RHOMB ? BC_RHOMB : BC_CUBIC), l, h);
}

// lrc of the potential scaled to the box (src/EnergyEvaluation.jl:420-474, :968-1008)
__device__ double tail(const PluginJob &j, const Coef &C, double rc, double l, double h)
{
    const int kind = j.bc == BC_RECT ? BC_RECT : (j.bc == BC_RHOMB ? BC_RHOMB : BC_CUBIC);
    if (j.pk == PK_ELJB) return box_lrc<PK_ELJB, 8>(kind, j.N, rc, l, h, C.c);
    if (j.pk == PK_LUT) {
        const double c6[1] = {j.pot.lut_c6};
        return box_lrc<PK_LUT, 1>(kind, j.N, rc, l, h, c6);
    }
    return box_lrc<PK_EVEN_R, 8>(kind, j.N, rc, l, h, C.c);
}

// invrexp (src/EnergyEvaluation.jl:1088-1095)
__device__ __forceinline__ void invrexp(double r2, double n, double m, double &f, double &g)
{
    if (r2 != 0.0) {
        const double rt = 1.0 / sqrt(r2);
        f = pow(rt, n);
        g = pow(rt, m);
    } else {
        f = 0.0;
        g = 0.0;
    }
}

// cutoff_function (MachineLearningPotential/Cutoff.jl:79-90)
__device__ __forceinline__ double fcut(double r, double rc)
{
    return r < rc ? 0.5 * (cos(3.14159265358979323846 * (r / rc)) + 1.0) : 0.0;
}

// thetacalc: xy . xz / (|xy| |xz|) (MachineLearningPotential/Cutoff.jl:31-68)
__device__ __forceinline__ double theta(double x0, double x1, double x2, double z0, double z1, double z2, double dxy,
                                        double dxz)
{
    return (x0 * z0 + x1 * z1 + x2 * z2) / (dxy * dxz);
}

// Cosine of the triangle (p, q, s) at vertex v (0: p, 1: q, 2: s), with the
// squared sides r_pq, r_ps, r_qs (all_angular_measure: the vectors from the
// vertex to the other two corners).
__device__ __forceinline__ double cos_at(int v, const double *p, const double *q, const double *s, double rpq,
                                         double rps, double rqs)
{
    if (v == 0)
        return theta(p[0] - q[0], p[1] - q[1], p[2] - q[2], p[0] - s[0], p[1] - s[1], p[2] - s[2], sqrt(rpq),
                     sqrt(rps));
    if (v == 1)
        return theta(q[0] - p[0], q[1] - p[1], q[2] - p[2], q[0] - s[0], q[1] - s[1], q[2] - s[2], sqrt(rpq),
                     sqrt(rqs));
    return theta(s[0] - p[0], s[1] - p[1], s[2] - p[2], s[0] - q[0], s[1] - q[1], s[2] - q[2], sqrt(rps), sqrt(rqs));
}

// ---------------------------------------------------------------- kernels

// squared distances of every configuration (get_distance2_mat: pair i < j
// once, as distance2(x_i, x_j), mirrored) or the caller's matrices; the
// network's cut-off function, or the xy / z extents of angle-dependent pairs
__global__ void __launch_bounds__(PT) plugin_geom_kernel(PluginJob j)
{
    const size_t M = (size_t)j.N * j.N;
    const size_t e = (size_t)blockIdx.x * PT + threadIdx.x;
    if (e >= (size_t)j.n_conf * M) return;
    const WS w = ws_of(j);
    const int c = (int)(e / M), r = (int)(e % M), a = r / j.N, b = r % j.N;
    double r2 = 0.0, dxy2 = 0.0, dz = 0.0;
    if (j.d2_in) {
        r2 = j.d2_in[e];
    } else if (a != b) {
        double l, h;
        const BoxC<double> B = box_of(j, c, l, h);
        const double *X = j.x + (size_t)c * j.N * 3;
        const int lo = a < b ? a : b, hi = a < b ? b : a;
        r2 = geom(j, X + 3 * lo, X + 3 * hi, B, dxy2, dz);
    }
    w.d2[e] = r2;
    if (j.kind == POT_NNP) {
        w.m1[e] = fcut(sqrt(r2), j.r_cut);
    } else if (w.m2) {
        w.m1[e] = dxy2;
        w.m2[e] = dz;
    }
    if (r == 0) w.flag[c] = 0.0;
}

// pair energies e(r_ab^2) of the pairs a < b (mirrored); pairs beyond the
// NPT cut-off contribute nothing (dimer_energy_config skips them)
__global__ void __launch_bounds__(PT) plugin_pair_kernel(PluginJob j)
{
    const size_t M = (size_t)j.N * j.N;
    const size_t e = (size_t)blockIdx.x * PT + threadIdx.x;
    if (e >= (size_t)j.n_conf * M) return;
    const WS w = ws_of(j);
    const int c = (int)(e / M), r = (int)(e % M), a = r / j.N, b = r % j.N;
    if (a > b) return;
    double v = 0.0;
    if (a < b) {
        double l, h;
        box_of(j, c, l, h);
        const double r2 = w.d2[e];
        if (!j.cutoff || r2 <= rcut_of(j, l, h)) {
            const Coef C = coef_of(j);
            v = pair_e(j, C, r2, w.m1 ? w.m1[e] : 0.0, w.m2 ? w.m2[e] : 0.0);
        }
        w.pe[(size_t)c * M + (size_t)b * j.N + a] = v;
    }
    w.pe[e] = v;
}

// dimer_energy_config: the per-atom rows (partners in order) and the total
// over pairs i < j in order, plus the tail correction under NPT
__global__ void __launch_bounds__(PT) plugin_pair_sum_kernel(PluginJob j)
{
    const int c = blockIdx.x, N = j.N;
    const WS w = ws_of(j);
    const double *pe = w.pe + (size_t)c * N * N;
    for (int a = threadIdx.x; a < N; a += PT) {
        double s = 0.0;
        for (int b = 0; b < N; ++b)
            if (b != a) s += pe[(size_t)a * N + b];
        j.en_atom[(size_t)c * N + a] = s;
    }
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int a = 0; a < N; ++a)
            for (int b = a + 1; b < N; ++b) tot += pe[(size_t)a * N + b];
        if (j.cutoff) {
            double l, h;
            box_of(j, c, l, h);
            const Coef C = coef_of(j);
            tot += tail(j, C, rcut_of(j, l, h), l, h);
        }
        j.en[c] = tot;
    }
}

// energy_update! of a pair potential, per partner k of the moved atom i:
// the new pair energy (trial position against x_k) and the old one
// (dist2_mat[i, k]); rows[0] new, rows[1] old
__global__ void __launch_bounds__(PT) plugin_pair_row_kernel(PluginJob j)
{
    const int N = j.N;
    const size_t t = (size_t)blockIdx.x * PT + threadIdx.x;
    if (t >= (size_t)j.n_conf * N) return;
    const WS w = ws_of(j);
    const int c = (int)(t / N), k = (int)(t % N), i = j.index[c];
    double en = 0.0, eo = 0.0;
    if (k != i) {
        double l, h;
        const BoxC<double> B = box_of(j, c, l, h);
        const double *X = j.x + (size_t)c * N * 3;
        const double rc = j.cutoff ? rcut_of(j, l, h) : 0.0;
        const Coef C = coef_of(j);
        double dxyn, dzn, dxyo = 0.0, dzo = 0.0, ro;
        const double rn = geom(j, j.trial + 3 * (size_t)c, X + 3 * k, B, dxyn, dzn);
        if (j.d2_in) {
            ro = j.d2_in[(size_t)c * N * N + (size_t)i * N + k];
        } else {
            const int lo = i < k ? i : k, hi = i < k ? k : i;
            ro = geom(j, X + 3 * lo, X + 3 * hi, B, dxyo, dzo);
        }
        if (!j.cutoff || rn <= rc) en = pair_e(j, C, rn, dxyn, dzn);
        if (!j.cutoff || ro <= rc) eo = pair_e(j, C, ro, dxyo, dzo);
    }
    double *R = w.rows + (size_t)c * 4 * N;
    R[k] = en;
    R[N + k] = eo;
}

// dimer_energy_update!: (sum of new - sum of old) + en_tot
__global__ void __launch_bounds__(PT) plugin_pair_delta_kernel(PluginJob j)
{
    const int c = blockIdx.x * PT + threadIdx.x;
    if (c >= j.n_conf) return;
    const int N = j.N, i = j.index[c];
    const double *R = ws_of(j).rows + (size_t)c * 4 * N;
    double sn = 0.0, so = 0.0;
    for (int k = 0; k < N; ++k)
        if (k != i) sn += R[k];
    for (int k = 0; k < N; ++k)
        if (k != i) so += R[N + k];
    j.en[c] = (sn - so) + j.en_in[c];
}

// set_variables(EmbeddedAtomPotential): phi / rho of each atom over its row
// of the distance matrix (diagonal included); the per-atom energy terms
__global__ void __launch_bounds__(PT) plugin_eam_comp_kernel(PluginJob j)
{
    const int N = j.N;
    const size_t t = (size_t)blockIdx.x * PT + threadIdx.x;
    if (t >= (size_t)j.n_conf * N) return;
    const int c = (int)(t / N), a = (int)(t % N);
    const double *d2 = ws_of(j).d2 + (size_t)c * N * N + (size_t)a * N;
    double f = 0.0, g = 0.0;
    for (int b = 0; b < N; ++b) {
        double pf, pg;
        invrexp(d2[b], j.pot.eam_n, j.pot.eam_m, pf, pg);
        f += pf;
        g += pg;
    }
    j.vars_out[2 * t] = f;
    j.vars_out[2 * t + 1] = g;
    j.en_atom[t] = j.pot.eam_ean * f - j.pot.eam_ecam * sqrt(g);
}

// calc_energies_from_components
__global__ void __launch_bounds__(PT) plugin_eam_energy_kernel(PluginJob j)
{
    const int c = blockIdx.x * PT + threadIdx.x;
    if (c >= j.n_conf) return;
    const double *V = j.vars_out + (size_t)c * j.N * 2;
    double e = 0.0;
    for (int a = 0; a < j.N; ++a) e += j.pot.eam_ean * V[2 * a] - j.pot.eam_ecam * sqrt(V[2 * a + 1]);
    j.en[c] = e;
}

// energy_update!(EmbeddedAtomPotential), per partner k: the term
// invrexp(new) - invrexp(old) (rows[0] phi, rows[1] rho) and the partner's
// new components comp[k] + term
__global__ void __launch_bounds__(PT) plugin_eam_row_kernel(PluginJob j)
{
    const int N = j.N;
    const size_t t = (size_t)blockIdx.x * PT + threadIdx.x;
    if (t >= (size_t)j.n_conf * N) return;
    const int c = (int)(t / N), k = (int)(t % N), i = j.index[c];
    double l, h;
    const BoxC<double> B = box_of(j, c, l, h);
    const double *X = j.x + (size_t)c * N * 3;
    double dxy, dz;
    const double rn = k == i ? 0.0 : geom(j, j.trial + 3 * (size_t)c, X + 3 * k, B, dxy, dz);
    double ro;
    if (j.d2_in) {
        ro = j.d2_in[(size_t)c * N * N + (size_t)i * N + k];
    } else if (k == i) {
        ro = 0.0;
    } else {
        const int lo = i < k ? i : k, hi = i < k ? k : i;
        ro = geom(j, X + 3 * lo, X + 3 * hi, B, dxy, dz);
    }
    double fn, gn, fo, go;
    invrexp(rn, j.pot.eam_n, j.pot.eam_m, fn, gn);
    invrexp(ro, j.pot.eam_n, j.pot.eam_m, fo, go);
    const double tf = fn - fo, tg = gn - go;
    double *R = ws_of(j).rows + (size_t)c * 4 * N;
    R[k] = tf;
    R[N + k] = tg;
    if (k != i) {
        const double nf = j.vars_in[2 * t] + tf, ng = j.vars_in[2 * t + 1] + tg;
        j.vars_out[2 * t] = nf;
        j.vars_out[2 * t + 1] = ng;
        j.en_atom[t] = j.pot.eam_ean * nf - j.pot.eam_ecam * sqrt(ng);
    }
}

// the moved atom's components: every partner's term added in order (the
// term of k = i twice, as calc_components adds it to row k and to row i),
// then the energy over all atoms in order
__global__ void __launch_bounds__(PT) plugin_eam_update_kernel(PluginJob j)
{
    const int c = blockIdx.x * PT + threadIdx.x;
    if (c >= j.n_conf) return;
    const int N = j.N, i = j.index[c];
    const double *R = ws_of(j).rows + (size_t)c * 4 * N;
    const size_t ti = (size_t)c * N + i;
    double vf = j.vars_in[2 * ti], vg = j.vars_in[2 * ti + 1];
    for (int k = 0; k < N; ++k) {
        if (k == i) {
            vf += R[k];
            vg += R[N + k];
        }
        vf += R[k];
        vg += R[N + k];
    }
    j.vars_out[2 * ti] = vf;
    j.vars_out[2 * ti + 1] = vg;
    j.en_atom[ti] = j.pot.eam_ean * vf - j.pot.eam_ecam * sqrt(vg);
    const double *V = j.vars_out + (size_t)c * N * 2;
    double e = 0.0;
    for (int a = 0; a < N; ++a) e += j.pot.eam_ean * V[2 * a] - j.pot.eam_ecam * sqrt(V[2 * a + 1]);
    j.en[c] = e;
}

// total_symm_calc: one thread per (configuration, function, atom).  Radial
// rows sum fc exp(-eta r^2) g_norm over the partners in order (pairs at
// fc = 0 or 1 add nothing); angular rows sum over the triples p < q < s that
// hold the atom, in the reference's triple order, exp(-eta (r_pq + r_ps +
// r_qs)) f_pq f_ps f_qs (1 + lambda cos)^zeta tpz with the cosine at the
// atom's corner; triples with a zero cut-off factor add exactly zero and are
// skipped.  Inactive functions (and, with q10_compat, the angular functions
// past n_ang - n_rad) leave zero rows.
__global__ void __launch_bounds__(PT) plugin_nnp_feature_kernel(PluginJob j)
{
    const int N = j.N, nf = plugin_nf(j);
    const size_t t = (size_t)blockIdx.x * PT + threadIdx.x;
    if (t >= (size_t)j.n_conf * nf * N) return;
    const int c = (int)(t / ((size_t)nf * N)), rem = (int)(t % ((size_t)nf * N)), q = rem / N, a = rem % N;
    const WS w = ws_of(j);
    const double *r2 = w.d2 + (size_t)c * N * N, *f = w.m1 + (size_t)c * N * N;
    const double *X = j.x + (size_t)c * N * 3;
    double g = 0.0;
    if (q < j.n_rad) {
        const double *R = j.rad + 4 * q;
        if (R[3] != 0.0) {
            const double eta = R[0], norm = R[1];
            double row = 0.0;
            for (int b = 0; b < N; ++b) {
                if (b == a) continue;
                const double fc = f[(size_t)a * N + b];
                row += (fc != 0.0 && fc != 1.0) ? fc * exp(-eta * r2[(size_t)a * N + b]) * norm : 0.0;
            }
            g = row + R[2];
        }
    } else {
        const int qa = q - j.n_rad;
        const double *A = j.ang + 6 * qa;
        if (A[5] != 0.0 && qa < j.n_ang_built) {
            const double eta = A[0], lam = A[1], zeta = A[2], tpz = A[3];
            double row = 0.0;
            auto term = [&](int p, int qq, int s, int v) {
                const double fpq = f[(size_t)p * N + qq], fps = f[(size_t)p * N + s], fqs = f[(size_t)qq * N + s];
                if (fpq == 0.0 || fps == 0.0 || fqs == 0.0) return;
                const double rpq = r2[(size_t)p * N + qq], rps = r2[(size_t)p * N + s], rqs = r2[(size_t)qq * N + s];
                const double e = exp(-eta * (rpq + rps + rqs)) * fpq * fps * fqs;
                const double th = cos_at(v, X + 3 * p, X + 3 * qq, X + 3 * s, rpq, rps, rqs);
                row += e * pow(1.0 + lam * th, zeta) * tpz;
            };
            for (int p = 0; p < a; ++p) {
                if (f[(size_t)p * N + a] == 0.0) continue;
                for (int qq = p + 1; qq < a; ++qq) term(p, qq, a, 2);
                for (int s = a + 1; s < N; ++s) term(p, a, s, 1);
            }
            for (int qq = a + 1; qq < N; ++qq) {
                if (f[(size_t)a * N + qq] == 0.0) continue;
                for (int s = qq + 1; s < N; ++s) term(a, qq, s, 0);
            }
            g = row + A[4];
        }
    }
    j.vars_out[((size_t)c * N + a) * nf + q] = g;
}

// forward pass of one atom: [nf] -> h1 (tanh) -> h2 (tanh) -> 1 (linear),
// RuNNer parameter order (each layer's input-major weights, then its
// biases); skipped where the hard-core test failed
__global__ void __launch_bounds__(PT) plugin_nnp_network_kernel(PluginJob j)
{
    const int N = j.N, nf = plugin_nf(j);
    const size_t t = (size_t)blockIdx.x * PT + threadIdx.x;
    if (t >= (size_t)j.n_conf * N) return;
    const int c = (int)(t / N);
    if (ws_of(j).flag[c] != 0.0) return;
    const double *g = j.vars_out + t * nf;
    const double *W1 = j.w, *B1 = W1 + (size_t)nf * j.h1, *W2 = B1 + j.h1, *B2 = W2 + (size_t)j.h1 * j.h2;
    const double *W3 = B2 + j.h2, *B3 = W3 + j.h2;
    double a1[PLUGIN_MAXH], a2[PLUGIN_MAXH];
    for (int o = 0; o < j.h1; ++o) {
        double s = 0.0;
        for (int k = 0; k < nf; ++k) s += g[k] * W1[(size_t)k * j.h1 + o];
        s += B1[o];
        a1[o] = tanh(s);
    }
    for (int o = 0; o < j.h2; ++o) {
        double s = 0.0;
        for (int k = 0; k < j.h1; ++k) s += a1[k] * W2[(size_t)k * j.h2 + o];
        s += B2[o];
        a2[o] = tanh(s);
    }
    double s = 0.0;
    for (int k = 0; k < j.h2; ++k) s += a2[k] * W3[k];
    s += B3[0];
    j.en_atom[t] = s;
}

// sum of the per-atom energies in order; 100 where the hard-core test
// failed (src/EnergyEvaluation.jl:1601-1606)
__global__ void __launch_bounds__(PT) plugin_nnp_sum_kernel(PluginJob j)
{
    const int c = blockIdx.x * PT + threadIdx.x;
    if (c >= j.n_conf) return;
    if (ws_of(j).flag[c] != 0.0) {
        j.en[c] = 100.0;
        return;
    }
    double e = 0.0;
    for (int a = 0; a < j.N; ++a) e += j.en_atom[(size_t)c * j.N + a];
    j.en[c] = e;
}

// energy_update!(RuNNerPotential), per partner k of the moved atom i: new
// and old squared distances and cut-off functions (rows 0-3); a partner
// inside the hard-core distance marks the configuration
__global__ void __launch_bounds__(PT) plugin_nnp_row_kernel(PluginJob j)
{
    const int N = j.N;
    const size_t t = (size_t)blockIdx.x * PT + threadIdx.x;
    if (t >= (size_t)j.n_conf * N) return;
    const WS w = ws_of(j);
    const int c = (int)(t / N), k = (int)(t % N), i = j.index[c];
    double l, h, dxy, dz;
    const BoxC<double> B = box_of(j, c, l, h);
    const double rn = k == i ? 0.0 : geom(j, j.trial + 3 * (size_t)c, j.x + ((size_t)c * N + k) * 3, B, dxy, dz);
    const size_t ik = (size_t)c * N * N + (size_t)i * N + k;
    double *R = w.rows + (size_t)c * 4 * N;
    R[k] = rn;
    R[N + k] = w.d2[ik];
    R[2 * N + k] = fcut(sqrt(rn), j.r_cut);
    R[3 * N + k] = w.m1[ik];
    if (k != i && rn < j.hardcore2) w.flag[c] = 1.0;
}

// total_thr_symm!: one thread per (configuration, function, atom) adds the
// atom's share of the change, in the reference's order: radial, the pair
// with the moved atom (new term, then the old one subtracted as a negative
// term); angular, the partner couples (k, l), k < l, both != i, that hold
// the atom (all of them for the moved atom), each adding the new triple's
// value and subtracting the old one.  Where the hard-core test failed the
// features are copied.
__global__ void __launch_bounds__(PT) plugin_nnp_delta_kernel(PluginJob j)
{
    const int N = j.N, nf = plugin_nf(j);
    const size_t t = (size_t)blockIdx.x * PT + threadIdx.x;
    if (t >= (size_t)j.n_conf * nf * N) return;
    const int c = (int)(t / ((size_t)nf * N)), rem = (int)(t % ((size_t)nf * N)), q = rem / N, a = rem % N;
    const WS w = ws_of(j);
    const size_t out = ((size_t)c * N + a) * nf + q;
    double v = j.vars_in[out];
    if (w.flag[c] != 0.0) {
        j.vars_out[out] = v;
        return;
    }
    const int i = j.index[c];
    const double *R = w.rows + (size_t)c * 4 * N;
    const double *rn = R, *ro = R + N, *fn = R + 2 * N, *fo = R + 3 * N;
    if (q < j.n_rad) {
        const double *Rd = j.rad + 4 * q;
        if (Rd[3] != 0.0) {
            const double eta = Rd[0], norm = Rd[1];
            if (a != i) {
                v += exp(-eta * rn[a]) * (fn[a] * norm);
                v += exp(-eta * ro[a]) * (-fo[a] * norm);
            } else {
                for (int k = 0; k < N; ++k) {
                    if (k == i) continue;
                    v += exp(-eta * rn[k]) * (fn[k] * norm);
                    v += exp(-eta * ro[k]) * (-fo[k] * norm);
                }
            }
        }
    } else {
        const double *A = j.ang + 6 * (q - j.n_rad);
        if (A[5] != 0.0) {
            const double eta = A[0], lam = A[1], zeta = A[2], tpz = A[3];
            const double *r2 = w.d2 + (size_t)c * N * N, *f = w.m1 + (size_t)c * N * N;
            const double *X = j.x + (size_t)c * N * 3, *T = j.trial + 3 * (size_t)c, *Xi = X + 3 * i;
            auto couple = [&](int k, int l, int vtx) {
                const double fkl = f[(size_t)k * N + l];
                const double pn = fn[k] * fn[l] * fkl, po = fo[k] * fo[l] * fkl;
                if (pn == 0.0 && po == 0.0) return;
                const double rkl = r2[(size_t)k * N + l];
                const double en = exp(-eta * (rn[k] + rn[l] + rkl)) * fn[k] * fn[l] * fkl;
                const double eo = exp(-eta * (ro[k] + ro[l] + rkl)) * fo[k] * fo[l] * fkl;
                const double tn = cos_at(vtx, T, X + 3 * k, X + 3 * l, rn[k], rn[l], rkl);
                const double to = cos_at(vtx, Xi, X + 3 * k, X + 3 * l, ro[k], ro[l], rkl);
                v += en * pow(1.0 + lam * tn, zeta) * tpz;
                v -= eo * pow(1.0 + lam * to, zeta) * tpz;
            };
            if (a != i) {
                for (int k = 0; k < a; ++k)
                    if (k != i) couple(k, a, 2);
                for (int l = a + 1; l < N; ++l)
                    if (l != i) couple(a, l, 1);
            } else {
                for (int k = 0; k < N; ++k) {
                    if (k == i) continue;
                    for (int l = k + 1; l < N; ++l)
                        if (l != i) couple(k, l, 0);
                }
            }
        }
    }
    j.vars_out[out] = v;
}

inline unsigned blocks(size_t n) { return (unsigned)((n + PT - 1) / PT); }

}  // namespace

size_t plugin_workspace_doubles(const PluginJob &j)
{
    const size_t C = (size_t)j.n_conf, N = (size_t)j.N;
    return C * N * N * plugin_mats(j.kind) + 4 * C * N + C;
}

hipError_t plugin_initialise(const PluginJob &j, hipStream_t s)
{
    const size_t C = (size_t)j.n_conf, N = (size_t)j.N;
    hipLaunchKernelGGL(plugin_geom_kernel, dim3(blocks(C * N * N)), dim3(PT), 0, s, j);
    if (j.kind == POT_EAM) {
        hipLaunchKernelGGL(plugin_eam_comp_kernel, dim3(blocks(C * N)), dim3(PT), 0, s, j);
        hipLaunchKernelGGL(plugin_eam_energy_kernel, dim3(blocks(C)), dim3(PT), 0, s, j);
    } else if (j.kind == POT_NNP) {
        hipLaunchKernelGGL(plugin_nnp_feature_kernel, dim3(blocks(C * plugin_nf(j) * N)), dim3(PT), 0, s, j);
        hipLaunchKernelGGL(plugin_nnp_network_kernel, dim3(blocks(C * N)), dim3(PT), 0, s, j);
        hipLaunchKernelGGL(plugin_nnp_sum_kernel, dim3(blocks(C)), dim3(PT), 0, s, j);
    } else {
        hipLaunchKernelGGL(plugin_pair_kernel, dim3(blocks(C * N * N)), dim3(PT), 0, s, j);
        hipLaunchKernelGGL(plugin_pair_sum_kernel, dim3((unsigned)C), dim3(PT), 0, s, j);
    }
    return hipGetLastError();
}

hipError_t plugin_update(const PluginJob &j, hipStream_t s)
{
    const size_t C = (size_t)j.n_conf, N = (size_t)j.N;
    if (j.kind == POT_EAM) {
        hipLaunchKernelGGL(plugin_eam_row_kernel, dim3(blocks(C * N)), dim3(PT), 0, s, j);
        hipLaunchKernelGGL(plugin_eam_update_kernel, dim3(blocks(C)), dim3(PT), 0, s, j);
    } else if (j.kind == POT_NNP) {
        hipLaunchKernelGGL(plugin_geom_kernel, dim3(blocks(C * N * N)), dim3(PT), 0, s, j);
        hipLaunchKernelGGL(plugin_nnp_row_kernel, dim3(blocks(C * N)), dim3(PT), 0, s, j);
        hipLaunchKernelGGL(plugin_nnp_delta_kernel, dim3(blocks(C * plugin_nf(j) * N)), dim3(PT), 0, s, j);
        hipLaunchKernelGGL(plugin_nnp_network_kernel, dim3(blocks(C * N)), dim3(PT), 0, s, j);
        hipLaunchKernelGGL(plugin_nnp_sum_kernel, dim3(blocks(C)), dim3(PT), 0, s, j);
    } else {
        hipLaunchKernelGGL(plugin_pair_row_kernel, dim3(blocks(C * N)), dim3(PT), 0, s, j);
        hipLaunchKernelGGL(plugin_pair_delta_kernel, dim3(blocks(C)), dim3(PT), 0, s, j);
    }
    return hipGetLastError();
}

}  // namespace ptmc
