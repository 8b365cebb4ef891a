// PT sweep kernel for the Sutton-Chen embedded-atom potential
// (src/EnergyEvaluation.jl:1059-1162, energy_update! :1567-1591):
//   E = sum_i ean * phi_i - eCam * sqrt(rho_i),  phi_i = sum_j r_ij^-n,
//   rho_i = sum_j r_ij^-m.
// A move of atom i changes phi_j, rho_j of every atom (one pair term each)
// and phi_i, rho_i (the sum of those terms).  Each lane keeps phi/rho of its
// atoms in two LDS buffers: the partner loop writes tentative values into the
// spare buffer, an accepted move makes it current (the reference swaps
// component_vector and new_component_vector, src/swap_config.jl:207-210).
// FP64: the new total is summed term by term (bit-identical to the oracle's
// engine arithmetic); binary32 sums the term differences (a true dE).
#pragma once
#include "ptmc_sweep.hpp"

namespace ptmc {

template <typename Real>
__device__ __forceinline__ void eam_terms(Real r2, Real n, Real m, Real &phi, Real &rho)
{
    if constexpr (sizeof(Real) == 8) {
        const double hl = -0.5 * det_log(r2);
        phi = det_exp(n * hl);
        rho = det_exp(m * hl);
    } else {
        const float hl = -0.5f * __builtin_amdgcn_logf(r2);  // log2
        phi = __builtin_amdgcn_exp2f(n * hl);
        rho = __builtin_amdgcn_exp2f(m * hl);
    }
}

// component slab: (buffer b, slab s, component c in {phi, rho}, lane)
template <int LANES>
__device__ __forceinline__ int coff(int A, int b, int s, int c, int lane)
{
    return ((b * A + s) * 2 + c) * LANES + lane;
}

// phi/rho of this lane's atoms from positions (FP64 accumulation) and the
// cluster's total energy (src/EnergyEvaluation.jl:1105-1162, :1800-1811).
template <typename Real, int BC, int KL, int LANES>
__device__ double eam_build(const Real *X, Real *Cp, int A, int cur, int gb, int q, int N, double box, double n,
                            double m, double ean, double ecam)
{
    const double invL = BC == BC_CUBIC ? 1.0 / box : 0.0;
    double part = 0.0;
    for (int s = 0; s < A; ++s) {
        const int j = s * KL + q;
        if (j >= N) {
            Cp[coff<LANES>(A, cur, s, 0, gb + q)] = 0;
            Cp[coff<LANES>(A, cur, s, 1, gb + q)] = 0;
            continue;
        }
        const double xj = (double)X[xoff<LANES>(s, 0, gb + q)];
        const double yj = (double)X[xoff<LANES>(s, 1, gb + q)];
        const double zj = (double)X[xoff<LANES>(s, 2, gb + q)];
        double f = 0.0, g = 0.0;
        for (int l = 0; l < N; ++l) {
            if (l == j) continue;
            const int sl = l / KL, ll = gb + (l % KL);
            const double r2 = dist2_64(xj, yj, zj, (double)X[xoff<LANES>(sl, 0, ll)], (double)X[xoff<LANES>(sl, 1, ll)],
                                       (double)X[xoff<LANES>(sl, 2, ll)], BC, box, invL, false);
            double pf, pg;
            eam_terms<double>(r2, n, m, pf, pg);
            f += pf;
            g += pg;
        }
        Cp[coff<LANES>(A, cur, s, 0, gb + q)] = (Real)f;
        Cp[coff<LANES>(A, cur, s, 1, gb + q)] = (Real)g;
        part += ean * (double)(Real)f - ecam * __builtin_sqrt((double)(Real)g);
    }
    return group_sum<KL>(part);
}

template <typename Real, int BC, int KL, int LANES>
__global__ void __launch_bounds__(LANES) eam_kernel(SweepArgs a)
{
    constexpr bool F32 = sizeof(Real) == 4;
    constexpr int NSLOT = LANES / KL;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const TileInfo tile = a.tiles[blockIdx.x];
    const int tid = threadIdx.x;
    const int lg = tid / KL;
    const int q = tid % KL;
    const int gb = lg * KL;
    const bool active = lg < tile.n_slots;
    const int N = tile.n_atoms;
    const int A = tile.n_slab;
    const uint32_t nsteps = (uint32_t)N;

    Real *X = reinterpret_cast<Real *>(lds_raw);
    Real *Cp = X + 3 * A * LANES;  // [2][A][2][LANES]
    double *exE = reinterpret_cast<double *>(lds_raw + sizeof(Real) * 7 * A * LANES);
    int *exT = reinterpret_cast<int *>(exE + NSLOT);
    int *exS = exT + NSLOT;

    const double n64 = a.pot.eam_n, m64 = a.pot.eam_m, ean64 = a.pot.eam_ean, ecam64 = a.pot.eam_ecam;
    const Real nr = (Real)n64, mr = (Real)m64, eanr = (Real)ean64, ecamr = (Real)ecam64;

    Real *gX = reinterpret_cast<Real *>(a.coords) + tile.coord_off;
    Real *gC = reinterpret_cast<Real *>(a.comps) + tile.comp_off;  // canonical [A][2][LANES]
    for (int e = tid; e < 3 * A * LANES; e += LANES) X[e] = gX[e];
    for (int e = tid; e < 2 * A * LANES; e += LANES) Cp[e] = gC[e];  // buffer 0

    const int slot = tile.slot0 + lg;
    int t = 0, R = 1, tb = 0, lb = 0, cur = 0;
    uint32_t run_key = 0;
    double E = 0.0, box = 1.0;
    Real radius2 = 0;
    if (active) {
        const RunInfo &ri = a.runs[a.slot_run[slot]];
        R = ri.n_temps;
        tb = ri.temp_base;
        lb = ri.local_base;
        run_key = ri.run_key;
        radius2 = (Real)ri.radius2;
        t = a.slot_temp[slot];
        E = a.slot_en[slot];
        if constexpr (BC == BC_CUBIC) box = a.slot_box[slot];
        if (q == 0) {
            exT[lg] = t;
            exS[lb + t] = lg;
        }
    }
    __syncthreads();

    for (int cyc = 0; cyc < a.n_cycles; ++cyc) {
        const uint32_t sweep = (uint32_t)(a.sweep0 + cyc);
        const long long pidx = a.cycle0 + cyc + 1;
        if (active) {
            const double beta = a.t_beta[tb + t];
            const Real delta_r = (Real)a.t_displ[2 * (tb + t)];
            const float beta_l2e = (float)beta * 1.4426950408889634f;
            const Real Lr = (Real)box;
            const Real invLr = (Real)(BC == BC_CUBIC ? 1.0 / box : 0.0);
            int nacc = 0;
            u32x4 mA = {0, 0, 0, 0}, mB = {0, 0, 0, 0}, nA = {0, 0, 0, 0}, nB = {0, 0, 0, 0};
            if constexpr (KL > 1) {
                if ((uint32_t)q < nsteps) move_words<Real>(a.seed, run_key, (uint32_t)t, sweep, (uint32_t)q, mA, mB);
                fetch_words<Real>(mA, mB, gb, nA, nB);
            }
            for (uint32_t mv = 0; mv < nsteps; ++mv) {
                MoveDraw<Real> d0;
                if constexpr (KL == 1) {
                    d0 = move_draw<Real>(a.seed, run_key, (uint32_t)t, sweep, mv, 0u, nsteps);
                } else {
                    d0 = decode_move<Real>(nA, nB, nsteps);
                    const uint32_t m1 = mv + 1;
                    if (m1 < nsteps) {
                        const uint32_t k1 = m1 & (KL - 1);
                        if (k1 == 0 && m1 + q < nsteps)
                            move_words<Real>(a.seed, run_key, (uint32_t)t, sweep, m1 + q, mA, mB);
                        fetch_words<Real>(mA, mB, gb + (int)k1, nA, nB);
                    }
                }
                const int i = (int)d0.index;
                const int si = i / KL, li = gb + (i % KL);
                Real xi = X[xoff<LANES>(si, 0, li)];
                Real yi = X[xoff<LANES>(si, 1, li)];
                Real zi = X[xoff<LANES>(si, 2, li)];
                Real tx, ty, tz;
                make_trial<Real, BC, false>(xi, yi, zi, d0, delta_r, Lr, invLr, tx, ty, tz);
                if constexpr (BC == BC_SPHERE) {
                    if (outside_sphere<Real>(tx, ty, tz, radius2)) {
                        if (!sphere_redraw<Real, KL, LANES>(X, tid, q, N, A, si, li, radius2, delta_r, a.seed, run_key,
                                                            (uint32_t)t, sweep, mv, nsteps, xi, yi, zi, tx, ty, tz) &&
                            q == 0)
                            a.run_error[a.slot_run[slot]] = 1;
                    }
                }
                const int nb = cur ^ 1;
                Real acc = 0, s1 = 0, s2 = 0;
                for (int s = 0; s < A; ++s) {
                    const int j = s * KL + q;
                    const bool valid = (j < N) & (j != i);
                    const Real xj = X[xoff<LANES>(s, 0, tid)];
                    const Real yj = X[xoff<LANES>(s, 1, tid)];
                    const Real zj = X[xoff<LANES>(s, 2, tid)];
                    const Real r2n = dist2<Real, BC, false>(tx, ty, tz, xj, yj, zj, Lr, invLr);
                    const Real r2o = dist2<Real, BC, false>(xi, yi, zi, xj, yj, zj, Lr, invLr);
                    Real fn, gn, fo, go;
                    eam_terms<Real>(r2n, nr, mr, fn, gn);
                    eam_terms<Real>(r2o, nr, mr, fo, go);
                    const Real t1 = valid ? fn - fo : Real(0);
                    const Real t2 = valid ? gn - go : Real(0);
                    const Real pj = Cp[coff<LANES>(A, cur, s, 0, tid)];
                    const Real rj = Cp[coff<LANES>(A, cur, s, 1, tid)];
                    const Real npj = pj + t1, nrj = rj + t2;
                    Cp[coff<LANES>(A, nb, s, 0, tid)] = npj;
                    Cp[coff<LANES>(A, nb, s, 1, tid)] = nrj;
                    if constexpr (F32) {
                        // ean dphi - eCam (sqrt(rho') - sqrt(rho)), the square-root
                        // difference as drho / (sqrt(rho') + sqrt(rho)) (no cancellation)
                        const float den = __builtin_sqrtf(nrj) + __builtin_sqrtf(rj);
                        const float dterm = eanr * t1 - ecamr * (t2 * __builtin_amdgcn_rcpf(den));
                        acc += valid ? dterm : Real(0);
                    } else {
                        acc += valid ? (eanr * npj - ecamr * __builtin_sqrt(nrj)) : Real(0);
                    }
                    s1 += t1;
                    s2 += t2;
                }
                acc = group_sum<KL>(acc);
                s1 = group_sum<KL>(s1);
                s2 = group_sum<KL>(s2);
                const Real pi = Cp[coff<LANES>(A, cur, si, 0, li)];
                const Real ri = Cp[coff<LANES>(A, cur, si, 1, li)];
                const Real npi = pi + s1, nri = ri + s2;
                bool accept;
                double dE_used;
                if constexpr (F32) {
                    const float den = __builtin_sqrtf(nri) + __builtin_sqrtf(ri);
                    const float dEf = acc + (eanr * s1 - ecamr * (s2 * __builtin_amdgcn_rcpf(den)));
                    dE_used = (double)dEf;
                    const float pr = __builtin_amdgcn_exp2f(-dEf * beta_l2e);
                    accept = (dEf <= 0.0f) || (pr >= d0.umet);
                    if (accept) E = E + (double)dEf;
                } else {
                    const double new_en = acc + (ean64 * npi - ecam64 * __builtin_sqrt(nri));
                    const double dd = new_en - E;
                    dE_used = dd;
                    accept = (dd <= 0.0) || (det_exp(-dd * beta) >= d0.umet);
                    if (accept) E = new_en;
                }
                if (accept) {
                    if (tid == li) {
                        Cp[coff<LANES>(A, nb, si, 0, li)] = npi;
                        Cp[coff<LANES>(A, nb, si, 1, li)] = nri;
                        X[xoff<LANES>(si, 0, li)] = tx;
                        X[xoff<LANES>(si, 1, li)] = ty;
                        X[xoff<LANES>(si, 2, li)] = tz;
                    }
                    cur = nb;
                    nacc += 1;
                }
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                if (a.trace_de != nullptr && q == 0 && cyc < a.trace_cycles) {
                    const size_t ti = ((size_t)(tb + t) * a.trace_cycles + cyc) * nsteps + mv;
                    a.trace_de[ti] = dE_used;
                    a.trace_acc[ti] = (int8_t)accept;
                    a.trace_idx[ti] = i;
                }
            }
            if (q == 0) {
                a.t_count_atom[2 * (tb + t)] += nacc;
                a.t_count_atom[2 * (tb + t) + 1] += nacc;
                exE[lg] = E;
            }
        }
        __syncthreads();

        // replica exchange, src/Exchange.jl:103-155 and src/MCRun.jl:203-205,
        // attempted every exchange_every cycles
        if (active && q == 0 && R > 1 && (uint32_t)(sweep + 1u) % (uint32_t)a.exchange_every == 0u) {
            int n = -1;
            if (a.exchange_mode == EXC_EVEN_ODD) {
                if ((t & 1) == (int)(sweep & 1u) && t + 1 < R) n = t;
            } else if (t == 0) {
                u32x4 X0 = draw(a.seed, run_key, RNG_EXCH, 0u, 0u, sweep, 0u);
                double u = u53(X0.x, X0.y);
                if (a.exchange_mode == EXC_EVERY || u < 0.1) n = (int)below(X0.z, (uint32_t)(R - 1));
            }
            if (n >= 0) {
                u32x4 X1 = draw(a.seed, run_key, RNG_EXCH, 0u, 0u, sweep, 1u + (uint32_t)n);
                const double u = u53(X1.x, X1.y);
                const int sa = exS[lb + n], sb = exS[lb + n + 1];
                const double de = exE[sa] - exE[sb];
                const double db = a.t_beta[tb + n] - a.t_beta[tb + n + 1];
                double acc = det_exp(db * de);
                if (acc > 1.0) acc = 1.0;
                a.t_count_exc[2 * (tb + n)] += 1;
                a.t_count_exc[2 * (tb + n + 1)] += 1;
                if (acc > u) {
                    a.t_count_exc[2 * (tb + n) + 1] += 1;
                    a.t_count_exc[2 * (tb + n + 1) + 1] += 1;
                    exT[sa] = n + 1;
                    exT[sb] = n;
                    exS[lb + n] = sb;
                    exS[lb + n + 1] = sa;
                }
            }
        }
        __syncthreads();

        if (active) {
            t = exT[lg];
            if (q == 0) {
                const int T = tb + t;
                if (pidx % a.n_adjust == 0) {
                    const double ar = (double)a.t_count_atom[2 * T + 1] / (double)((long long)a.n_adjust * N);
                    if (ar < a.min_acc) a.t_displ[2 * T] *= 0.9;
                    else if (ar > a.max_acc) a.t_displ[2 * T] *= 1.1;
                    a.t_count_atom[2 * T + 1] = 0;
                }
                if (a.sampling && pidx % a.mc_sample == 0) {
                    const RunInfo &ri = a.runs[a.slot_run[slot]];
                    a.t_ham[2 * T] += E;
                    a.t_ham[2 * T + 1] += E * E;
                    const double hi = (E - ri.en_min) / ri.delta_en + 1.0;
                    const int h = hi < 1.0 ? 1 : (hi > (double)a.n_bin ? a.n_bin + 2 : (int)__builtin_floor(hi) + 1);
                    a.t_hist[(size_t)T * a.hist_stride + (h - 1)] += 1u;
                }
            }
            if (a.recompute_every > 0 && ((long long)sweep + 1) % a.recompute_every == 0) {
                const double ef = eam_build<Real, BC, KL, LANES>(X, Cp, A, cur, gb, q, N, box, n64, m64, ean64, ecam64);
                const double dr = __builtin_fabs(E - ef);
                if (q == 0 && dr > a.slot_drift[slot]) a.slot_drift[slot] = dr;
                E = ef;
            }
        }
    }

    __syncthreads();
    for (int e = tid; e < 3 * A * LANES; e += LANES) gX[e] = X[e];
    // canonical components: each lane copies its own current buffer
    for (int s = 0; s < A; ++s) {
        gC[(s * 2 + 0) * LANES + tid] = Cp[coff<LANES>(A, cur, s, 0, tid)];
        gC[(s * 2 + 1) * LANES + tid] = Cp[coff<LANES>(A, cur, s, 1, tid)];
    }
    if (active && q == 0) {
        a.slot_temp[slot] = t;
        a.slot_en[slot] = E;
        if constexpr (BC == BC_CUBIC) a.slot_box[slot] = box;
    }
}


// set_variables + initialise_energy for EAM (src/EnergyEvaluation.jl:1800-1811,
// :1662-1692): phi/rho of every atom and the FP64 total of each slot.
template <typename Real, int BC, int KL, int LANES>
__global__ void __launch_bounds__(LANES) eam_init_kernel(SweepArgs a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const TileInfo tile = a.tiles[blockIdx.x];
    const int tid = threadIdx.x;
    const int lg = tid / KL;
    const int q = tid % KL;
    const int N = tile.n_atoms;
    const int A = tile.n_slab;
    Real *X = reinterpret_cast<Real *>(lds_raw);
    Real *Cp = X + 3 * A * LANES;
    const Real *gX = reinterpret_cast<const Real *>(a.coords) + tile.coord_off;
    Real *gC = reinterpret_cast<Real *>(a.comps) + tile.comp_off;
    for (int e = tid; e < 3 * A * LANES; e += LANES) X[e] = gX[e];
    __syncthreads();
    if (lg >= tile.n_slots) return;  // whole clusters leave together
    const int slot = tile.slot0 + lg;
    if (a.init_run >= 0 && a.slot_run[slot] != a.init_run) return;
    const double box = BC == BC_CUBIC ? a.slot_box[slot] : 1.0;
    const double E = eam_build<Real, BC, KL, LANES>(X, Cp, A, 0, lg * KL, q, N, box, a.pot.eam_n, a.pot.eam_m,
                                                    a.pot.eam_ean, a.pot.eam_ecam);
    for (int s = 0; s < A; ++s) {
        gC[(s * 2 + 0) * LANES + tid] = Cp[coff<LANES>(A, 0, s, 0, tid)];
        gC[(s * 2 + 1) * LANES + tid] = Cp[coff<LANES>(A, 0, s, 1, tid)];
    }
    if (q == 0) a.slot_en[slot] = E;
}

}  // namespace ptmc
