// Pair potentials one workgroup per configuration ("pair slots"), the layout
// of delayed acceptance with the learned surrogate for an ELJ / LJ cluster.
//
// The tile kernels (sweep_kernel, spread_sweep_kernel) pack many
// configurations into a workgroup; the surrogate needs per configuration
// 64 descriptors of every atom and the network in LDS, so here a workgroup
// of 256 lanes owns one configuration of up to 64 atoms for a cycle, like
// the network potential's slots.  A move (mc_step!, src/MCRun.jl:169-182):
// draws and trial position (src/MCMoves.jl:115-135), the partner terms of the
// moved atom one lane per atom, dE = sum_j e(r'_ij) - e(r_ij)
// (src/EnergyEvaluation.jl:804-821) as a wave sum in binary32 added to the
// FP64 total, and Metropolis (src/Exchange.jl:53-57).  With delayed
// acceptance the surrogate's change dE_hat decides a first stage on the
// move's uniform and the exact dE a second stage on min(1, e^{-beta (dE -
// dE_hat)}) with a second uniform (Christen & Fox 2005), as the network
// kernel does.  Exchange, adaptation and sampling follow in nnp_cycle_kernel.
#pragma once
#include "ptmc_surrogate.hpp"
#include "ptmc_sweep.hpp"

namespace ptmc {

// LDS floats: X[3][64] | rn2, ro2 [64] | surrogate (DA): Sd, Sn [64][SS] f32,
// W1, W2 [K][H] bf16, Hb [64][HS] bf16, atom energies current / trial [64],
// b1, b2, w3 [H], geometry [4][64], layer scratch Y [64][YS]
__host__ __device__ inline size_t pair_slot_lds_floats(bool da)
{
    size_t n = (size_t)3 * NNP_NP + 2 * NNP_NP;
    if (da)
        n += (size_t)2 * NNP_NP * SUR_SS + (size_t)SUR_K * SUR_H + (size_t)NNP_NP * SUR_HS / 2 + 2 * NNP_NP +
             3 * SUR_H + 4 * NNP_NP + (size_t)NNP_NP * SUR_YS;
    return n;
}

template <int PK, int KC>
__device__ __forceinline__ void pair_slot_coeffs(const SweepArgs &a, float (&c)[KC], double (&c64)[KC])
{
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        const double v = k < a.pot.ncoeff ? a.pot.coeff[k] : 0.0;
        c64[k] = v;
        c[k] = (float)v;
    }
    if constexpr (PK == PK_LJ) {
        c[0] = (float)a.pot.coeff[0];
        c[1] = (float)a.pot.coeff[3];
    }
}

// FP64 total energy of the slot's positions (initialise_energy,
// src/EnergyEvaluation.jl:1662-1692): pairs i < j dealt to the threads, the
// per-thread sums added in thread order by thread 0.
template <int PK, int KC>
__device__ double pair_slot_energy64(const float *X, int N, const double (&c64)[KC], double *red)
{
    const int tid = threadIdx.x;
    double part = 0.0;
    for (int i = tid; i < N; i += PSL_NT)
        for (int j = i + 1; j < N; ++j) {
            const double dx = (double)X[i] - (double)X[j], dy = (double)X[NNP_NP + i] - (double)X[NNP_NP + j],
                         dz = (double)X[2 * NNP_NP + i] - (double)X[2 * NNP_NP + j];
            part += pair_energy64<PK, KC>(dx * dx + dy * dy + dz * dz, c64);
        }
    red[tid] = part;
    __syncthreads();
    double tot = 0.0;
    if (tid == 0)
        for (int t = 0; t < PSL_NT; ++t) tot += red[t];
    __syncthreads();
    if (tid == 0) red[0] = tot;
    __syncthreads();
    tot = red[0];
    __syncthreads();
    return tot;
}

template <int PK, int KC>
__global__ void __launch_bounds__(PSL_NT) pair_slot_init_kernel(SweepArgs a)
{
    __shared__ float X[3 * NNP_NP];
    __shared__ double red[PSL_NT];
    const int slot = blockIdx.x;
    if (a.init_run >= 0 && a.slot_run[slot] != a.init_run) return;  // whole workgroup
    const int N = a.runs[a.slot_run[slot]].n_atoms;
    const float *gX = reinterpret_cast<const float *>(a.coords) + (size_t)slot * 3 * NNP_NP;
    for (int e = threadIdx.x; e < 3 * NNP_NP; e += PSL_NT) X[e] = gX[e];
    __syncthreads();
    float c[KC];
    double c64[KC];
    pair_slot_coeffs<PK, KC>(a, c, c64);
    const double E = pair_slot_energy64<PK, KC>(X, N, c64, red);
    if (threadIdx.x == 0) a.slot_en[slot] = E;
}

// One cycle (N moves) of one configuration per workgroup.
template <bool DA, int PK, int KC>
__global__ void __launch_bounds__(PSL_NT) pair_slot_kernel(SweepArgs a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ int s_err;
    __shared__ float s_de;
    __shared__ float s_acci[SUR_K];
    __shared__ double s_ehat;
    __shared__ double red[PSL_NT];
    const int slot = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    float *X = reinterpret_cast<float *>(lds_raw);
    float *rn2 = X + 3 * NNP_NP, *ro2 = rn2 + NNP_NP;
    const RunInfo &ri = a.runs[a.slot_run[slot]];
    const int N = ri.n_atoms;
    const uint32_t nsteps = (uint32_t)N;
    const int t = a.slot_temp[slot];
    const int Tg = ri.temp_base + t;
    const uint32_t sweep = (uint32_t)a.sweep0;
    const double beta = a.t_beta[Tg];
    const float delta_r = (float)a.t_displ[2 * Tg];
    const float radius2 = (float)ri.radius2;
    float c[KC];
    double c64[KC];
    pair_slot_coeffs<PK, KC>(a, c, c64);

    const float *gXr = reinterpret_cast<const float *>(a.coords) + (size_t)slot * 3 * NNP_NP;
    for (int e = tid; e < 3 * NNP_NP; e += PSL_NT) X[e] = gXr[e];
    double E = a.slot_en[slot];
    if (tid == 0) s_err = 0;
    // delayed acceptance: surrogate descriptors and atom energies (rebuilt
    // from the positions at every launch)
    float *Sd = nullptr, *Sn = nullptr, *sat = nullptr, *satn = nullptr, *sgeo = nullptr, *Y = nullptr;
    SurNet net{};
    const SurGeom sg = sur_geom(a.sur_rcut);
    double Ehat = 0.0;
    long long n_prop = 0, n_acc1 = 0;
    __syncthreads();
    if constexpr (DA) {
        Sd = ro2 + NNP_NP;
        Sn = Sd + NNP_NP * SUR_SS;
        net.W1 = reinterpret_cast<short *>(Sn + NNP_NP * SUR_SS);
        net.W2 = net.W1 + SUR_K * SUR_H;
        net.Hb = net.W2 + SUR_H * SUR_H;
        sat = reinterpret_cast<float *>(net.Hb + NNP_NP * SUR_HS);
        satn = sat + NNP_NP;
        net.b1 = satn + NNP_NP;
        net.b2 = net.b1 + SUR_H;
        net.w3 = net.b2 + SUR_H;
        sgeo = net.w3 + SUR_H;  // per atom: r_new, r_old, f_c(r_new), f_c(r_old)
        Y = sgeo + 4 * NNP_NP;
        sur_load_net(net, a.sur_w, PSL_NT);
        sur_descriptors(sg, X, N, Sd, SUR_SS, PSL_NT);
        for (int e = tid; e < NNP_NP * SUR_SS; e += PSL_NT) Sn[e] = 0.0f;
        __syncthreads();
        sur_atom_energies(net, Sd, N, Y, sat, &s_ehat);
        Ehat = s_ehat;
    }
    __syncthreads();
    int nacc = 0;

    for (uint32_t m = 0; m < nsteps; ++m) {
        // draws and trial position (every thread the same), src/MCMoves.jl:115-135
        MoveDraw<float> d0 = move_draw<float>(a.seed, ri.run_key, (uint32_t)t, sweep, m, 0u, nsteps);
        const int i = (int)d0.index;
        float xi = X[i], yi = X[NNP_NP + i], zi = X[2 * NNP_NP + i];
        float tx, ty, tz;
        make_trial<float, BC_SPHERE, false>(xi, yi, zi, d0, delta_r, 0.0f, 0.0f, tx, ty, tz);
        if (outside_sphere<float>(tx, ty, tz, radius2)) {
            int count = 0;
            uint32_t attempt = 0;
            while (outside_sphere<float>(tx, ty, tz, radius2)) {
                count += 1;
                if (count == 50) {
                    __syncthreads();
                    if (tid == 0) {
                        double cm[3] = {0.0, 0.0, 0.0};
                        for (int b = 0; b < N; ++b)
                            for (int k = 0; k < 3; ++k) cm[k] += (double)X[k * NNP_NP + b];
                        for (int k = 0; k < 3; ++k) cm[k] /= (double)N;
                        for (int b = 0; b < N; ++b)
                            for (int k = 0; k < 3; ++k) X[k * NNP_NP + b] = (float)((double)X[k * NNP_NP + b] - cm[k]);
                    }
                    __syncthreads();
                } else {
                    attempt += 1;
                    MoveDraw<float> dd = move_draw<float>(a.seed, ri.run_key, (uint32_t)t, sweep, m, attempt, nsteps);
                    xi = X[i];
                    yi = X[NNP_NP + i];
                    zi = X[2 * NNP_NP + i];
                    make_trial<float, BC_SPHERE, false>(xi, yi, zi, dd, delta_r, 0.0f, 0.0f, tx, ty, tz);
                    if (count == 100) {
                        if (tid == 0) s_err = 1;
                        break;
                    }
                }
            }
            xi = X[i];
            yi = X[NNP_NP + i];
            zi = X[2 * NNP_NP + i];
        }
        // partner terms, one lane of wave 0 per atom: squared distances to
        // the new and old position and dE = sum_j e(r'^2) - e(r^2)
        if (tid < 64) {
            const int j = lane;
            float d = 0.0f;
            if (j < N && j != i) {
                const float px = X[j], py = X[NNP_NP + j], pz = X[2 * NNP_NP + j];
                float dx = tx - px, dy = ty - py, dz = tz - pz;
                const float rn = dx * dx + dy * dy + dz * dz;
                dx = xi - px;
                dy = yi - py;
                dz = zi - pz;
                const float ro = dx * dx + dy * dy + dz * dz;
                rn2[j] = rn;
                ro2[j] = ro;
                d = pair_energy<float, PK, KC>(rn, c) - pair_energy<float, PK, KC>(ro, c);
            } else if (j < NNP_NP) {
                rn2[j] = 0.0f;
                ro2[j] = 0.0f;
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) d += __shfl_xor(d, o, 64);
            if (lane == 0) s_de = d;
        }
        if (DA && tid < SUR_K) s_acci[tid] = 0.0f;
        __syncthreads();
        // delayed acceptance, stage 1: Metropolis on the surrogate energy change
        bool stage1 = true;
        double dEh = 0.0, Ehn = 0.0;
        if constexpr (DA) {
            if (tid < N) {
                const float rn = sqrtf(rn2[tid]), ro = sqrtf(ro2[tid]);
                sgeo[tid] = rn;
                sgeo[NNP_NP + tid] = ro;
                sgeo[2 * NNP_NP + tid] = tid == i ? 0.0f : sur_cutoff(rn, sg.r_cut);
                sgeo[3 * NNP_NP + tid] = tid == i ? 0.0f : sur_cutoff(ro, sg.r_cut);
            }
            __syncthreads();
            float accq = 0.0f;
            const int kq = tid & (SUR_K - 1);  // PSL_NT is a multiple of SUR_K: fixed k per thread
            const float mu = sur_mu(sg, kq);
            for (int j = tid / SUR_K; j < N; j += PSL_NT / SUR_K) {
                if (j == i) continue;
                const float fcn = sgeo[2 * NNP_NP + j], fco = sgeo[3 * NNP_NP + j];
                const float pn = sur_phi_mu(sg, sgeo[j], fcn, mu);
                const float pq = sur_phi_mu(sg, sgeo[NNP_NP + j], fco, mu);
                Sn[j * SUR_SS + kq] = __builtin_fmaf(SUR_SIN, pn - pq, Sd[j * SUR_SS + kq]);
                accq += pn;
            }
            atomicAdd(&s_acci[kq], accq);
            __syncthreads();
            if (tid < SUR_K) Sn[i * SUR_SS + tid] = SUR_SIN * s_acci[tid];
            __syncthreads();
            sur_atom_energies(net, Sn, N, Y, satn, &s_ehat);
            Ehn = s_ehat;
            dEh = Ehn - Ehat;
            stage1 = (dEh <= 0.0) || (det_exp(-dEh * beta) >= (double)d0.umet);
            n_prop += 1;
            n_acc1 += stage1 ? 1 : 0;
        }
        // Metropolis, src/MCRun.jl:122-132 and src/Exchange.jl:53-57; with
        // delayed acceptance the second stage corrects the surrogate's error
        const double dd = (double)s_de;
        bool accept;
        if constexpr (!DA) {
            accept = (dd <= 0.0) || (det_exp(-dd * beta) >= (double)d0.umet);
        } else if (!stage1) {
            accept = false;
        } else {
            const u32x4 X2 = draw(a.seed, ri.run_key, RNG_MOVE, 255u, (uint32_t)t, sweep, m);
            const double x2 = dd - dEh;
            accept = (x2 <= 0.0) || (det_exp(-x2 * beta) >= (double)u24(X2.x));
        }
        if (tid == 0 && a.trace_de != nullptr && a.trace_cycle < a.trace_cycles) {
            const size_t ti = ((size_t)Tg * a.trace_cycles + a.trace_cycle) * nsteps + m;
            a.trace_de[ti] = DA && !stage1 ? __builtin_nan("") : dd;
            a.trace_acc[ti] = (int8_t)accept;
            a.trace_idx[ti] = i;
            if (DA && a.trace_aux != nullptr) a.trace_aux[ti] = dEh;
        }
        __syncthreads();  // every thread has read s_de and the trial's partners
        if (accept) {
            // swap_atom_config!, src/swap_config.jl:40-51
            if constexpr (DA) {
                float *t1 = Sd;
                Sd = Sn;
                Sn = t1;
                t1 = sat;
                sat = satn;
                satn = t1;
                Ehat = Ehn;
            }
            if (tid == 0) {
                X[i] = tx;
                X[NNP_NP + i] = ty;
                X[2 * NNP_NP + i] = tz;
            }
            E = E + dd;
            nacc += 1;
        }
        __syncthreads();
    }

    // periodic FP64 recompute bounds the drift of the binary32 sums
    if (a.recompute_every > 0 && ((long long)sweep + 1) % a.recompute_every == 0) {
        const double ef = pair_slot_energy64<PK, KC>(X, N, c64, red);
        const double dr = __builtin_fabs(E - ef);
        if (tid == 0 && dr > a.slot_drift[slot]) a.slot_drift[slot] = dr;
        E = ef;
    }

    float *gX = reinterpret_cast<float *>(a.coords) + (size_t)slot * 3 * NNP_NP;
    for (int e = tid; e < 3 * NNP_NP; e += PSL_NT) gX[e] = X[e];
    if (tid == 0) {
        a.slot_en[slot] = E;
        a.t_count_atom[2 * Tg] += nacc;
        a.t_count_atom[2 * Tg + 1] += nacc;
        if (s_err) a.run_error[a.slot_run[slot]] = 1;
        if (DA) {
            a.da_stats[3 * slot] += n_prop;
            a.da_stats[3 * slot + 1] += n_acc1;
            a.da_stats[3 * slot + 2] += nacc;
        }
    }
}

}  // namespace ptmc
