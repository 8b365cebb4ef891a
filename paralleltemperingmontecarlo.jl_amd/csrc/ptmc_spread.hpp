// Spread layout of the pair-potential sweep (spherical container in NVT; cubic
// and rectangular boxes in NVT and NPT with uniform volume moves).
//
// The ladder-per-workgroup kernel (sweep_kernel) keeps every temperature of a
// run in one workgroup so the exchange can run in LDS, many cycles per
// launch.  A 64-temperature ladder of a 98-atom cluster then fills one CU
// (16 waves, 4 per SIMD) while most of the chip idles, and the cycle takes as
// long as that CU needs.  Here a ladder is cut into tiles of at most
// LANES / KL configurations (LANES = 64, 128 or 256 lanes, KL lanes per
// configuration chosen per run and read from the tile at run time), so its
// configurations spread over as many CUs as there are tiles, and the tiles of
// all runs go out in one launch, heaviest first.  One launch is one cycle of
// moves (src/MCRun.jl:169-182); the exchange, step-size adaptation and
// sampling of the cycle (src/MCRun.jl:203-212, :248-250) follow in
// spread_cycle_kernel, one workgroup per ladder.  The arithmetic of a move,
// the random stream and the order of the per-cluster sums are those of
// sweep_kernel, so both layouts give the same results bit for bit.
#pragma once
#include "ptmc_sweep.hpp"

namespace ptmc {

// Value of the lane whose index differs in bit 4 (ROW = 16) or bit 5 (ROW =
// 32), from gfx950's row swaps: v_permlane16_swap(a, b) exchanges the odd
// 16-lane rows of a with the even rows of b, v_permlane32_swap(a, b) the upper
// half of a with the lower half of b.  With a = b = v, the returned a holds
// the even (lower) rows twice and b the odd (upper) rows twice, so a lane of
// an even row finds its partner in b and a lane of an odd row in a.
template <int ROW>
__device__ __forceinline__ uint32_t row_xor_u32(uint32_t v)
{
    const bool odd = (threadIdx.x & ROW) != 0;
    if constexpr (ROW == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return odd ? r[0] : r[1];
    } else {
        static_assert(ROW == 32, "row swap of 16 or 32 lanes");
        const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return odd ? r[0] : r[1];
    }
}
template <int ROW>
__device__ __forceinline__ float row_xor(float v) { return __uint_as_float(row_xor_u32<ROW>(__float_as_uint(v))); }

// v + (v of the lane across the row boundary), from one row swap of v with
// a copy of itself: after v_permlane16_swap(a = v, b = v) (or the 32-lane
// swap) a lane of an even row holds its own value in a and its partner's in
// b, a lane of an odd row the reverse, so a + b is the pair sum on every lane
// (the same bits as v + partner: the two operands are the same, in either
// order).
template <int ROW>
__device__ __forceinline__ float row_pair_sum(float v)
{
    const uint32_t u = __float_as_uint(v);
    if constexpr (ROW == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(u, u, false, false);
        return __uint_as_float(r[0]) + __uint_as_float(r[1]);
    } else {
        static_assert(ROW == 32, "row swap of 16 or 32 lanes");
        const auto r = __builtin_amdgcn_permlane32_swap(u, u, false, false);
        return __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
}
template <int ROW>
__device__ __forceinline__ double row_xor(double v)
{
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = row_xor_u32<ROW>((uint32_t)b), hi = row_xor_u32<ROW>((uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// group_sum for a cluster width kl known at run time (uniform across the
// workgroup, so the branches are scalar); the same steps as group_sum<KL>,
// then for clusters of 32 and 64 lanes the sums of the two 16-lane rows and
// of the two halves of the wave.  Every step adds the same two operands on
// both lanes of a pair (an xor butterfly), so all lanes end with equal bits.
template <typename T>
__device__ __forceinline__ T group_sum_rt(T v, int kl)
{
    if (kl >= 2) v = v + dpp_perm<0xB1>(v);
    if (kl >= 4) v = v + dpp_perm<0x4E>(v);
    if (kl >= 8) v = v + dpp_perm<0x141>(v);
    if (kl >= 16) v = v + dpp_perm<0x140>(v);
    if constexpr (sizeof(T) == 4) {
        if (kl >= 32) v = row_pair_sum<16>(v);
        if (kl >= 64) v = row_pair_sum<32>(v);
    } else {
        if (kl >= 32) v = v + row_xor<16>(v);
        if (kl >= 64) v = v + row_xor<32>(v);
    }
    return v;
}

// pair_sum64 for a run-time cluster width: rows j of this lane, partners
// l > j, cluster sum.  Positions scaled by (px, px, pz), squared distances in
// box bx (minimum image in a periodic box) times m2; NPT keeps pairs with
// r^2 <= rc.  The same operations and order as pair_sum64.
template <typename Real, int PK, int KC, int BC, int ENS, int LANES>
__device__ double spread_pair_sum64(const Real *X, int gb, int q, int N, int kl, int kls, const BoxC<double> &bx,
                                    double px, double pz, double m2, double rc, const double (&c64)[KC])
{
    const int A = (N + kl - 1) >> kls;
    double part = 0.0;
    for (int s = 0; s < A; ++s) {
        const int j = (s << kls) + q;
        if (j >= N) break;
        const double xj = (double)X[xoff<LANES>(s, 0, gb + q)] * px;
        const double yj = (double)X[xoff<LANES>(s, 1, gb + q)] * px;
        const double zj = (double)X[xoff<LANES>(s, 2, gb + q)] * pz;
        for (int l = j + 1; l < N; ++l) {
            const int sl = l >> kls, ll = gb + (l & (kl - 1));
            const double xl = (double)X[xoff<LANES>(sl, 0, ll)] * px;
            const double yl = (double)X[xoff<LANES>(sl, 1, ll)] * px;
            const double zl = (double)X[xoff<LANES>(sl, 2, ll)] * pz;
            double r2 = box_d2<double, BC, false>(xj, yj, zj, xl, yl, zl, bx);
            r2 = r2 * m2;
            if (ENS == ENS_NPT) {
                if (r2 <= rc) part += pair_energy64<PK, KC>(r2, c64);
            } else {
                part += pair_energy64<PK, KC>(r2, c64);
            }
        }
    }
    return group_sum_rt(part, kl);
}

// wave_pair_sum64 for a run-time cluster width: the NPT pair sum of the
// configuration of the cluster whose first lane is wave lane `cl`, computed
// by every executing lane of the wave together; pairs (j, j + d mod N) dealt
// round-robin over the `live` lanes, cluster sums, then the totals of the
// `leaders` in lane order (the same operations and order as wave_pair_sum64).
template <int PK, int KC, int BC, int LANES, typename Real>
__device__ double wave_pair_sum64_rt(const Real *X, int cl, int N, int kl, int kls, const BoxC<double> &bx,
                                     double m2, double rc, const double (&c64)[KC], uint64_t live, uint64_t leaders)
{
    const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(live >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)live, 0u));
    const int nl = __popcll(live);
    const int cb = (threadIdx.x & ~63) + cl;
    auto pos = [&](int j, double &x, double &y, double &z) {
        const int sj = j >> kls, lj = cb + (j & (kl - 1));
        x = (double)X[xoff<LANES>(sj, 0, lj)] * 1.0;
        y = (double)X[xoff<LANES>(sj, 1, lj)] * 1.0;
        z = (double)X[xoff<LANES>(sj, 2, lj)] * 1.0;
    };
    auto energy = [&](double xj, double yj, double zj, double xl, double yl, double zl) -> double {
        const double r2 = box_d2<double, BC, false>(xj, yj, zj, xl, yl, zl, bx) * m2;
        return r2 <= rc ? pair_energy64_vol<PK, KC>(r2, c64) : 0.0;
    };
    double part = wave_pair_walk(N, rank, nl, pos, energy);
    part = group_sum_rt(part, kl);
    double tot = 0.0;
    for (uint64_t b = leaders; b != 0; b &= b - 1) tot += readlane_d(part, __builtin_ctzll(b));
    return tot;
}

// sphere_redraw for a run-time cluster width (src/MCMoves.jl:115-135).
template <typename Real, int LANES>
__device__ bool sphere_redraw_rt(Real *X, int tid, int q, int N, int kl, int kls, int A, int xo,
                                 Real radius2, Real delta_r, uint64_t seed, uint32_t run_key, uint32_t temp,
                                 uint32_t sweep, uint32_t m, uint32_t nsteps, Real &xi, Real &yi, Real &zi, Real &tx,
                                 Real &ty, Real &tz)
{
    int count = 0;
    uint32_t attempt = 0;
    bool ok = true;
    while (outside_sphere<Real>(tx, ty, tz, radius2)) {
        count += 1;
        if (count == 50) {
            // every lane of the cluster forms the same sequential mean, atom by atom
            const int gb = tid - q;
            double sx = 0.0, sy = 0.0, sz = 0.0;
            for (int j = 0; j < N; ++j) {
                const int sj = j >> kls, lj = gb + (j & (kl - 1));
                sx += (double)X[xoff<LANES>(sj, 0, lj)];
                sy += (double)X[xoff<LANES>(sj, 1, lj)];
                sz += (double)X[xoff<LANES>(sj, 2, lj)];
            }
            sx /= (double)N;
            sy /= (double)N;
            sz /= (double)N;
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            for (int s = 0; s < A; ++s) {
                if ((s << kls) + q >= N) break;
                X[xoff<LANES>(s, 0, tid)] = (Real)((double)X[xoff<LANES>(s, 0, tid)] - sx);
                X[xoff<LANES>(s, 1, tid)] = (Real)((double)X[xoff<LANES>(s, 1, tid)] - sy);
                X[xoff<LANES>(s, 2, tid)] = (Real)((double)X[xoff<LANES>(s, 2, tid)] - sz);
            }
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
        } else {
            attempt += 1;
            MoveDraw<Real> dd = move_draw<Real>(seed, run_key, temp, sweep, m, attempt, nsteps);
            xi = X[xo];
            yi = X[xo + LANES];
            zi = X[xo + 2 * LANES];
            make_trial<Real, BC_SPHERE, false>(xi, yi, zi, dd, delta_r, Real(0), Real(0), tx, ty, tz);
            if (count == 100) {
                ok = false;
                break;
            }
        }
    }
    xi = X[xo];
    yi = X[xo + LANES];
    zi = X[xo + 2 * LANES];
    return ok;
}

// One cycle of moves of every configuration of the launch's tiles.  With
// n_cycles = 0 the launch only applies a pending FP64 recompute.
//
// The FP64 recompute that sweep_kernel makes at the end of cycle c (after
// its exchange and sampling) is made here at the start of the launch of cycle
// c + 1, when sweep0 % recompute_every == 0; the engine adds a launch with
// n_cycles = 0 after a call's last cycle so that it is never left pending,
// and marks the next call's first launch (recompute_done) not to repeat it.
// BC_CUBIC serves cubic and rectangular boxes (box kind at run time); NPT
// adds the uniform volume move as the cycle's (N + 1)-th move
// (src/MCMoves.jl:169-177, :277-287; energy src/EnergyEvaluation.jl:397-474;
// acceptance src/Exchange.jl:58-71).
//
// MC (exchange after every k-th cycle, k > 1): one launch runs the cycles up
// to the next exchange; the temperature label of a configuration stays in
// between, so its step sizes, window counters and energy sums live in
// registers and the cycles without an exchange adapt and sample here.  MC =
// false compiles the one-cycle launch of the default cadence without that
// state (fewer registers, more waves per SIMD).
template <typename Real, int PK, int KC, int BC, int ENS, int LANES, bool MC>
__global__ void __launch_bounds__(LANES) spread_sweep_kernel(SweepArgs a)
{
    static_assert((BC == BC_SPHERE && ENS == ENS_NVT) || BC == BC_CUBIC,
                  "spread layout: spherical container (NVT) or cubic / rectangular boxes");
    constexpr bool F32 = sizeof(Real) == 4;
    constexpr bool PBC = BC != BC_SPHERE;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const TileInfo tile = a.tiles[blockIdx.x];
    const int KL = tile.kl;
    const int kls = __builtin_ctz((unsigned)KL);
    const int tid = threadIdx.x;
    const int lg = tid >> kls;
    const int q = tid & (KL - 1);
    const int gb = lg << kls;
    const bool active = lg < tile.n_slots;
    const int N = tile.n_atoms;
    const int A = tile.n_slab;
    const uint32_t nsteps = (uint32_t)(ENS == ENS_NPT ? N + 1 : N);
    const int kind = PBC ? a.box_kind : (int)BC_SPHERE;
    Real *X = reinterpret_cast<Real *>(lds_raw);

    Real c[KC];
    double c64[KC];
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        const double v = k < a.pot.ncoeff ? a.pot.coeff[k] : 0.0;
        c64[k] = v;
        c[k] = (Real)v;
    }
    if constexpr (PK == PK_LJ) {
        c[0] = (Real)a.pot.coeff[0];
        c[1] = (Real)a.pot.coeff[3];
    }

    Real *gX = reinterpret_cast<Real *>(a.coords) + tile.coord_off;
    for (int e = tid; e < 3 * A * LANES; e += LANES) X[e] = gX[e];
    __syncthreads();

    const int slot = tile.slot0 + lg;
    if (active) {
        const RunInfo &ri = a.runs[a.slot_run[slot]];
        const int tb = ri.temp_base;
        const uint32_t run_key = ri.run_key;
        const Real radius2 = (Real)ri.radius2;
        const int t = a.slot_temp[slot];
        double E = a.slot_en[slot];
        double box = 1.0, boxh = 1.0;
        if constexpr (PBC) {
            box = a.slot_box[slot];
            boxh = a.slot_boxh[slot];
        }
        // The temperature label stays until the exchange after the launch's
        // last cycle (MC: its step sizes, window counters and energy sums in
        // registers, every lane of the cluster holding the same values, back
        // to memory at the end).
        const int T = tb + t;
        const double beta = a.t_beta[T];
        double delta = a.t_displ[2 * T];
        double dvol = ENS == ENS_NPT ? a.t_displ[2 * T + 1] : 0.0;
        long long win_atom = 0, add_atom = 0, win_vol = 0, add_vol = 0;
        double ham0 = 0.0, ham1 = 0.0;
        if constexpr (MC) {
            win_atom = a.t_count_atom[2 * T + 1];
            if (ENS == ENS_NPT) win_vol = a.t_count_vol[2 * T + 1];
            ham0 = a.t_ham[2 * T];
            ham1 = a.t_ham[2 * T + 1];
        }
        // n_cycles = 0: the pending recompute only
        const int ncyc = MC ? (a.n_cycles > 0 ? a.n_cycles : 1) : 1;
        for (int cc = 0; cc < ncyc; ++cc) {
          const uint32_t sweep = (uint32_t)(a.sweep0 + cc);
          // FP64 recompute due after the previous cycle (after its sampling)
          if ((cc > 0 || !a.recompute_done) && a.recompute_every > 0 && sweep > 0u &&
              sweep % (uint32_t)a.recompute_every == 0u) {
            const double rc64 = PBC ? box_rcut(kind, box, boxh) : 0.0;
            double ef = spread_pair_sum64<Real, PK, KC, BC, ENS, LANES>(X, gb, q, N, KL, kls,
                                                                       make_box<double>(box, boxh), 1.0, 1.0, 1.0,
                                                                       rc64, c64);
            if constexpr (ENS == ENS_NPT) ef = ef + box_lrc<PK, KC>(kind, N, rc64, box, boxh, c64);
            const double dr = __builtin_fabs(E - ef);
            if (q == 0 && dr > a.slot_drift[slot]) a.slot_drift[slot] = dr;
            E = ef;
          }
          if (a.n_cycles > 0) {
            const Real delta_r = (Real)delta;
            const float beta_l2e = (float)beta * 1.4426950408889634f;
            // decoded draws carry the Metropolis threshold -log2(u) / (beta log2 e):
            // exp2(-beta log2e dE) >= u  <=>  dE <= threshold
            const float inv_bl2e = 1.0f / beta_l2e;
            // binary32 displacement (u - 1/2) delta as one fma of the 24-bit
            // integer of u: u - 1/2 and the products by powers of two are
            // exact, so the result is the rounding of the same exact value
            const float dsc = (float)delta_r * 0x1.0p-24f, dof = -0.5f * (float)delta_r;
            BoxC<Real> bxr{};
            Real rc = 0;
            if constexpr (PBC) {
                bxr = make_box<Real>(box, boxh);
                rc = (Real)box_rcut(kind, box, boxh);
            }
            int nacc = 0, nvacc = 0;
            u32x4 mA = {0, 0, 0, 0}, mB = {0, 0, 0, 0};
            u32x4 nA = {0, 0, 0, 0}, nB = {0, 0, 0, 0};
            double mthr = 0.0, nthr = 0.0;  // FP64: Metropolis thresholds of the lane's and the next move
            // binary32 in the spherical container: the lane that draws a move's
            // words also decodes them (LDS offset of the atom, the displacement,
            // the Metropolis uniform), so per move the cluster's lanes fetch five
            // ready values instead of decoding the words each
            constexpr bool PRE = F32 && !PBC;
            const int gb4 = (gb & 63) << 2;  // byte address of the cluster's first lane for ds_bpermute
            uint32_t p_off = 0u;
            float p_dx = 0.0f, p_dy = 0.0f, p_dz = 0.0f, p_um = 0.0f;
            if (!PRE && KL > 1) {
                if ((uint32_t)q < nsteps) {
                    move_words<Real>(a.seed, run_key, (uint32_t)t, sweep, (uint32_t)q, mA, mB);
                    if constexpr (!F32) mthr = metro_thr(u53(mB.z, mB.w), beta);
                }
                fetch_words<Real>(mA, mB, gb, nA, nB);
                if constexpr (!F32) nthr = lane_get_d(mthr, gb);
            }
            for (uint32_t m = 0; m < nsteps; ++m) {
                // draws of this move (src/MCRun.jl:145, src/MCMoves.jl:34-38, src/MCRun.jl:123)
                u32x4 cur = nA, curB = nB;
                double thr = nthr;
                MoveDraw<Real> d0{};
                uint32_t i = 0u;  // PRE: formed from xo where needed
                int xo = 0;       // LDS offset of the moved atom's x
                float pdx = 0.0f, pdy = 0.0f, pdz = 0.0f, fthr = 0.0f;
                if constexpr (PRE) {
                    const uint32_t k = m & (uint32_t)(KL - 1);
                    if (k == 0u && m + (uint32_t)q < nsteps) {
                        u32x4 w, wb;
                        move_words<Real>(a.seed, run_key, (uint32_t)t, sweep, m + (uint32_t)q, w, wb);
                        const MoveDraw<Real> dw = decode_move<Real>(w, w, nsteps);
                        p_off = (uint32_t)xoff<LANES>((int)(dw.index >> kls), 0,
                                                      gb + (int)(dw.index & (uint32_t)(KL - 1)));
                        p_dx = __builtin_fmaf((float)(w.y >> 8), dsc, dof);
                        p_dy = __builtin_fmaf((float)(w.z >> 8), dsc, dof);
                        p_dz = __builtin_fmaf((float)(w.w >> 8), dsc, dof);
                        p_um = -__builtin_amdgcn_logf((float)dw.umet) * inv_bl2e;
                    }
                    // (also with one lane per configuration: LDS-pipe permutes
                    // in place of register copies on the VALU; the cluster lies
                    // inside one wave, so the source never wraps)
                    const int src4 = gb4 + (int)(k << 2);
                    xo = __builtin_amdgcn_ds_bpermute(src4, (int)p_off);
                    pdx = __int_as_float(__builtin_amdgcn_ds_bpermute(src4, __float_as_int(p_dx)));
                    pdy = __int_as_float(__builtin_amdgcn_ds_bpermute(src4, __float_as_int(p_dy)));
                    pdz = __int_as_float(__builtin_amdgcn_ds_bpermute(src4, __float_as_int(p_dz)));
                    fthr = __int_as_float(__builtin_amdgcn_ds_bpermute(src4, __float_as_int(p_um)));
                } else if (KL == 1) {
                    move_words<Real>(a.seed, run_key, (uint32_t)t, sweep, m, cur, curB);
                    if constexpr (!F32) thr = metro_thr(u53(curB.z, curB.w), beta);
                } else {
                    const uint32_t m1 = m + 1;
                    if (m1 < nsteps) {
                        const uint32_t k1 = m1 & (uint32_t)(KL - 1);
                        if (k1 == 0 && m1 + q < nsteps) {
                            move_words<Real>(a.seed, run_key, (uint32_t)t, sweep, m1 + q, mA, mB);
                            if constexpr (!F32) mthr = metro_thr(u53(mB.z, mB.w), beta);
                        }
                        fetch_words<Real>(mA, mB, gb + (int)k1, nA, nB);
                        if constexpr (!F32) nthr = lane_get_d(mthr, gb + (int)k1);
                    }
                }
                if constexpr (!PRE) {
                    d0 = decode_move<Real>(cur, curB, nsteps);
                    i = d0.index;
                    if (i < (uint32_t)N) xo = xoff<LANES>((int)(i >> kls), 0, gb + (int)(i & (uint32_t)(KL - 1)));
                }
                bool accept = false;
                double dE_used = 0.0;
                float dE_f32 = 0.0f;  // binary32 atom moves: dE, widened only where traced
                bool atom_f32 = false;
                // NVT: every index is an atom; NPT: index N is the volume move
                if (PRE || i < (uint32_t)N) {
                    const int li = xo & (LANES - 1);
                    Real xi = X[xo];
                    Real yi = X[xo + LANES];
                    Real zi = X[xo + 2 * LANES];
                    Real tx, ty, tz;
                    if constexpr (PBC) {
                        box_trial<Real, BC, false>(xi, yi, zi, d0, delta_r, bxr, tx, ty, tz);
                    } else if constexpr (F32) {
                        tx = xi + pdx;
                        ty = yi + pdy;
                        tz = zi + pdz;
                    } else {
                        make_trial<Real, BC_SPHERE, false>(xi, yi, zi, d0, delta_r, Real(1), Real(1), tx, ty, tz);
                    }
                    if constexpr (!PBC) {
                        // src/MCMoves.jl:115-135: redraw until inside, recentre at 50, error at 100
                        if (outside_sphere<Real>(tx, ty, tz, radius2)) {
                            if (!sphere_redraw_rt<Real, LANES>(X, tid, q, N, KL, kls, A, xo, radius2, delta_r,
                                                               a.seed, run_key, (uint32_t)t, sweep, m, nsteps, xi, yi,
                                                               zi, tx, ty, tz) &&
                                q == 0)
                                a.run_error[a.slot_run[slot]] = 1;
                        }
                        // park the moved atom so its self pair contributes zero
                        if (tid == li) {
                            X[xo] = far_point<Real>();
                            X[xo + LANES] = far_point<Real>();
                            X[xo + 2 * LANES] = far_point<Real>();
                        }
                        __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    }
                    // dE = sum_j e(r'_ij) - sum_j e(r_ij), src/EnergyEvaluation.jl:804-821
                    Real sn = 0, so = 0;  // FP64: the two sums; binary32: sum of differences in sn
                    const Real *Xl = X + tid;
                    // binary32 sphere: trial distances from the old ones (dist2_pair_f32)
                    float m2dx = 0.0f, m2dy = 0.0f, m2dz = 0.0f, ddd = 0.0f;
                    if constexpr (F32 && !PBC) {
                        const float dx = (float)tx - (float)xi, dy = (float)ty - (float)yi, dz = (float)tz - (float)zi;
                        m2dx = -2.0f * dx;
                        m2dy = -2.0f * dy;
                        m2dz = -2.0f * dz;
                        ddd = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                    }
#pragma unroll 2
                    for (int s = 0; s < A; ++s) {
                        const Real xj = Xl[xoff<LANES>(s, 0, 0)];
                        const Real yj = Xl[xoff<LANES>(s, 1, 0)];
                        const Real zj = Xl[xoff<LANES>(s, 2, 0)];
                        if constexpr (!PBC) {
                            Real r2n, r2o;
                            if constexpr (F32) {
                                float fn, fo;
                                dist2_pair_f32((float)xj, (float)yj, (float)zj, (float)xi, (float)yi, (float)zi, m2dx,
                                               m2dy, m2dz, ddd, fn, fo);
                                r2n = (Real)fn;
                                r2o = (Real)fo;
                            } else {
                                r2n = box_d2<Real, BC_SPHERE, false>(tx, ty, tz, xj, yj, zj, BoxC<Real>{});
                                r2o = box_d2<Real, BC_SPHERE, false>(xi, yi, zi, xj, yj, zj, BoxC<Real>{});
                            }
                            if constexpr (F32 && PK == PK_LJ) {
                                // e(u) = u^3 (c12 u^3 + c6): e_new - e_old = (a - b)(c12 (a + b) + c6),
                                // a = u_new^3, b = u_old^3; one reciprocal serves both pairs
                                // (a reciprocal issues at a fraction of the multiply rate)
                                const float ip = __builtin_amdgcn_rcpf(r2n * r2o);
                                const float un = ip * r2o, uo = ip * r2n;
                                const float an = un * un * un, bo = uo * uo * uo;
                                sn = __builtin_fmaf(an - bo, __builtin_fmaf(c[1], an + bo, c[0]), sn);
                            } else if constexpr (F32 && PK == PK_EVEN_H) {
                                const float ip = __builtin_amdgcn_rcpf(r2n * r2o);
                                sn += pair_energy_u<Real, PK, KC>(ip * r2o, c) - pair_energy_u<Real, PK, KC>(ip * r2n, c);
                            } else if constexpr (F32) {
                                sn += pair_energy<Real, PK, KC>(r2n, c) - pair_energy<Real, PK, KC>(r2o, c);
                            } else {
                                sn += pair_energy<Real, PK, KC>(r2n, c);
                                so += pair_energy<Real, PK, KC>(r2o, c);
                            }
                        } else {
                            // minimum image; the moved atom and padding atoms are masked, NPT
                            // keeps pairs inside the cut-off (src/EnergyEvaluation.jl:323-336)
                            const Real r2n = box_d2<Real, BC, false>(tx, ty, tz, xj, yj, zj, bxr);
                            const Real r2o = box_d2<Real, BC, false>(xi, yi, zi, xj, yj, zj, bxr);
                            Real en, eo;
                            if constexpr (F32 && (PK == PK_LJ || PK == PK_EVEN_H)) {
                                const float ip = __builtin_amdgcn_rcpf(r2n * r2o);
                                en = pair_energy_u<Real, PK, KC>(ip * r2o, c);
                                eo = pair_energy_u<Real, PK, KC>(ip * r2n, c);
                            } else {
                                en = pair_energy<Real, PK, KC>(r2n, c);
                                eo = pair_energy<Real, PK, KC>(r2o, c);
                            }
                            const int j = (s << kls) + q;
                            const bool ok = (j < N) & (j != (int)i);
                            bool okn = ok, oko = ok;
                            if constexpr (ENS == ENS_NPT) {
                                okn = ok & (r2n <= rc);
                                oko = ok & (r2o <= rc);
                            }
                            if constexpr (F32) {
                                sn += (okn ? en : Real(0)) - (oko ? eo : Real(0));
                            } else {
                                sn += okn ? en : Real(0);
                                so += oko ? eo : Real(0);
                            }
                        }
                    }
                    if constexpr (F32) {
                        // Metropolis in binary32 (src/Exchange.jl:53-57, src/MCRun.jl:122-132)
                        const float dEf = group_sum_rt((float)sn, KL);
                        dE_f32 = dEf;
                        atom_f32 = true;
                        if constexpr (PRE) {
                            accept = dEf <= fthr;  // the threshold is >= 0: dE <= 0 accepts
                        } else {
                            const float pr = __builtin_amdgcn_exp2f(-dEf * beta_l2e);
                            accept = (dEf <= 0.0f) || (pr >= (float)d0.umet);
                        }
                        // E + 0 = E: the same bits as adding only on accept; the select on
                        // the binary32 value (the empty asm keeps the compiler from moving
                        // the widening before it: one select instead of two)
                        float dacc = accept ? dEf : 0.0f;
                        __asm__("" : "+v"(dacc));
                        E = E + (double)dacc;
                        nacc += (int)accept;
                    } else {
                        sn = group_sum_rt(sn, KL);
                        so = group_sum_rt(so, KL);
                        const double new_en = (sn - so) + E;
                        const double dd = new_en - E;
                        dE_used = dd;
                        accept = dd <= thr;
                        if (accept) {
                            E = new_en;
                            nacc += 1;
                        }
                    }
                    // commit (src/swap_config.jl:40-51) or restore the parked atom
                    if (tid == li) {
                        X[xo] = accept ? tx : xi;
                        X[xo + LANES] = accept ? ty : yi;
                        X[xo + 2 * LANES] = accept ? tz : zi;
                    }
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                }
                // uniform volume move: r^2 of the current positions scaled by s^2
                // (binary32: skipped by a wave none of whose clusters makes one)
                if (ENS == ENS_NPT && (!F32 || __ballot(i >= (uint32_t)N) != 0)) {
                    const bool vm = i >= (uint32_t)N;
                    double sx = 1.0;
                    if (vm) {
                        const double y = ((double)d0.ux - 0.5) * dvol;
                        double scale = det_exp(y / 3.0);
                        if (box >= a.t_maxbox[T] && scale > 1.0) scale = 1.0;
                        sx = scale;
                    }
                    const double nbox = box * sx, nh = boxh * sx;
                    const double nrc = box_rcut(kind, nbox, nh);
                    const double m2 = sx * sx;
                    double new_en = 0.0;
                    if constexpr (F32) {
                        // the clusters of a wave evaluate each pending volume move together
                        __atomic_signal_fence(__ATOMIC_SEQ_CST);
                        const uint64_t live = __ballot(1);
                        const uint64_t leaders = __ballot(q == 0);
                        uint64_t todo = __ballot(vm && q == 0);
                        while (todo != 0) {
                            const int cl = __builtin_ctzll(todo);
                            todo &= todo - 1;
                            const double e = wave_pair_sum64_rt<PK, KC, BC, LANES>(
                                X, cl, N, KL, kls, make_box<double>(readlane_d(box, cl), readlane_d(boxh, cl)),
                                readlane_d(m2, cl), readlane_d(nrc, cl), c64, live, leaders);
                            if ((gb & 63) == cl) new_en = e;
                        }
                    } else {
                        if (vm)
                            new_en = spread_pair_sum64<Real, PK, KC, BC, ENS, LANES>(
                                X, gb, q, N, KL, kls, make_box<double>(box, boxh), 1.0, 1.0, m2, nrc, c64);
                    }
                    if (vm) {
                        new_en = new_en + box_lrc<PK, KC>(kind, N, nrc, nbox, nh, c64);
                        const double pressure = ri.pressure;
                        const double dd = new_en - E;
                        dE_used = dd;
                        const double vn = box_volume(kind, nbox, nh), vo = box_volume(kind, box, boxh);
                        const double dh = dd + pressure * (vn - vo);
                        const double arg = -dh * beta + (double)N * det_log(vn / vo);
                        double prob = det_exp(arg);
                        if (prob > 1.0) prob = 1.0;
                        accept = prob >= (double)d0.umet;
                        if (accept) {
                            for (int s = 0; s < A; ++s) {
                                Real *pxp = &X[xoff<LANES>(s, 0, tid)];
                                Real *pyp = &X[xoff<LANES>(s, 1, tid)];
                                Real *pzp = &X[xoff<LANES>(s, 2, tid)];
                                *pxp = (Real)((double)*pxp * sx);
                                *pyp = (Real)((double)*pyp * sx);
                                *pzp = (Real)((double)*pzp * sx);
                            }
                            __atomic_signal_fence(__ATOMIC_SEQ_CST);
                            box = nbox;
                            boxh = nh;
                            E = new_en;
                            nvacc += 1;
                            bxr = make_box<Real>(box, boxh);
                            rc = (Real)box_rcut(kind, box, boxh);
                        }
                    }
                }
                if (a.trace_de != nullptr && q == 0 && a.trace_cycle + cc < a.trace_cycles) {
                    const size_t ti = ((size_t)(tb + t) * a.trace_cycles + a.trace_cycle + cc) * nsteps + m;
                    a.trace_de[ti] = atom_f32 ? (double)dE_f32 : dE_used;
                    a.trace_acc[ti] = (int8_t)accept;
                    a.trace_idx[ti] = PRE ? (int32_t)(((xo / (3 * LANES)) << kls) + ((xo & (LANES - 1)) - gb))
                                          : (int32_t)i;
                }
            }
            if constexpr (!MC) {
                if (q == 0) {
                    a.t_count_atom[2 * T] += nacc;
                    a.t_count_atom[2 * T + 1] += nacc;
                    if (ENS == ENS_NPT) {
                        a.t_count_vol[2 * T] += nvacc;
                        a.t_count_vol[2 * T + 1] += nvacc;
                    }
                }
            } else {
            win_atom += nacc;
            add_atom += nacc;
            win_vol += nvacc;
            add_vol += nvacc;
            if ((sweep + 1u) % (uint32_t)a.exchange_every != 0u) {
                // no exchange after this cycle: its step-size adaptation and
                // sampling here (src/Exchange.jl:198-263, src/Sampling.jl:29-125,
                // :258-280, :330-371), as spread_cycle_kernel makes them
                const long long pidx = a.cycle0 + cc + 1;
                if (pidx % a.n_adjust == 0) {
                    const double ar = (double)win_atom / (double)((long long)a.n_adjust * N);
                    if (ar < a.min_acc) delta *= 0.9;
                    else if (ar > a.max_acc) delta *= 1.1;
                    win_atom = 0;
                    if (ENS == ENS_NPT) {
                        const double av = (double)win_vol / (double)a.n_adjust;
                        if (av < a.min_acc) dvol *= 0.9;
                        else if (av > a.max_acc) dvol *= 1.1;
                        win_vol = 0;
                    }
                }
                if (a.sampling && pidx % a.mc_sample == 0) {
                    if (ENS == ENS_NPT) {
                        const double h = E + ri.pressure * box_volume(kind, box, boxh);
                        ham0 += h;
                        ham1 += h * h;
                        const double ie = __builtin_floor((E - ri.en_min) / ri.delta_en) + 1.0;
                        const double iv =
                            __builtin_floor((box_hist_volume(kind, box, boxh) - ri.v_min) / ri.delta_v) + 1.0;
                        const int he = ie < 1.0 ? 1 : (ie > (double)a.n_bin ? a.n_bin + 2 : (int)__builtin_floor(ie) + 1);
                        const int hv = iv < 1.0 ? 1 : (iv > (double)a.n_bin ? a.n_bin + 2 : (int)__builtin_floor(iv) + 1);
                        if (q == 0)
                            atomicAdd(&a.t_hist[(size_t)T * a.hist_stride + (size_t)(he - 1) * (a.n_bin + 2) + (hv - 1)],
                                      1u);
                    } else {
                        ham0 += E;
                        ham1 += E * E;
                        const double hi = (E - ri.en_min) / ri.delta_en + 1.0;
                        const int h = hi < 1.0 ? 1 : (hi > (double)a.n_bin ? a.n_bin + 2 : (int)__builtin_floor(hi) + 1);
                        if (q == 0) atomicAdd(&a.t_hist[(size_t)T * a.hist_stride + (h - 1)], 1u);
                    }
                }
            }
            }
          }
        }
        if (q == 0) {
            a.slot_en[slot] = E;
            if constexpr (PBC) {
                a.slot_box[slot] = box;
                a.slot_boxh[slot] = boxh;
            }
            if (MC && a.n_cycles > 0) {
                a.t_displ[2 * T] = delta;
                a.t_count_atom[2 * T] += add_atom;
                a.t_count_atom[2 * T + 1] = win_atom;
                if (ENS == ENS_NPT) {
                    a.t_displ[2 * T + 1] = dvol;
                    a.t_count_vol[2 * T] += add_vol;
                    a.t_count_vol[2 * T + 1] = win_vol;
                }
                if (a.sampling) {
                    a.t_ham[2 * T] = ham0;
                    a.t_ham[2 * T + 1] = ham1;
                }
            }
        }
    }
    __syncthreads();
    if (a.n_cycles > 0)
        for (int e = tid; e < 3 * A * LANES; e += LANES) gX[e] = X[e];
}

// initialise_energy (src/EnergyEvaluation.jl:1662-1692) for the spread
// layout: FP64 total energy of each configuration of the tile (with the
// box's cut-off and tail correction in NPT).
template <typename Real, int PK, int KC, int BC, int ENS, int LANES>
__global__ void __launch_bounds__(LANES) spread_init_kernel(SweepArgs a)
{
    constexpr bool PBC = BC != BC_SPHERE;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const TileInfo tile = a.tiles[blockIdx.x];
    const int KL = tile.kl;
    const int kls = __builtin_ctz((unsigned)KL);
    const int tid = threadIdx.x;
    const int lg = tid >> kls;
    const int q = tid & (KL - 1);
    const int N = tile.n_atoms;
    const int A = tile.n_slab;
    Real *X = reinterpret_cast<Real *>(lds_raw);
    const Real *gX = reinterpret_cast<const Real *>(a.coords) + tile.coord_off;
    for (int e = tid; e < 3 * A * LANES; e += LANES) X[e] = gX[e];
    __syncthreads();
    if (lg >= tile.n_slots) return;  // whole clusters leave together
    const int slot = tile.slot0 + lg;
    if (a.init_run >= 0 && a.slot_run[slot] != a.init_run) return;
    double c64[KC];
#pragma unroll
    for (int k = 0; k < KC; ++k) c64[k] = k < a.pot.ncoeff ? a.pot.coeff[k] : 0.0;
    const int kind = PBC ? a.box_kind : (int)BC_SPHERE;
    const double box = PBC ? a.slot_box[slot] : 1.0, boxh = PBC ? a.slot_boxh[slot] : 1.0;
    const double rc64 = PBC ? box_rcut(kind, box, boxh) : 0.0;
    double E = spread_pair_sum64<Real, PK, KC, BC, ENS, LANES>(X, lg << kls, q, N, KL, kls,
                                                              make_box<double>(box, boxh), 1.0, 1.0, 1.0, rc64, c64);
    if constexpr (ENS == ENS_NPT) E = E + box_lrc<PK, KC>(kind, N, rc64, box, boxh, c64);
    if (q == 0) a.slot_en[slot] = E;
}

// Exchange, step-size adaptation and sampling of one cycle of every ladder
// (src/MCRun.jl:203-212, :248-250), one 64-lane workgroup per run, after the
// moves of a cycle that ends with an exchange (every cycle by default).  The ladder's energies (NPT: and volumes) and slot map are
// staged in LDS; the even/odd pairs of a cycle are disjoint, so lanes take
// them in parallel.  Draws, tests and counters are those of the exchange in
// sweep_kernel (src/Exchange.jl:103-155, :198-263; src/Sampling.jl:29-125,
// :258-280).
__global__ void __launch_bounds__(64) spread_cycle_kernel(SweepArgs a)
{
    __shared__ double s_e[SPREAD_MAX_TEMPS];
    __shared__ double s_v[SPREAD_MAX_TEMPS];
    __shared__ int32_t s_slot[SPREAD_MAX_TEMPS];
    const RunInfo ri = a.runs[a.run_list != nullptr ? a.run_list[blockIdx.x] : (int)blockIdx.x];
    const int R = ri.n_temps, tb = ri.temp_base, N = ri.n_atoms;
    const bool npt = a.ensemble == ENS_NPT;
    const int kind = a.box_kind;
    const int lane = threadIdx.x;
    const uint32_t sweep = (uint32_t)a.sweep0;
    const long long pidx = a.cycle0 + 1;
    for (int t = lane; t < R; t += 64) {
        const int s = a.temp_slot[tb + t];
        s_slot[t] = s;
        s_e[t] = a.slot_en[s];
        s_v[t] = npt ? box_volume(kind, a.slot_box[s], a.slot_boxh[s]) : 0.0;
    }
    __syncthreads();
    auto try_pair = [&](int n) {
        const u32x4 X1 = draw(a.seed, ri.run_key, RNG_EXCH, 0u, 0u, sweep, 1u + (uint32_t)n);
        const double u = u53(X1.x, X1.y);
        double e1 = s_e[n], e2 = s_e[n + 1];
        if (npt) {
            // enthalpies E + PV (src/Exchange.jl:115-124)
            e1 = e1 + ri.pressure * s_v[n];
            e2 = e2 + ri.pressure * s_v[n + 1];
        }
        const double de = e1 - e2;
        const double db = a.t_beta[tb + n] - a.t_beta[tb + n + 1];
        double acc = det_exp(db * de);
        if (acc > 1.0) acc = 1.0;
        a.t_count_exc[2 * (tb + n)] += 1;
        a.t_count_exc[2 * (tb + n + 1)] += 1;
        if (acc > u) {
            a.t_count_exc[2 * (tb + n) + 1] += 1;
            a.t_count_exc[2 * (tb + n + 1) + 1] += 1;
            const int s0 = s_slot[n];
            s_slot[n] = s_slot[n + 1];
            s_slot[n + 1] = s0;
            const double e0 = s_e[n];
            s_e[n] = s_e[n + 1];
            s_e[n + 1] = e0;
            const double v0 = s_v[n];
            s_v[n] = s_v[n + 1];
            s_v[n + 1] = v0;
        }
    };
    if (R > 1 && (sweep + 1u) % (uint32_t)a.exchange_every == 0u) {
        if (a.exchange_mode == EXC_EVEN_ODD) {
            for (int n = (int)(sweep & 1u) + 2 * lane; n + 1 < R; n += 128) try_pair(n);
        } else if (lane == 0) {
            const u32x4 X0 = draw(a.seed, ri.run_key, RNG_EXCH, 0u, 0u, sweep, 0u);
            const double u = u53(X0.x, X0.y);
            if (a.exchange_mode == EXC_EVERY || u < 0.1) try_pair((int)below(X0.z, (uint32_t)(R - 1)));
        }
    }
    __syncthreads();
    for (int t = lane; t < R; t += 64) {
        const int T = tb + t;
        const int s = s_slot[t];
        a.temp_slot[T] = s;
        a.slot_temp[s] = t;
        // update_max_stepsize!, src/Exchange.jl:198-263
        if (pidx % a.n_adjust == 0) {
            const double ar = (double)a.t_count_atom[2 * T + 1] / (double)((long long)a.n_adjust * N);
            if (ar < a.min_acc) a.t_displ[2 * T] *= 0.9;
            else if (ar > a.max_acc) a.t_displ[2 * T] *= 1.1;
            a.t_count_atom[2 * T + 1] = 0;
            if (npt) {
                const double av = (double)a.t_count_vol[2 * T + 1] / (double)a.n_adjust;
                if (av < a.min_acc) a.t_displ[2 * T + 1] *= 0.9;
                else if (av > a.max_acc) a.t_displ[2 * T + 1] *= 1.1;
                a.t_count_vol[2 * T + 1] = 0;
            }
        }
        // sampling_step!, src/Sampling.jl:29-125, :258-280, :330-371
        if (a.sampling && pidx % a.mc_sample == 0) {
            const double E = s_e[t];
            if (npt) {
                const double box = a.slot_box[s], boxh = a.slot_boxh[s];
                const double h = E + ri.pressure * box_volume(kind, box, boxh);
                a.t_ham[2 * T] += h;
                a.t_ham[2 * T + 1] += h * h;
                const double ie = __builtin_floor((E - ri.en_min) / ri.delta_en) + 1.0;
                const double iv = __builtin_floor((box_hist_volume(kind, box, boxh) - ri.v_min) / ri.delta_v) + 1.0;
                const int he = ie < 1.0 ? 1 : (ie > (double)a.n_bin ? a.n_bin + 2 : (int)__builtin_floor(ie) + 1);
                const int hv = iv < 1.0 ? 1 : (iv > (double)a.n_bin ? a.n_bin + 2 : (int)__builtin_floor(iv) + 1);
                a.t_hist[(size_t)T * a.hist_stride + (size_t)(he - 1) * (a.n_bin + 2) + (hv - 1)] += 1u;
            } else {
                a.t_ham[2 * T] += E;
                a.t_ham[2 * T + 1] += E * E;
                const double hi = (E - ri.en_min) / ri.delta_en + 1.0;
                const int h = hi < 1.0 ? 1 : (hi > (double)a.n_bin ? a.n_bin + 2 : (int)__builtin_floor(hi) + 1);
                a.t_hist[(size_t)T * a.hist_stride + (h - 1)] += 1u;
            }
        }
    }
}

}  // namespace ptmc
