// Sweep kernels for the RuNNer-style neural-network potential
// (src/EnergyEvaluation.jl:1182-1275, :1593-1621, :1735-1747, :1812-1822;
// MachineLearningPotential/{Cutoff,SymmFunc,DeltaMatrix}.jl).
//
// One workgroup owns one configuration for a whole cycle.  Its positions,
// feature matrix G (nf x atoms) and per-atom energies live in LDS.  A move
// builds the change D of G caused by moving atom i (radial terms by pair,
// angular terms by partner pair (j, k) with the exponentials shared by all
// functions of one eta), evaluates the network on G + D for every atom, and
// on acceptance folds D into G.  Exchange, step adaptation and sampling run
// between cycles in nnp_cycle_kernel, so one launch of the sweep kernel is
// one cycle of every configuration.
#pragma once
#include "ptmc_surrogate.hpp"
#include "ptmc_sweep.hpp"

#include "ptmc_nnp_layout.hpp"

namespace ptmc {

struct NnpLds {
    float *X, *G, *D, *W1, *H1, *H2, *rn2, *fn, *ro2, *fo, *eat, *eatn;
    // tables copied from the host-prepared arrays: radial functions, angular
    // group descriptors and their waves, second-layer weights, b1 | b2 | w3
    float *rad, *W2, *B;
    int32_t *gd, *wg;
};

// static LDS holding the function tables (declared in each kernel)
#define NNP_TABLES_LDS                                                         \
    __shared__ float s_rad[4 * NNP_MAXF], s_w2[NNP_MAXH * NNP_MAXH], s_bias[3 * NNP_MAXH]; \
    __shared__ __align__(16) int32_t s_gdesc[NNP_MAXG * NNP_GD];              \
    __shared__ int32_t s_wgrp[(NNP_NT / 64) * (NNP_WG + 1)]

template <int NP>
__device__ inline NnpLds nnp_lds(float *base, int nf, int h1)
{
    NnpLds l;
    l.X = base;
    l.G = l.X + 3 * NP;
    l.D = l.G + nf * NP;
    l.W1 = l.D + nf * NP;
    l.H1 = l.W1 + nf * h1;
    l.H2 = l.H1 + NP * NNP_HS;
    l.rn2 = l.H2 + NP * NNP_HS;
    l.fn = l.rn2 + NP;
    l.ro2 = l.fn + NP;
    l.fo = l.ro2 + NP;
    l.eat = l.fo + NP;
    l.eatn = l.eat + NP;
    l.rad = l.W2 = l.B = nullptr;
    l.gd = l.wg = nullptr;
    return l;
}

// type-2 cutoff, MachineLearningPotential/Cutoff.jl:79-90
__device__ __forceinline__ float nnp_cutoff(float r, float rc) { return cutoff_cos2(r, rc); }

// (1 + lambda cos)^zeta of three cosines for zeta = 2^k: k squarings (k is
// uniform across the wave, so the loop is a scalar branch).
__device__ __forceinline__ void nnp_theta3(float lam, float kpow, float c0, float c1, float c2, float &p0,
                                           float &p1, float &p2)
{
    p0 = 1.0f + lam * c0;
    p1 = 1.0f + lam * c1;
    p2 = 1.0f + lam * c2;
    const int k = __builtin_amdgcn_readfirstlane((int)kpow);
    for (int s = 0; s < k; ++s) {
        p0 *= p0;
        p1 *= p1;
        p2 *= p2;
    }
}

// cosine of the angle at c between c->p and c->q (Cutoff.jl:48-58)
__device__ __forceinline__ float nnp_cos(const float *c, const float *p, const float *q, float dcp, float dcq)
{
    const float ux = p[0] - c[0], uy = p[1] - c[1], uz = p[2] - c[2];
    const float wx = q[0] - c[0], wy = q[1] - c[1], wz = q[2] - c[2];
    return (ux * wx + uy * wy + uz * wz) / (dcp * dcq);
}

// number of set bits of a 64-lane mask below this lane
__device__ __forceinline__ int mbcnt64(uint64_t m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// cosines of the triangle (a, b, c) at a, b and c (Cutoff.jl:48-58)
__device__ __forceinline__ void nnp_angles(const float *a, const float *b, const float *c, float dab, float dac,
                                           float dbc, float &ta, float &tb, float &tc)
{
    const float abx = a[0] - b[0], aby = a[1] - b[1], abz = a[2] - b[2];
    const float acx = a[0] - c[0], acy = a[1] - c[1], acz = a[2] - c[2];
    const float bcx = b[0] - c[0], bcy = b[1] - c[1], bcz = b[2] - c[2];
    ta = (abx * acx + aby * acy + abz * acz) / (dab * dac);
    tb = (-abx * bcx - aby * bcy - abz * bcz) / (dab * dbc);
    tc = (acx * bcx + acy * bcy + acz * bcz) / (dac * dbc);
}

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// DPP lane permutation of a float (all 64 lanes active)
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, true));
}

// Sum over each row of 16 lanes, left in every lane of the row: quad
// exchanges (quad_perm [1,0,3,2] and [2,3,0,1]), then half-row and row mirrors.
__device__ __forceinline__ float row_sum16(float v)
{
    v += dpp_f<0xB1>(v);
    v += dpp_f<0x4E>(v);
    v += dpp_f<0x141>(v);
    v += dpp_f<0x140>(v);
    return v;
}

// Sum over all 64 lanes, complete in lane 63: row sums, then the last lane
// of rows 0 and 2 broadcast into rows 1 and 3, then lane 31 into rows 2 and 3.
__device__ __forceinline__ float wave_sum_last(float v)
{
    v = row_sum16(v);
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x142, 0xa, 0xf, false));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x143, 0xc, 0xf, false));
    return v;
}


// Accumulate e_new (1 + lam c_new)^(2^k) + e_old (1 + lam c_old)^(2^k) into
// acc[S0 + k] for k = 0..kmax (kmax uniform; -1: no function of this lambda).
// eo carries the sign of the old term.
template <int S0>
__device__ __forceinline__ void nnp_chain(int kmax, float lam, float cn, float co, float en, float eo, float *acc)
{
    if (kmax < 0) return;
    float pn = 1.0f + lam * cn, po = 1.0f + lam * co;
#pragma unroll
    for (int k = 0; k < NNP_KSLOT; ++k) {
        if (k <= kmax) {
            acc[S0 + k] = __builtin_fmaf(en, pn, acc[S0 + k]);
            acc[S0 + k] = __builtin_fmaf(eo, po, acc[S0 + k]);
            pn *= pn;
            po *= po;
        }
    }
}

// One eta group of angular functions as the move update reads it: e^{-eta x}
// = 2^{-eta2 x}; slots (lambda, k) = (+1, 0..4), (-1, 0..4); kp / km the
// largest k per lambda (-1: none); per slot tpz and the row (-1: none).
struct NnpGroup {
    float eta2;
    int row0, kp, km;
    int rk[NNP_GMAX];
    float tpz[NNP_GMAX];
    __device__ void none()
    {
        eta2 = 0.0f;
        row0 = 0;
        kp = km = -1;
#pragma unroll
        for (int q = 0; q < NNP_GMAX; ++q) {
            rk[q] = -1;
            tpz[q] = 0.0f;
        }
    }
};

// Descriptor of group g from LDS in 16-byte reads; every word is uniform.
__device__ __forceinline__ void nnp_group(const int32_t *gd, int g, NnpGroup &G)
{
    int dw[NNP_GD];
#pragma unroll
    for (int u = 0; u < NNP_GD / 4; ++u) {
        const int4 w4 = reinterpret_cast<const int4 *>(gd + NNP_GD * g)[u];
        dw[4 * u] = w4.x;
        dw[4 * u + 1] = w4.y;
        dw[4 * u + 2] = w4.z;
        dw[4 * u + 3] = w4.w;
    }
    G.eta2 = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(dw[0]));
    G.row0 = __builtin_amdgcn_readfirstlane(dw[1]);
    G.kp = __builtin_amdgcn_readfirstlane(dw[2]);
    G.km = __builtin_amdgcn_readfirstlane(dw[3]);
#pragma unroll
    for (int q = 0; q < NNP_GMAX; ++q) {
        G.tpz[q] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(dw[4 + q]));
        G.rk[q] = __builtin_amdgcn_readfirstlane(dw[4 + NNP_GMAX + q]);
    }
}

// One visit (j, k) for one group: the terms centred on j into aj and, for
// own visits, those centred on i into ai (per slot, tpz applied at the store).
__device__ __forceinline__ void nnp_visit(const NnpGroup &G, float rsn, float rso, float fpn, float fpo, float cjn,
                                          float cjo, float cin, float cio, bool own, float *aj, float *ai)
{
    const float en = __builtin_amdgcn_exp2f(-G.eta2 * rsn) * fpn;
    const float eo = -(__builtin_amdgcn_exp2f(-G.eta2 * rso) * fpo);
    nnp_chain<0>(G.kp, 1.0f, cjn, cjo, en, eo, aj);
    nnp_chain<NNP_KSLOT>(G.km, -1.0f, cjn, cjo, en, eo, aj);
    if (own) {
        nnp_chain<0>(G.kp, 1.0f, cin, cio, en, eo, ai);
        nnp_chain<NNP_KSLOT>(G.km, -1.0f, cin, cio, en, eo, ai);
    }
}

// Stores of one group: the 2^lp lanes of a target (an aligned block)
// combine their parts by quad exchanges and the half-row mirror, the first
// writes the target's entry; the wave's sum of the terms centred on i goes
// to the moved atom's column (added to the earlier chunks' sum when the
// targets come in several chunks of 64).  All lanes call.
template <int NP>
__device__ __forceinline__ void nnp_store(const NnpGroup &G, int lp, bool writer, int j, int i, int lane, float *Dp,
                                          float *aj, float *ai, bool first)
{
#pragma unroll
    for (int q = 0; q < NNP_GMAX; ++q) {
        if (G.rk[q] >= 0) {
            float v = aj[q];
            const float t1 = dpp_f<0xB1>(v);
            v = lp >= 1 ? v + t1 : v;
            const float t2 = dpp_f<0x4E>(v);
            v = lp >= 2 ? v + t2 : v;
            const float t3 = dpp_f<0x141>(v);
            v = lp >= 3 ? v + t3 : v;
            float *Dr = Dp + (G.row0 + G.rk[q]) * NP;
            if (writer) Dr[j] = v * G.tpz[q];
            const float vi = wave_sum_last(ai[q]);
            if (lane == 63) Dr[i] = first ? vi * G.tpz[q] : Dr[i] + vi * G.tpz[q];
        }
    }
}

// Full feature matrix from positions into Gout (total_symm_calc,
// SymmFunc.jl:494-510; radial :259-278, angular :280-312).  All threads call.
template <int NP>
__device__ void nnp_full_build(const NnpTables &T, const NnpLds &L, float *Gout, int N)
{
    const int tid = threadIdx.x;
    for (int e = tid; e < T.nf * NP; e += NNP_NT) Gout[e] = 0.0f;
    __syncthreads();
    // radial: one (function, atom) per item
    for (int e = tid; e < T.n_rad_active * N; e += NNP_NT) {
        const int q = e / N, a = e % N;
        const float eta = L.rad[4 * q], norm = L.rad[4 * q + 1];
        const int row = (int)L.rad[4 * q + 2];
        float s = 0.0f;
        for (int b = 0; b < N; ++b) {
            if (b == a) continue;
            const float dx = L.X[a] - L.X[b], dy = L.X[NP + a] - L.X[NP + b], dz = L.X[2 * NP + a] - L.X[2 * NP + b];
            const float r2 = dx * dx + dy * dy + dz * dz;
            const float fc = nnp_cutoff(sqrtf(r2), T.r_cut);
            if (fc != 0.0f && fc != 1.0f) s += fc * expf(-eta * r2) * norm;
        }
        Gout[row * NP + a] = s;
    }
    // angular: one triangle a < b < c per item, all functions
    const int ntri = N * (N - 1) * (N - 2) / 6;
    for (int e = tid; e < ntri; e += NNP_NT) {
        // decode e -> (a, b, c)
        int a = 0, rem = e;
        while (rem >= (N - 1 - a) * (N - 2 - a) / 2) {
            rem -= (N - 1 - a) * (N - 2 - a) / 2;
            ++a;
        }
        int b = a + 1;
        while (rem >= N - 1 - b) {
            rem -= N - 1 - b;
            ++b;
        }
        const int c = b + 1 + rem;
        const float pa[3] = {L.X[a], L.X[NP + a], L.X[2 * NP + a]};
        const float pb[3] = {L.X[b], L.X[NP + b], L.X[2 * NP + b]};
        const float pc[3] = {L.X[c], L.X[NP + c], L.X[2 * NP + c]};
        auto d2 = [](const float *u, const float *v) {
            const float x = u[0] - v[0], y = u[1] - v[1], z = u[2] - v[2];
            return x * x + y * y + z * z;
        };
        const float rab = d2(pa, pb), rac = d2(pa, pc), rbc = d2(pb, pc);
        const float dab = sqrtf(rab), dac = sqrtf(rac), dbc = sqrtf(rbc);
        const float fp = nnp_cutoff(dab, T.r_cut) * nnp_cutoff(dac, T.r_cut) * nnp_cutoff(dbc, T.r_cut);
        if (fp == 0.0f) continue;
        float ta, tb, tc;
        nnp_angles(pa, pb, pc, dab, dac, dbc, ta, tb, tc);
        const float rs = rab + rac + rbc;
        for (int g = 0; g < T.n_groups; ++g) {
            const float ex = expf(-T.grp_eta[g] * rs) * fp;
            const int f0 = T.grp[2 * g], cnt = T.grp[2 * g + 1];
            for (int s = 0; s < cnt; ++s) {
                const int f = f0 + s;
                if (!T.abuild[f]) continue;
                const float lam = T.afn[4 * f], zc = T.afn[4 * f + 1], tpz = T.afn[4 * f + 2];
                const int row = (int)T.afn[4 * f + 3];
                float pa_, pb_, pc_;
                nnp_theta3(lam, zc, ta, tb, tc, pa_, pb_, pc_);
                atomicAdd(&Gout[row * NP + a], ex * pa_ * tpz);
                atomicAdd(&Gout[row * NP + b], ex * pb_ * tpz);
                atomicAdd(&Gout[row * NP + c], ex * pc_ * tpz);
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < T.nf * N; e += NNP_NT) {
        const int q = e / N, a = e % N;
        Gout[q * NP + a] += T.row_off[q];
    }
    __syncthreads();
}

typedef float nnp_f32x16 __attribute__((ext_vector_type(16)));

// tanh x = 1 - 2 / (1 + e^{2x}), e^{2x} = 2^{2 log2(e) x}: absolute error below
// about 4e-7 over the whole range (the limits +-1 come out exactly)
__device__ __forceinline__ float nnp_tanh(float x)
{
    const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);
    return __builtin_fmaf(-2.0f, __builtin_amdgcn_rcpf(1.0f + e), 1.0f);
}

// Per-atom network energies of the features Gin into out[N]
// (n_features -> tanh h1 -> tanh h2 -> linear 1, ForwardPass.jl:23-39).
// The two hidden layers are NP x 32 products on the matrix cores in binary32
// (v_mfma_f32_32x32x2_f32; wave w takes atoms 32w to 32w + 31, two
// accumulators over alternating k pairs); lane layout: A[row l%32][k l/32],
// B[k l/32][col l%32], output value v of lane l at row 8(v/4) + 4(l/32) + v%4,
// column l%32.  The output layer is one dot product per atom; the FP64 total
// is a butterfly over each wave of 64 atoms, then the waves' sums in order.
// All threads call.
template <int NP>
__device__ void nnp_forward(const NnpTables &T, const NnpLds &L, const float *Gin, int N, float *out,
                            double *total = nullptr)
{
    constexpr int NW = NP / 64;
    __shared__ double s_fpart[NNP_NP_WIDE / 64];
    const int tid = threadIdx.x, l = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lr = l & 31, lh = l >> 5;
    const int nf = T.nf, h1 = T.h1, h2 = T.h2;
    if (w < NP / 32) {
        const int r0 = 32 * w;
        const int c1 = lr < h1 ? lr : 0;
        const float m1 = lr < h1 ? 1.0f : 0.0f;
        nnp_f32x16 za = {}, zb = {};
        for (int kk = 0; kk < nf; kk += 4) {
            const int ka = kk + lh, kb = kk + 2 + lh;
            const int kac = ka < nf ? ka : nf - 1, kbc = kb < nf ? kb : nf - 1;
            const float ma = ka < nf ? 1.0f : 0.0f, mb = kb < nf ? 1.0f : 0.0f;
            const float xa = Gin[kac * NP + r0 + lr] * ma, xb = Gin[kbc * NP + r0 + lr] * mb;
            const float wa = L.W1[kac * h1 + c1] * m1, wb = L.W1[kbc * h1 + c1] * m1;
            za = __builtin_amdgcn_mfma_f32_32x32x2f32(xa, wa, za, 0, 0, 0);
            zb = __builtin_amdgcn_mfma_f32_32x32x2f32(xb, wb, zb, 0, 0, 0);
        }
        if (lr < h1) {
            const float bias = L.B[lr];
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = r0 + 8 * (v >> 2) + 4 * lh + (v & 3);
                L.H1[row * NNP_HS + lr] = nnp_tanh(za[v] + zb[v] + bias);
            }
        }
    }
    __syncthreads();
    if (w < NP / 32) {
        const int r0 = 32 * w;
        const int c2 = lr < h2 ? lr : 0;
        const float m2 = lr < h2 ? 1.0f : 0.0f;
        nnp_f32x16 za = {}, zb = {};
        for (int kk = 0; kk < h1; kk += 4) {
            const int ka = kk + lh, kb = kk + 2 + lh;
            const int kac = ka < h1 ? ka : h1 - 1, kbc = kb < h1 ? kb : h1 - 1;
            const float ma = ka < h1 ? 1.0f : 0.0f, mb = kb < h1 ? 1.0f : 0.0f;
            const float xa = L.H1[(r0 + lr) * NNP_HS + kac] * ma, xb = L.H1[(r0 + lr) * NNP_HS + kbc] * mb;
            const float wa = L.W2[kac * h2 + c2] * m2, wb = L.W2[kbc * h2 + c2] * m2;
            za = __builtin_amdgcn_mfma_f32_32x32x2f32(xa, wa, za, 0, 0, 0);
            zb = __builtin_amdgcn_mfma_f32_32x32x2f32(xb, wb, zb, 0, 0, 0);
        }
        if (lr < h2) {
            // second hidden layer times the output weights
            const float bias = L.B[NNP_MAXH + lr], w3 = L.B[2 * NNP_MAXH + lr];
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = r0 + 8 * (v >> 2) + 4 * lh + (v & 3);
                L.H2[row * NNP_HS + lr] = nnp_tanh(za[v] + zb[v] + bias) * w3;
            }
        }
    }
    __syncthreads();
    if (w < NW) {
        const int at = 64 * w + l;
        float y = 0.0f;
        if (at < N) {
            for (int k = 0; k < h2; ++k) y += L.H2[at * NNP_HS + k];
            y += T.b3;
            out[at] = y;
        }
        // FP64 total of the atomic energies (a butterfly over the wave)
        if (total != nullptr) {
            double t = at < N ? (double)y : 0.0;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
            if (NW == 1 && l == 0) *total = t;
            if (NW > 1 && l == 0) s_fpart[w] = t;
        }
    }
    if (NW > 1 && total != nullptr) {
        __syncthreads();
        if (tid == 0) {
            double t = s_fpart[0];
            for (int v = 1; v < NW; ++v) t += s_fpart[v];
            *total = t;
        }
    }
    __syncthreads();
}

template <int NP>
__device__ __forceinline__ void nnp_load(const SweepArgs &a, const NnpLds &L, int slot)
{
    const NnpTables &T = a.nnp;
    const float *gX = reinterpret_cast<const float *>(a.coords) + (size_t)slot * 3 * NP;
    const int tid = threadIdx.x;
    for (int e = tid; e < 3 * NP; e += NNP_NT) L.X[e] = gX[e];
    for (int e = tid; e < T.nf * T.h1; e += NNP_NT) L.W1[e] = T.w1[e];
    for (int e = tid; e < 4 * T.n_rad_active; e += NNP_NT) L.rad[e] = T.rad[e];
    for (int e = tid; e < NNP_GD * T.n_groups; e += NNP_NT) L.gd[e] = T.gdesc[e];
    for (int e = tid; e < (NNP_NT / 64) * (NNP_WG + 1); e += NNP_NT) L.wg[e] = T.wgroups[e];
    for (int e = tid; e < T.h1 * T.h2; e += NNP_NT) L.W2[e] = T.w2[e];
    for (int e = tid; e < NNP_MAXH; e += NNP_NT) {
        L.B[e] = e < T.h1 ? T.b1[e] : 0.0f;
        L.B[NNP_MAXH + e] = e < T.h2 ? T.b2[e] : 0.0f;
        L.B[2 * NNP_MAXH + e] = e < T.h2 ? T.w3[e] : 0.0f;
    }
}

#define NNP_BIND_TABLES(L) \
    L.rad = s_rad;         \
    L.W2 = s_w2;           \
    L.B = s_bias;          \
    L.gd = s_gdesc;        \
    L.wg = s_wgrp

// set_variables + initialise_energy (src/EnergyEvaluation.jl:1812-1822,
// :1735-1747): features, per-atom energies and the FP64 total of each slot.
template <int NP>
__global__ void __launch_bounds__(NNP_NT) nnp_init_kernel(SweepArgs a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const NnpTables &T = a.nnp;
    const int slot = blockIdx.x;
    if (a.init_run >= 0 && a.slot_run[slot] != a.init_run) return;  // whole workgroup
    NNP_TABLES_LDS;
    NnpLds L = nnp_lds<NP>(reinterpret_cast<float *>(lds_raw), T.nf, T.h1);
    NNP_BIND_TABLES(L);
    const int N = a.runs[a.slot_run[slot]].n_atoms;
    nnp_load<NP>(a, L, slot);
    __syncthreads();
    nnp_full_build<NP>(T, L, L.G, N);
    __shared__ double s_tot;
    nnp_forward<NP>(T, L, L.G, N, L.eat, &s_tot);
    float *gG = a.feat + (size_t)slot * T.nf * NP;
    for (int e = threadIdx.x; e < T.nf * NP; e += NNP_NT) gG[e] = L.G[e];
    for (int e = threadIdx.x; e < NP; e += NNP_NT) a.eatom[(size_t)slot * NP + e] = e < N ? L.eat[e] : 0.0f;
    if (threadIdx.x == 0) a.slot_en[slot] = s_tot;
}

// One Monte-Carlo cycle (N moves) of one configuration per workgroup
// (mc_step!, src/MCRun.jl:169-182, with energy_update!(RuNNerPotential)).
// DA: delayed acceptance with the learned surrogate (one workgroup per CU by
// LDS, so the register budget is relaxed for that variant).  NP: atoms of a
// slot, 64 (two workgroups per CU) or 128 (one; the near atoms of a move are
// then up to 127, their partner masks two words, and the angular targets are
// taken in chunks of 64 lanes).
template <bool DA, int NP>
__global__ void __launch_bounds__(NNP_NT) __attribute__((amdgpu_waves_per_eu((DA || NP > NNP_NP) ? 2 : 4)))
nnp_sweep_kernel(SweepArgs a)
{
    static_assert(NP == NNP_NP || (NP == NNP_NP_WIDE && !DA), "slots of 64 atoms, or 128 without the surrogate");
    constexpr int NW = NP / 64;  // 64-bit words of an atom mask
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ int s_flag, s_err, s_nnear;
    __shared__ uint64_t s_wmask[NW];
    __shared__ uint8_t s_near[NP];
    __shared__ double s_en;
    __shared__ float s_trial[3], s_old[3];
    const NnpTables &T = a.nnp;
    const int slot = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    NNP_TABLES_LDS;
    NnpLds L = nnp_lds<NP>(reinterpret_cast<float *>(lds_raw), T.nf, T.h1);
    NNP_BIND_TABLES(L);
    const RunInfo &ri = a.runs[a.slot_run[slot]];
    const int N = ri.n_atoms;
    const uint32_t nsteps = (uint32_t)N;
    const int t = a.slot_temp[slot];
    const int Tg = ri.temp_base + t;
    const uint32_t sweep = (uint32_t)a.sweep0;
    const double beta = a.t_beta[Tg];
    const float delta_r = (float)a.t_displ[2 * Tg];
    const float radius2 = (float)ri.radius2;

    nnp_load<NP>(a, L, slot);
    const float *gG = a.feat + (size_t)slot * T.nf * NP;
    for (int e = tid; e < T.nf * NP; e += NNP_NT) L.G[e] = gG[e];
    for (int e = tid; e < NP; e += NNP_NT) L.eat[e] = a.eatom[(size_t)slot * NP + e];
    double E = a.slot_en[slot];
    if (tid == 0) s_err = 0;
    // delayed acceptance: surrogate descriptors and atom energies in LDS
    // (rebuilt from the positions at every launch)
    __shared__ float s_acci[SUR_K];
    __shared__ double s_ehat;
    constexpr bool da = DA;
    float *Sd = nullptr, *Sn = nullptr, *sat = nullptr, *satn = nullptr, *sgeo = nullptr;
    SurNet net{};
    const SurGeom sg = sur_geom(a.sur_rcut);
    double Ehat = 0.0;
    long long n_prop = 0, n_acc1 = 0;
    if (da) {
        float *base = reinterpret_cast<float *>(lds_raw) + nnp_lds_floats(T.nf, T.h1);
        Sd = base;
        Sn = Sd + NP * SUR_SS;
        net.W1 = reinterpret_cast<short *>(Sn + NP * SUR_SS);
        net.W2 = net.W1 + SUR_K * SUR_H;
        net.Hb = net.W2 + SUR_H * SUR_H;
        sat = reinterpret_cast<float *>(net.Hb + NP * SUR_HS);
        satn = sat + NP;
        net.b1 = satn + NP;
        net.b2 = net.b1 + SUR_H;
        net.w3 = net.b2 + SUR_H;
        sgeo = net.w3 + SUR_H;  // per atom: r_new, r_old, f_c(r_new), f_c(r_old)
        sur_load_net(net, a.sur_w, NNP_NT);
        sur_descriptors(sg, L.X, N, Sd, SUR_SS, NNP_NT);
        for (int e = tid; e < NP * SUR_SS; e += NNP_NT) Sn[e] = 0.0f;
        __syncthreads();
        sur_atom_energies(net, Sd, N, L.H1, sat, &s_ehat);
        Ehat = s_ehat;
    }
    __syncthreads();
    int nacc = 0;
    // current features and the scratch that becomes the trial features; an
    // accepted move swaps the two (swap_vars!(NNPVariables), src/swap_config.jl:213-221)
    float *Gp = L.G, *Dp = L.D;
    float4 *nb_a = reinterpret_cast<float4 *>((reinterpret_cast<uintptr_t>(L.H2) + 15) & ~(uintptr_t)15);
    float4 *nb_b = nb_a + NP;
    float *nb_fo = reinterpret_cast<float *>(nb_b + NP);

    for (uint32_t m = 0; m < nsteps; ++m) {
        // draws and trial position (every thread the same), src/MCMoves.jl:115-135
        MoveDraw<float> d0 = move_draw<float>(a.seed, ri.run_key, (uint32_t)t, sweep, m, 0u, nsteps);
        const int i = (int)d0.index;
        float xi = L.X[i], yi = L.X[NP + i], zi = L.X[2 * NP + i];
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
                        double c[3] = {0.0, 0.0, 0.0};
                        for (int b = 0; b < N; ++b)
                            for (int k = 0; k < 3; ++k) c[k] += (double)L.X[k * NP + b];
                        for (int k = 0; k < 3; ++k) c[k] /= (double)N;
                        for (int b = 0; b < N; ++b)
                            for (int k = 0; k < 3; ++k)
                                L.X[k * NP + b] = (float)((double)L.X[k * NP + b] - c[k]);
                    }
                    __syncthreads();
                } else {
                    attempt += 1;
                    MoveDraw<float> dd = move_draw<float>(a.seed, ri.run_key, (uint32_t)t, sweep, m, attempt, nsteps);
                    xi = L.X[i];
                    yi = L.X[NP + i];
                    zi = L.X[2 * NP + i];
                    make_trial<float, BC_SPHERE, false>(xi, yi, zi, dd, delta_r, 0.0f, 0.0f, tx, ty, tz);
                    if (count == 100) {
                        if (tid == 0) s_err = 1;
                        break;
                    }
                }
            }
            xi = L.X[i];
            yi = L.X[NP + i];
            zi = L.X[2 * NP + i];
        }
        const float pt[3] = {tx, ty, tz}, po[3] = {xi, yi, zi};
        // distances and cutoffs to the new and old position; hard-core test
        // (one lane per atom) and the list of the atoms within the cut-off of
        // either position, the only partners of angular terms, in atom order
        if (tid == 0) s_flag = 0;
        __syncthreads();
        if (tid < NP) {
            const int j = tid;
            bool near = false;
            if (j < N) {
                const float px = L.X[j], py = L.X[NP + j], pz = L.X[2 * NP + j];
                float dx = tx - px, dy = ty - py, dz = tz - pz;
                const float rn = j == i ? 0.0f : dx * dx + dy * dy + dz * dz;
                dx = xi - px;
                dy = yi - py;
                dz = zi - pz;
                const float ro = j == i ? 0.0f : dx * dx + dy * dy + dz * dz;
                const float fnj = nnp_cutoff(sqrtf(rn), T.r_cut), foj = nnp_cutoff(sqrtf(ro), T.r_cut);
                L.rn2[j] = rn;
                L.ro2[j] = ro;
                L.fn[j] = fnj;
                L.fo[j] = foj;
                if (j != i && (double)rn < T.hardcore2) s_flag = 1;
                near = j != i && (fnj != 0.0f || foj != 0.0f);
            }
            const uint64_t b = __ballot(near);
            if constexpr (NW == 1) {
                if (near) {
                    // packed per near atom for the angular visits (in the H2 scratch):
                    // {x, y, z, f_new}, {r2_new, r2_old, r_new, r_old}, f_old
                    const int q = mbcnt64(b);
                    s_near[q] = (uint8_t)j;
                    nb_a[q] = make_float4(L.X[j], L.X[NP + j], L.X[2 * NP + j], L.fn[j]);
                    nb_b[q] = make_float4(L.rn2[j], L.ro2[j], sqrtf(L.rn2[j]), sqrtf(L.ro2[j]));
                    nb_fo[q] = L.fo[j];
                }
                if (lane == 0) s_nnear = __popcll(b);
            } else {
                if (lane == 0) s_wmask[tid >> 6] = b;
            }
        }
        if constexpr (NW > 1) {
            // the near atoms of lower waves come first: list positions from
            // the waves' masks once every wave has written its own
            __syncthreads();
            if (tid < NP) {
                const int j = tid;
                const uint64_t b = s_wmask[tid >> 6];
                if ((b >> lane) & 1ull) {
                    int q = mbcnt64(b);
                    for (int v = 0; v < (tid >> 6); ++v) q += __popcll(s_wmask[v]);
                    s_near[q] = (uint8_t)j;
                    nb_a[q] = make_float4(L.X[j], L.X[NP + j], L.X[2 * NP + j], L.fn[j]);
                    nb_b[q] = make_float4(L.rn2[j], L.ro2[j], sqrtf(L.rn2[j]), sqrtf(L.ro2[j]));
                    nb_fo[q] = L.fo[j];
                }
            }
            if (tid == 0) {
                int t = 0;
                for (int v = 0; v < NW; ++v) t += __popcll(s_wmask[v]);
                s_nnear = t;
            }
        }
        for (int e = tid; e < T.nf * NP; e += NNP_NT) Dp[e] = 0.0f;
        if (da && tid < SUR_K) s_acci[tid] = 0.0f;
        __syncthreads();
        // delayed acceptance, stage 1: Metropolis on the surrogate energy change
        bool stage1 = true;
        double dEh = 0.0, Ehn = 0.0;
        if (da) {
            // distances and cutoffs of every atom to the two positions, once
            if (tid < N) {
                const float rn = sqrtf(L.rn2[tid]), ro = sqrtf(L.ro2[tid]);
                sgeo[tid] = rn;
                sgeo[NP + tid] = ro;
                sgeo[2 * NP + tid] = tid == i ? 0.0f : sur_cutoff(rn, sg.r_cut);
                sgeo[3 * NP + tid] = tid == i ? 0.0f : sur_cutoff(ro, sg.r_cut);
            }
            __syncthreads();
            float accq = 0.0f;
            const int kq = tid & (SUR_K - 1);  // NNP_NT is a multiple of SUR_K: fixed k per thread
            const float mu = sur_mu(sg, kq);
            for (int j = tid / SUR_K; j < N; j += NNP_NT / SUR_K) {
                if (j == i) continue;
                const float fcn = sgeo[2 * NP + j], fco = sgeo[3 * NP + j];
                const float pn = sur_phi_mu(sg, sgeo[j], fcn, mu);
                const float pq = sur_phi_mu(sg, sgeo[NP + j], fco, mu);
                Sn[j * SUR_SS + kq] = __builtin_fmaf(SUR_SIN, pn - pq, Sd[j * SUR_SS + kq]);
                accq += pn;
            }
            atomicAdd(&s_acci[kq], accq);
            __syncthreads();
            if (tid < SUR_K) Sn[i * SUR_SS + tid] = SUR_SIN * s_acci[tid];
            __syncthreads();
            sur_atom_energies(net, Sn, N, L.H1, satn, &s_ehat);
            Ehn = s_ehat;
            dEh = Ehn - Ehat;
            stage1 = (dEh <= 0.0) || (det_exp(-dEh * beta) >= (double)d0.umet);
            n_prop += 1;
            n_acc1 += stage1 ? 1 : 0;
        }
        double new_en = 100.0;  // the hard-core sentinel, src/EnergyEvaluation.jl:1604-1606
        if (da && !stage1) new_en = __builtin_nan("");
        if (stage1 && !s_flag) {
            if constexpr (DA) {
                // the surrogate's layer scratch covered the packed near-atom data
                if (tid < s_nnear) {
                    const int j = s_near[tid];
                    nb_a[tid] = make_float4(L.X[j], L.X[NP + j], L.X[2 * NP + j], L.fn[j]);
                    nb_b[tid] = make_float4(L.rn2[j], L.ro2[j], sqrtf(L.rn2[j]), sqrtf(L.ro2[j]));
                    nb_fo[tid] = L.fo[j];
                }
                __syncthreads();
            }
            // radial changes, DeltaMatrix.jl:794-822: one wave per function,
            // one lane per atom (NW atoms per lane in the wide layout)
            for (int q = tid >> 6; q < T.n_rad_active; q += NNP_NT / 64) {
                const float eta = L.rad[4 * q], norm = L.rad[4 * q + 1];
                const int row = (int)L.rad[4 * q + 2];
                float dsum = 0.0f;
#pragma unroll
                for (int v = 0; v < NW; ++v) {
                    const int at = 64 * v + lane;
                    float d = 0.0f;
                    if (at < N && at != i)
                        d = expf(-eta * L.rn2[at]) * (L.fn[at] * norm) + expf(-eta * L.ro2[at]) * (-L.fo[at] * norm);
                    if (at < N && at != i) Dp[row * NP + at] = d;
                    dsum += d;
                }
                const float s = wave_sum(dsum);
                if (lane == 0) Dp[row * NP + i] = s;
            }
            // angular changes, DeltaMatrix.jl:827-872.  Every partner pair
            // (j, k) of the moved atom i whose triangle lies inside the cut-off
            // (old or new position) is visited from j and from k; a visit adds
            // the terms centred on its own atom and, from the atom of smaller
            // index, the terms centred on i.  Each wave owns a few eta groups
            // (their feature rows), each lane one target atom j (or a half or
            // quarter of its partners when few atoms are near), so every
            // feature entry is written by one lane with a plain store: no LDS
            // atomics (float LDS adds run far below the store rate).
            const int n_near = s_nnear;
            // partner masks [near atom][NW words] (scratch until the network pass)
            uint64_t *s_mask = reinterpret_cast<uint64_t *>(L.H1);
            const float rc2 = T.r_cut * T.r_cut * 1.0001f;  // loose: the cut-off function decides below
            for (int a = tid >> 6; a < n_near; a += NNP_NT / 64) {
                const float4 pa = nb_a[a];
                const float foa = nb_fo[a];
#pragma unroll
                for (int v = 0; v < NW; ++v) {
                    const int b = 64 * v + lane;
                    bool act = false;
                    if (b < n_near && b != a) {
                        const float4 pb = nb_a[b];
                        if (pa.w * pb.w != 0.0f || foa * nb_fo[b] != 0.0f) {
                            const float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
                            act = dx * dx + dy * dy + dz * dz < rc2;
                        }
                    }
                    const uint64_t m = __ballot(act);
                    if (lane == 0) s_mask[a * NW + v] = m;
                }
            }
            __syncthreads();
            // Targets in chunks of 64 near atoms; per chunk, lanes per target:
            // target a gets P_a = 1, 2, 4 or 8 lanes, the fewest that bring its
            // share of visits down to the load T, T the smallest load for which
            // the shares fit 64 lanes (bisection, the lane count summed by
            // ballots).  Blocks of P lanes are laid out by P (8, 4, 2, 1), so
            // each is aligned for the quad and half-row exchanges that combine
            // them.  Every wave builds the same tables in its own LDS slice (no
            // barrier needed within a wave).
            uint16_t *s_lt = reinterpret_cast<uint16_t *>(L.H1 + 2 * NW * NP) + NP * (tid >> 6);
            const int nch = (n_near + 63) >> 6;
            int used_c[NW];
#pragma unroll
            for (int c = 0; c < NW; ++c) {
                used_c[c] = 0;
                if (c >= nch) continue;
                const int nc = n_near - 64 * c < 64 ? n_near - 64 * c : 64;
                int La = 0;
                if (lane < nc) {
#pragma unroll
                    for (int v = 0; v < NW; ++v) La += __popcll(s_mask[(64 * c + lane) * NW + v]);
                }
                auto lanes_for = [&](int load) {
                    if (lane >= nc) return 0;
                    const int q = (La + load - 1) / load;
                    return q <= 1 ? 1 : (q <= 2 ? 2 : (q <= 4 ? 4 : 8));
                };
                auto lanes_used = [&](int pa) {
                    int t = 0;
#pragma unroll
                    for (int cc = 0; cc < 4; ++cc) t += __popcll(__ballot(pa == (1 << cc))) << cc;
                    return t;
                };
                int lo = 1, hi = 64 * NW;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (lanes_used(lanes_for(mid)) <= 64) hi = mid;
                    else lo = mid + 1;
                }
                const int Pa = lanes_for(lo);
                int used = 0;
#pragma unroll
                for (int cc = 3; cc >= 0; --cc) {
                    const bool in = Pa == (1 << cc);
                    const uint64_t b = __ballot(in);
                    if (in)
                        for (int q = 0; q < Pa; ++q)
                            s_lt[64 * c + used + (mbcnt64(b) << cc) + q] = (uint16_t)(lane | (q << 6) | (cc << 9));
                    used += __popcll(b) << cc;
                }
                used_c[c] = used;
            }
            // this lane's target in chunk c: lanes, part, atom, partner masks
            // (and those of larger list index, the visits that also carry the
            // terms centred on i), packed data
            struct Target {
                int lp, P, part, j;
                bool tact;
                uint64_t m0, m1, after0, after1;
                float4 ja, jb;
                float jfo;
            };
            auto target = [&](int c) {
                Target g;
                const int ent = lane < used_c[c] ? (int)s_lt[64 * c + lane] : 0;
                g.lp = ent >> 9;
                g.P = 1 << g.lp;
                const int ta = 64 * c + (ent & 63);
                g.part = (ent >> 6) & 7;
                g.tact = lane < used_c[c];
                g.j = g.tact ? (int)s_near[ta] : 0;
                g.m0 = g.tact ? s_mask[ta * NW] : 0ull;
                g.m1 = 0ull;
                g.after0 = ta < 63 ? ~((2ull << ta) - 1ull) : 0ull;
                g.after1 = 0ull;
                if constexpr (NW > 1) {
                    g.m1 = g.tact ? s_mask[ta * NW + 1] : 0ull;
                    const int sh = ta - 64;
                    g.after1 = sh < 0 ? ~0ull : (sh < 63 ? ~((2ull << sh) - 1ull) : 0ull);
                }
                g.ja = nb_a[ta];
                g.jb = nb_b[ta];
                g.jfo = nb_fo[ta];
                return g;
            };
            Target t1{};
            if constexpr (NW == 1) t1 = target(0);  // one chunk: read once for all group passes
            const int w = tid >> 6;
            const int ngw = L.wg[w * (NNP_WG + 1)];
            for (int gi = 0; gi < ngw; gi += 2) {
                // up to two groups per pass over the visits
                const int ga = L.wg[w * (NNP_WG + 1) + 1 + gi];
                const int gb = gi + 1 < ngw ? L.wg[w * (NNP_WG + 1) + 2 + gi] : -1;
                NnpGroup G0, G1;
                nnp_group(L.gd, ga, G0);
                if (gb >= 0) nnp_group(L.gd, gb, G1);
                else G1.none();
#pragma unroll
                for (int c = 0; c < NW; ++c) {
                    if (c >= nch) continue;
                    const Target tg = NW == 1 ? t1 : target(c);
                    const int lp = tg.lp, P = tg.P, part = tg.part, j = tg.j;
                    const bool tact = tg.tact;
                    uint64_t m0 = tg.m0, m1 = tg.m1;
                    const uint64_t after0 = tg.after0, after1 = tg.after1;
                    const float4 ja = tg.ja, jb = tg.jb;
                    const float jfo = tg.jfo;
                    const float pj3[3] = {ja.x, ja.y, ja.z};
                    const float dijn = jb.z, dijo = jb.w;
                    float aj0[NNP_GMAX], ai0[NNP_GMAX], aj1[NNP_GMAX], ai1[NNP_GMAX];
#pragma unroll
                    for (int q = 0; q < NNP_GMAX; ++q) aj0[q] = ai0[q] = aj1[q] = ai1[q] = 0.0f;
                    auto pop = [&]() {  // drop the lowest partner
                        if constexpr (NW == 1) {
                            m0 &= m0 - 1ull;
                        } else {
                            if (m0 != 0ull) m0 &= m0 - 1ull;
                            else m1 &= m1 - 1ull;
                        }
                    };
                    for (int t = 0; t < part; ++t) pop();  // this lane's first partner
                    while (NW == 1 ? m0 != 0ull : (m0 | m1) != 0ull) {
                        const int bk = (NW == 1 || m0 != 0ull) ? __builtin_ctzll(m0) : 64 + __builtin_ctzll(m1);
                        for (int t = 0; t < P; ++t) pop();  // next partner of this lane
                        const bool own = (NW == 1 || bk < 64) ? ((after0 >> bk) & 1ull) != 0ull
                                                              : ((after1 >> (bk - 64)) & 1ull) != 0ull;
                        const float4 ka = nb_a[bk], kb = nb_b[bk];
                        const float kfo = nb_fo[bk];
                        const float pk3[3] = {ka.x, ka.y, ka.z};
                        const float dx = pj3[0] - pk3[0], dy = pj3[1] - pk3[1], dz = pj3[2] - pk3[2];
                        const float rjk = dx * dx + dy * dy + dz * dz;
                        const float djk = sqrtf(rjk);
                        const float fjk = nnp_cutoff(djk, T.r_cut);
                        const float fpn = ja.w * ka.w * fjk, fpo = jfo * kfo * fjk;
                        if (fpn == 0.0f && fpo == 0.0f) continue;
                        const float rsn = jb.x + kb.x + rjk, rso = jb.y + kb.y + rjk;
                        // cosines at j (and at i), new and old
                        float cjn = 0.0f, cjo = 0.0f, cin = 0.0f, cio = 0.0f;
                        const float dikn = kb.z, diko = kb.w;
                        if (fpn != 0.0f) {
                            cjn = nnp_cos(pj3, pt, pk3, dijn, djk);
                            if (own) cin = nnp_cos(pt, pj3, pk3, dijn, dikn);
                        }
                        if (fpo != 0.0f) {
                            cjo = nnp_cos(pj3, po, pk3, dijo, djk);
                            if (own) cio = nnp_cos(po, pj3, pk3, dijo, diko);
                        }
                        nnp_visit(G0, rsn, rso, fpn, fpo, cjn, cjo, cin, cio, own, aj0, ai0);
                        if (gb >= 0) nnp_visit(G1, rsn, rso, fpn, fpo, cjn, cjo, cin, cio, own, aj1, ai1);
                    }
                    nnp_store<NP>(G0, lp, tact && part == 0, j, i, lane, Dp, aj0, ai0, c == 0);
                    if (gb >= 0) nnp_store<NP>(G1, lp, tact && part == 0, j, i, lane, Dp, aj1, ai1, c == 0);
                }
            }
            __syncthreads();
            for (int e = tid; e < T.nf * NP; e += NNP_NT) Dp[e] += Gp[e];
            __syncthreads();
            nnp_forward<NP>(T, L, Dp, N, L.eatn, &s_en);
            new_en = s_en;
        }
        // Metropolis, src/MCRun.jl:122-132 and src/Exchange.jl:53-57; with
        // delayed acceptance the second stage corrects the surrogate's error:
        // accept with min(1, exp(-beta ((E' - E) - (E_hat' - E_hat))))
        const double dd = new_en - E;
        bool accept;
        if (!da) {
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
            a.trace_de[ti] = dd;
            a.trace_acc[ti] = (int8_t)accept;
            a.trace_idx[ti] = i;
            if (da && a.trace_aux != nullptr) a.trace_aux[ti] = dEh;
        }
        if (accept) {
            // swap_atom_config! + swap_vars!(NNPVariables), src/swap_config.jl:40-51, :213-221
            float *tmp = Gp;
            Gp = Dp;
            Dp = tmp;
            for (int e = tid; e < N; e += NNP_NT) L.eat[e] = L.eatn[e];
            if (da) {
                float *t1 = Sd;
                Sd = Sn;
                Sn = t1;
                t1 = sat;
                sat = satn;
                satn = t1;
                Ehat = Ehn;
            }
            if (tid == 0) {
                L.X[i] = tx;
                L.X[NP + i] = ty;
                L.X[2 * NP + i] = tz;
            }
            E = new_en;
            nacc += 1;
        }
        __syncthreads();
    }

    // periodic full rebuild bounds the drift of the incremental features
    if (a.recompute_every > 0 && ((long long)sweep + 1) % a.recompute_every == 0) {
        nnp_full_build<NP>(T, L, Gp, N);
        nnp_forward<NP>(T, L, Gp, N, L.eat, &s_en);
        const double dr = __builtin_fabs(E - s_en);
        if (tid == 0 && dr > a.slot_drift[slot]) a.slot_drift[slot] = dr;
        E = s_en;
    }

    float *gX = reinterpret_cast<float *>(a.coords) + (size_t)slot * 3 * NP;
    float *gGw = a.feat + (size_t)slot * T.nf * NP;
    for (int e = tid; e < 3 * NP; e += NNP_NT) gX[e] = L.X[e];
    for (int e = tid; e < T.nf * NP; e += NNP_NT) gGw[e] = Gp[e];
    for (int e = tid; e < NP; e += NNP_NT) a.eatom[(size_t)slot * NP + e] = L.eat[e];
    if (tid == 0) {
        a.slot_en[slot] = E;
        a.t_count_atom[2 * Tg] += nacc;
        a.t_count_atom[2 * Tg + 1] += nacc;
        if (s_err) a.run_error[a.slot_run[slot]] = 1;
        if (da) {
            a.da_stats[3 * slot] += n_prop;
            a.da_stats[3 * slot + 1] += n_acc1;
            a.da_stats[3 * slot + 2] += nacc;
        }
    }
}

// Between cycles: replica exchange (src/Exchange.jl:103-155, src/MCRun.jl:203-205),
// step adaptation (src/Exchange.jl:198-263) and sampling (src/Sampling.jl:330-371),
// one thread per run.
__global__ void nnp_cycle_kernel(SweepArgs a, int n_runs)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_runs) return;
    const RunInfo &ri = a.runs[r];
    const int R = ri.n_temps, tb = ri.temp_base, N = ri.n_atoms;
    const uint32_t sweep = (uint32_t)a.sweep0;
    const long long pidx = a.cycle0 + 1;
    auto try_pair = [&](int n) {
        u32x4 X1 = draw(a.seed, ri.run_key, RNG_EXCH, 0u, 0u, sweep, 1u + (uint32_t)n);
        const double u = u53(X1.x, X1.y);
        const int sa = a.temp_slot[tb + n], sb = a.temp_slot[tb + n + 1];
        const double de = a.slot_en[sa] - a.slot_en[sb];
        const double db = a.t_beta[tb + n] - a.t_beta[tb + n + 1];
        double acc = det_exp(db * de);
        if (acc > 1.0) acc = 1.0;
        a.t_count_exc[2 * (tb + n)] += 1;
        a.t_count_exc[2 * (tb + n + 1)] += 1;
        if (acc > u) {
            a.t_count_exc[2 * (tb + n) + 1] += 1;
            a.t_count_exc[2 * (tb + n + 1) + 1] += 1;
            a.slot_temp[sa] = n + 1;
            a.slot_temp[sb] = n;
            a.temp_slot[tb + n] = sb;
            a.temp_slot[tb + n + 1] = sa;
        }
    };
    if (R > 1 && (sweep + 1u) % (uint32_t)a.exchange_every == 0u) {
        if (a.exchange_mode == EXC_EVEN_ODD) {
            for (int n = (int)(sweep & 1u); n + 1 < R; n += 2) try_pair(n);
        } else {
            u32x4 X0 = draw(a.seed, ri.run_key, RNG_EXCH, 0u, 0u, sweep, 0u);
            const double u = u53(X0.x, X0.y);
            if (a.exchange_mode == EXC_EVERY || u < 0.1) try_pair((int)below(X0.z, (uint32_t)(R - 1)));
        }
    }
    for (int t = 0; t < R; ++t) {
        const int T = tb + t;
        if (pidx % a.n_adjust == 0) {
            const double ar = (double)a.t_count_atom[2 * T + 1] / (double)((long long)a.n_adjust * N);
            if (ar < a.min_acc) a.t_displ[2 * T] *= 0.9;
            else if (ar > a.max_acc) a.t_displ[2 * T] *= 1.1;
            a.t_count_atom[2 * T + 1] = 0;
        }
        if (a.sampling && pidx % a.mc_sample == 0) {
            const double E = a.slot_en[a.temp_slot[T]];
            a.t_ham[2 * T] += E;
            a.t_ham[2 * T + 1] += E * E;
            const double hi = (E - ri.en_min) / ri.delta_en + 1.0;
            const int h = hi < 1.0 ? 1 : (hi > (double)a.n_bin ? a.n_bin + 2 : (int)__builtin_floor(hi) + 1);
            a.t_hist[(size_t)T * a.hist_stride + (h - 1)] += 1u;
        }
    }
}

}  // namespace ptmc
