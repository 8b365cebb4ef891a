// Sizes shared by the host engine and the neural-network kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace ptmc {

constexpr int NNP_NP = 64;     // atoms per configuration of the narrow slot layout
constexpr int NNP_NP_WIDE = 128;  // atoms per configuration of the wide layout (one workgroup per CU)
constexpr int NNP_NT = 512;    // threads per workgroup
constexpr int PSL_NT = 256;    // threads per workgroup of the pair slots (the surrogate network uses waves 0-3)
constexpr int NNP_MAXH = 32;   // hidden width (maximum)
constexpr int NNP_HS = NNP_MAXH + 1;  // padded hidden row stride
constexpr int NNP_KSLOT = 5;   // zeta = 2^k, k = 0..4 (zeta up to 16)
constexpr int NNP_GMAX = 2 * NNP_KSLOT;  // functions per eta group (one per lambda = +-1 and k)
// group descriptor words: eta log2(e), first row, max k for lambda = +1 and
// -1 (-1: none), then per slot (lambda, k) = (+1, 0..4), (-1, 0..4) the
// function's tpz and its row within the group (-1: no function)
constexpr int NNP_GD = 4 + 2 * NNP_GMAX;
constexpr int NNP_MAXG = 32;  // angular groups (maximum)
constexpr int NNP_WG = 4;     // angular groups per wave in the move update (maximum)
static_assert(NNP_GD % 4 == 0, "group descriptors are read in 16-byte words");
constexpr int NNP_MAXF = 128;  // symmetry functions (maximum)

// LDS layout in floats for slots of np atoms (64 or 128): X[3][np] |
// G[nf][np] | D[nf][np] | W1[nf][h1] | H1[np][HS] | H2[np][HS] | rn2, fn,
// ro2, fo, eat, eatn [np each]
__host__ __device__ inline size_t nnp_lds_floats(int nf, int h1, int np = NNP_NP)
{
    return (size_t)3 * np + 2 * (size_t)nf * np + (size_t)nf * h1 + 2 * (size_t)np * NNP_HS + 6 * (size_t)np;
}

}  // namespace ptmc

namespace ptmc {

// Taylor coefficient (-1)^k (pi/2)^(2k) / (2k)! of cos(pi t / 2) in u = t^2
constexpr double cutoff_cos_coef(int k)
{
    double c = 1.0;
    for (int q = 1; q <= 2 * k; ++q) c *= 1.5707963267948966 / q;
    return (k & 1) ? -c : c;
}

// Cosine cutoff 0.5 (cos(pi r / rc) + 1) for r < rc, else 0
// (MachineLearningPotential/Cutoff.jl:79-90): cos^2(pi t / 2), t = r / rc,
// the cosine by its Taylor series in u = t^2 through u^7 (remainder below
// 1e-8 on [0, 1])
__device__ __forceinline__ float cutoff_cos2(float r, float rc)
{
    if (!(r < rc)) return 0.0f;
    const float t = r * (1.0f / rc);
    const float u = t * t;
    float c = (float)cutoff_cos_coef(7);
    c = __builtin_fmaf(c, u, (float)cutoff_cos_coef(6));
    c = __builtin_fmaf(c, u, (float)cutoff_cos_coef(5));
    c = __builtin_fmaf(c, u, (float)cutoff_cos_coef(4));
    c = __builtin_fmaf(c, u, (float)cutoff_cos_coef(3));
    c = __builtin_fmaf(c, u, (float)cutoff_cos_coef(2));
    c = __builtin_fmaf(c, u, (float)cutoff_cos_coef(1));
    c = __builtin_fmaf(c, u, 1.0f);
    return c * c;
}

// Learned surrogate (delayed acceptance): 64 Gaussian radial descriptors per
// atom -> 64 -> 64 (SiLU) -> 1.
constexpr int SUR_K = 64;
constexpr int SUR_H = 64;
constexpr int SUR_P = SUR_K * SUR_H + SUR_H + SUR_H * SUR_H + SUR_H + SUR_H + 1;
constexpr float SUR_RMIN = 2.0f;
constexpr float SUR_SIN = 0.25f;

// padded LDS row strides (bank-conflict-free column access)
constexpr int SUR_SS = SUR_K + 1;  // descriptors, f32
constexpr int SUR_HS = SUR_H + 2;  // hidden activations, bf16
constexpr int SUR_YS = SUR_H + 1;  // output products, f32

// extra LDS floats of a delayed-acceptance sweep: descriptors S, S' [NP][SS]
// (f32), bf16 W1, W2 [K][H] and H [NP][HS], surrogate atom energies (current
// and trial), b1, b2, w3
__host__ __device__ inline size_t sur_lds_floats()
{
    return (size_t)2 * NNP_NP * SUR_SS + (size_t)SUR_K * SUR_H + (size_t)NNP_NP * SUR_HS / 2 + 2 * NNP_NP +
           3 * SUR_H + 4 * NNP_NP;
}

// Two-species NNVT kernels: threads per workgroup, chunks of the partner
// pairs of a fully rebuilt feature column, and the LDS layout in doubles:
// X[N][3] | XT[N][3] | R2[N][N] | FC[N][N] | RN[2][N] | FN[2][N] |
// D[N][ng] | eat[N] | eatn[N] | W[2][n_params] (both networks) | scratch
// (layer activations, or the partial sums of rebuilt columns) | ints:
// moved index[N], changed list[N]
constexpr int NNP2_NT = 512;
constexpr int NNP2_NCH = 20;
constexpr int NNP2_GMAX = 6;  // angular functions per eta group (larger groups are split)
__host__ __device__ inline size_t nnp2_scratch_doubles(int N, int nfn, int h1, int h2)
{
    const size_t act = (size_t)N * (h1 + h2), part = (size_t)2 * nfn * NNP2_NCH * 3;
    return act > part ? act : part;
}
__host__ __device__ inline size_t nnp2_params(int ng, int h1, int h2)
{
    return (size_t)ng * h1 + h1 + (size_t)h1 * h2 + h2 + h2 + 1;
}
// esz: bytes per element (8 for the FP64 kernels, 4 for binary32)
__host__ __device__ inline size_t nnp2_lds_bytes(int N, int ng, int nfn, int h1, int h2, int esz)
{
    const size_t d = (size_t)6 * N + 2 * (size_t)N * N + 4 * (size_t)N + (size_t)N * ng + 2 * (size_t)N +
                     2 * nnp2_params(ng, h1, h2) + nnp2_scratch_doubles(N, nfn, h1, h2);
    return d * (size_t)esz + 2 * (size_t)N * sizeof(int);
}

}  // namespace ptmc
