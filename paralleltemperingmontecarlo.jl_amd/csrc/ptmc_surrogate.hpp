// Learned surrogate for delayed-acceptance Metropolis (Christen & Fox,
// J. Comput. Graph. Stat. 14 (2005) 795); it has no counterpart in the
// reference.  Per atom a:
//   d_a[k] = s_in sum_{b != a} exp(-((r_ab - mu_k) / sigma)^2) f_c(r_ab)
//   mu_k = r_min + k (r_cut - r_min) / (K - 1), sigma = 1.5 (r_cut - r_min) / (K - 1)
//   e_a = w3 . silu(W2^T silu(W1^T d_a + b1) + b2) + b3,  E_hat = sum_a e_a
// Parameters: one f32 vector W1 [K][H] | b1 [H] | W2 [H][H] | b2 [H] | w3 [H] | b3.
//
// Inference inside the sweep multiplies on the matrix cores in bf16
// (v_mfma_f32_32x32x8_bf16): the surrogate only filters proposals, the exact
// potential decides, so its rounding changes efficiency, not the sampled
// distribution.  Training runs forward and backward on v_mfma_f32_32x32x2_f32.
//
// MFMA lane layouts (one wave, 32x32 output tile):
//   32x32x8 bf16: A[row l%32][k 4(l/32)+i], B[k 4(l/32)+i][col l%32], i = 0..3
//   32x32x2 f32:  A[row l%32][k l/32],      B[k l/32][col l%32]
//   output value v of lane l: row 8(v/4) + 4(l/32) + v%4, column l%32
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ptmc_nnp_layout.hpp"
#include "ptmc_types.hpp"

namespace ptmc {

typedef short sur_bf16x4 __attribute__((ext_vector_type(4)));
typedef float sur_f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float sur_silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float sur_dsilu(float z)
{
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// round-to-nearest-even bf16 bits of a finite float
__device__ __forceinline__ short sur_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

struct SurGeom {
    float mu0, dmu, c2, r_cut;  // c2 = log2(e) / sigma^2
};

__device__ __forceinline__ SurGeom sur_geom(float r_cut)
{
    SurGeom g;
    g.mu0 = SUR_RMIN;
    g.dmu = (r_cut - SUR_RMIN) / (float)(SUR_K - 1);
    const float inv_sig = 1.0f / (1.5f * g.dmu);
    g.c2 = 1.4426950408889634f * inv_sig * inv_sig;
    g.r_cut = r_cut;
    return g;
}

__device__ __forceinline__ float sur_mu(const SurGeom &g, int k) { return g.mu0 + g.dmu * (float)k; }

// one descriptor term exp(-((r - mu_k)/sigma)^2) f_c(r) = 2^{-c2 (r - mu_k)^2} f_c(r)
__device__ __forceinline__ float sur_phi_mu(const SurGeom &g, float r, float fc, float mu)
{
    const float t = r - mu;
    return __builtin_amdgcn_exp2f(-g.c2 * t * t) * fc;
}

__device__ __forceinline__ float sur_phi(const SurGeom &g, float r, float fc, int k)
{
    return sur_phi_mu(g, r, fc, sur_mu(g, k));
}

__device__ __forceinline__ float sur_cutoff(float r, float rc) { return cutoff_cos2(r, rc); }

// Full descriptors of N atoms (positions X[3][NNP_NP]) into S[NNP_NP][stride];
// pad rows are zero.  nthreads threads call.
__device__ void sur_descriptors(const SurGeom &g, const float *X, int N, float *S, int stride, int nthreads)
{
    for (int e = threadIdx.x; e < NNP_NP * SUR_K; e += nthreads) {
        const int a = e / SUR_K, k = e % SUR_K;
        float s = 0.0f;
        if (a < N) {
            for (int b = 0; b < N; ++b) {
                if (b == a) continue;
                const float dx = X[a] - X[b], dy = X[NNP_NP + a] - X[NNP_NP + b],
                            dz = X[2 * NNP_NP + a] - X[2 * NNP_NP + b];
                const float r = sqrtf(dx * dx + dy * dy + dz * dz);
                const float fc = sur_cutoff(r, g.r_cut);
                if (fc != 0.0f) s += sur_phi(g, r, fc, k);
            }
        }
        S[a * stride + k] = SUR_SIN * s;
    }
}

// LDS of the bf16 inference network
struct SurNet {
    short *W1, *W2;      // [K][H], [H][H] bf16
    short *Hb;           // [NNP_NP][SUR_HS] bf16 hidden activations
    float *b1, *b2, *w3; // [H]
    float b3;
};

__device__ void sur_load_net(SurNet &n, const float *w, int nthreads)
{
    const float *W1 = w, *b1 = W1 + SUR_K * SUR_H, *W2 = b1 + SUR_H, *b2 = W2 + SUR_H * SUR_H, *w3 = b2 + SUR_H;
    for (int e = threadIdx.x; e < SUR_K * SUR_H; e += nthreads) {
        n.W1[e] = sur_bf16(W1[e]);
        n.W2[e] = sur_bf16(W2[e]);
    }
    for (int e = threadIdx.x; e < SUR_H; e += nthreads) {
        n.b1[e] = b1[e];
        n.b2[e] = b2[e];
        n.w3[e] = w3[e];
    }
    n.b3 = w3[SUR_H];
}

// Surrogate atom energies of descriptors S [NNP_NP][SUR_SS] (f32) into
// out[a], a < N, and optionally their FP64 total.  Waves 0-3 each own one
// 32x32 tile (atom half, hidden half); scratch Y [NNP_NP][SUR_YS] f32.  All
// threads of the workgroup call.
__device__ void sur_atom_energies(const SurNet &n, const float *S, int N, float *Y, float *out,
                                  double *total = nullptr)
{
    const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
    const int r0 = (w & 1) * 32, c0 = ((w >> 1) & 1) * 32;
    const int lr = l & 31, lh = l >> 5;
    if (w < 4) {
        sur_f32x16 c = {};
#pragma unroll
        for (int ks = 0; ks < SUR_K / 8; ++ks) {
            sur_bf16x4 A, B;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = ks * 8 + lh * 4 + i;
                A[i] = sur_bf16(S[(r0 + lr) * SUR_SS + k]);
                B[i] = n.W1[k * SUR_H + c0 + lr];
            }
            c = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(A, B, c, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int row = r0 + 8 * (v >> 2) + 4 * lh + (v & 3), col = c0 + lr;
            n.Hb[row * SUR_HS + col] = sur_bf16(sur_silu(c[v] + n.b1[col]));
        }
    }
    __syncthreads();
    if (w < 4) {
        sur_f32x16 c = {};
#pragma unroll
        for (int ks = 0; ks < SUR_H / 8; ++ks) {
            sur_bf16x4 A, B;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = ks * 8 + lh * 4 + i;
                A[i] = n.Hb[(r0 + lr) * SUR_HS + k];
                B[i] = n.W2[k * SUR_H + c0 + lr];
            }
            c = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(A, B, c, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int row = r0 + 8 * (v >> 2) + 4 * lh + (v & 3), col = c0 + lr;
            Y[row * SUR_YS + col] = sur_silu(c[v] + n.b2[col]) * n.w3[col];
        }
    }
    __syncthreads();
    if (w == 0) {
        // row sums with eight independent partial sums, then the FP64 total
        float p[8] = {};
#pragma unroll
        for (int h = 0; h < SUR_H; ++h) p[h & 7] += Y[l * SUR_YS + h];
        const float e = ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7])) + n.b3;
        if (l < N) out[l] = e;
        if (total != nullptr) {
            double t = l < N ? (double)e : 0.0;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
            if (l == 0) *total = t;
        }
    }
    __syncthreads();
}

}  // namespace ptmc
