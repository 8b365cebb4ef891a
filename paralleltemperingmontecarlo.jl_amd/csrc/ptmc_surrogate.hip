// Surrogate data collection and training kernels (see ptmc_surrogate.hpp).
//
// Samples are (descriptors, exact energy) pairs of the configurations the PT
// run visits.  Training minimises (1/B) sum_c ((E_hat_c - E_c) / N_c)^2 with
// Adam; one workgroup computes the forward pass and the gradient of one
// configuration on the matrix cores (f32), a second kernel averages the
// per-configuration gradients in a fixed order, a third applies Adam.
#include "ptmc_surrogate.hpp"

namespace ptmc {

constexpr int SUR_TT = 256;  // threads of the training kernels (4 waves, one 32x32 tile each)

// offsets in the packed parameter vector
constexpr int SO_W1 = 0;
constexpr int SO_B1 = SO_W1 + SUR_K * SUR_H;
constexpr int SO_W2 = SO_B1 + SUR_H;
constexpr int SO_B2 = SO_W2 + SUR_H * SUR_H;
constexpr int SO_W3 = SO_B2 + SUR_H;
constexpr int SO_B3 = SO_W3 + SUR_H;

// descriptors of sample configurations x [n][3][NNP_NP] into bd [n][NNP_NP][K]
__global__ void __launch_bounds__(SUR_TT) sur_desc_kernel(const float *x, const int32_t *n_atoms, float *bd,
                                                          float r_cut)
{
    __shared__ float X[3 * NNP_NP];
    const int s = blockIdx.x;
    for (int e = threadIdx.x; e < 3 * NNP_NP; e += SUR_TT) X[e] = x[(size_t)s * 3 * NNP_NP + e];
    __syncthreads();
    sur_descriptors(sur_geom(r_cut), X, n_atoms[s], bd + (size_t)s * NNP_NP * SUR_K, SUR_K, SUR_TT);
}

// copy every slot's configuration into the sample ring: descriptors and the
// exact energy (slot s goes to ring position (base + s) mod cap)
__global__ void __launch_bounds__(SUR_TT) sur_collect_kernel(SweepArgs a, float *bd, double *be, int32_t *bn,
                                                             long long base, int cap)
{
    __shared__ float X[3 * NNP_NP];
    const int slot = blockIdx.x;
    const int N = a.runs[a.slot_run[slot]].n_atoms;
    const long long s = (base + slot) % cap;
    const float *gX = reinterpret_cast<const float *>(a.coords) + (size_t)slot * 3 * NNP_NP;
    for (int e = threadIdx.x; e < 3 * NNP_NP; e += SUR_TT) X[e] = gX[e];
    __syncthreads();
    sur_descriptors(sur_geom(a.sur_rcut), X, N, bd + (size_t)s * NNP_NP * SUR_K, SUR_K, SUR_TT);
    if (threadIdx.x == 0) {
        be[s] = a.slot_en[slot];
        bn[s] = N;
    }
}

__device__ __forceinline__ sur_f32x16 mfma_f32(float a, float b, sur_f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}

// Forward pass (and, with grad, the gradient) of sample idx[b]:
// pred[b] = E_hat; partial[b] = d loss_b / d params with
// loss_b = ((E_hat - E) / N)^2 (the batch mean is taken by sur_reduce_kernel).
__global__ void __launch_bounds__(SUR_TT) sur_train_kernel(const float *bd, const double *be, const int32_t *bn,
                                                           const int32_t *idx, const float *w, float *partial,
                                                           double *pred, int grad)
{
    extern __shared__ __align__(16) float sm[];
    constexpr int SS = SUR_SS;
    float *D = sm;                     // [NP][SS] descriptors
    float *W1 = D + NNP_NP * SS;       // [K][SS]
    float *W2 = W1 + SUR_K * SS;       // [H][SS]
    float *Z1 = W2 + SUR_H * SS;       // [NP][SS]
    float *Z2 = Z1 + NNP_NP * SS;      // [NP][SS], later dZ1
    float *dZ2 = Z2 + NNP_NP * SS;     // [NP][SS]
    float *vb1 = dZ2 + NNP_NP * SS, *vb2 = vb1 + SUR_H, *vw3 = vb2 + SUR_H, *ea = vw3 + SUR_H;
    __shared__ double s_e;
    const int b = blockIdx.x, s = idx[b];
    const int N = bn[s];
    const int tid = threadIdx.x, l = tid & 63, wv = tid >> 6;
    const int r0 = (wv & 1) * 32, c0 = (wv >> 1) * 32, lr = l & 31, lh = l >> 5;
    const float *gD = bd + (size_t)s * NNP_NP * SUR_K;
    for (int e = tid; e < NNP_NP * SUR_K; e += SUR_TT) {
        const int r = e / SUR_K, c = e % SUR_K;
        D[r * SS + c] = gD[e];
        W1[r * SS + c] = w[SO_W1 + e];
        W2[r * SS + c] = w[SO_W2 + e];
    }
    for (int e = tid; e < SUR_H; e += SUR_TT) {
        vb1[e] = w[SO_B1 + e];
        vb2[e] = w[SO_B2 + e];
        vw3[e] = w[SO_W3 + e];
    }
    const float b3 = w[SO_B3];
    __syncthreads();
    // Z1 = D W1 + b1
    {
        sur_f32x16 c = {};
        for (int st = 0; st < SUR_K / 2; ++st) {
            const int k = 2 * st + lh;
            c = mfma_f32(D[(r0 + lr) * SS + k], W1[k * SS + c0 + lr], c);
        }
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int row = r0 + 8 * (v >> 2) + 4 * lh + (v & 3), col = c0 + lr;
            Z1[row * SS + col] = c[v] + vb1[col];
        }
    }
    __syncthreads();
    // Z2 = silu(Z1) W2 + b2
    {
        sur_f32x16 c = {};
        for (int st = 0; st < SUR_H / 2; ++st) {
            const int k = 2 * st + lh;
            c = mfma_f32(sur_silu(Z1[(r0 + lr) * SS + k]), W2[k * SS + c0 + lr], c);
        }
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int row = r0 + 8 * (v >> 2) + 4 * lh + (v & 3), col = c0 + lr;
            Z2[row * SS + col] = c[v] + vb2[col];
        }
    }
    __syncthreads();
    if (wv == 0) {
        float e = 0.0f;
        if (l < N) {
            for (int h = 0; h < SUR_H; ++h) e += sur_silu(Z2[l * SS + h]) * vw3[h];
            e += b3;
        }
        ea[l] = e;
    }
    __syncthreads();
    if (tid == 0) {
        double E = 0.0;
        for (int q = 0; q < N; ++q) E += (double)ea[q];
        s_e = E;
        pred[b] = E;
    }
    __syncthreads();
    if (!grad) return;  // uniform
    const float g = (float)(2.0 * (s_e - be[s]) / ((double)N * (double)N));
    float *P = partial + (size_t)b * SUR_P;
    // dZ2 = g w3 silu'(Z2) on real atoms
    for (int e = tid; e < NNP_NP * SUR_H; e += SUR_TT) {
        const int q = e / SUR_H, h = e % SUR_H;
        dZ2[q * SS + h] = q < N ? g * vw3[h] * sur_dsilu(Z2[q * SS + h]) : 0.0f;
    }
    __syncthreads();
    if (tid < SUR_H) {
        float sw3 = 0.0f, sb2 = 0.0f;
        for (int q = 0; q < N; ++q) {
            sw3 += g * sur_silu(Z2[q * SS + tid]);
            sb2 += dZ2[q * SS + tid];
        }
        P[SO_W3 + tid] = sw3;
        P[SO_B2 + tid] = sb2;
    }
    if (tid == 0) P[SO_B3] = g * (float)N;
    // dW2 = silu(Z1)^T dZ2 (reduction over atoms)
    {
        sur_f32x16 c = {};
        for (int st = 0; st < NNP_NP / 2; ++st) {
            const int q = 2 * st + lh;
            c = mfma_f32(sur_silu(Z1[q * SS + r0 + lr]), dZ2[q * SS + c0 + lr], c);
        }
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int row = r0 + 8 * (v >> 2) + 4 * lh + (v & 3), col = c0 + lr;
            P[SO_W2 + row * SUR_H + col] = c[v];
        }
    }
    __syncthreads();  // Z2 is read above and overwritten below
    // dZ1 = (dZ2 W2^T) silu'(Z1), stored in Z2
    {
        sur_f32x16 c = {};
        for (int st = 0; st < SUR_H / 2; ++st) {
            const int k = 2 * st + lh;
            c = mfma_f32(dZ2[(r0 + lr) * SS + k], W2[(c0 + lr) * SS + k], c);
        }
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int row = r0 + 8 * (v >> 2) + 4 * lh + (v & 3), col = c0 + lr;
            Z2[row * SS + col] = c[v] * sur_dsilu(Z1[row * SS + col]);
        }
    }
    __syncthreads();
    // dW1 = D^T dZ1, db1 = column sums of dZ1
    {
        sur_f32x16 c = {};
        for (int st = 0; st < NNP_NP / 2; ++st) {
            const int q = 2 * st + lh;
            c = mfma_f32(D[q * SS + r0 + lr], Z2[q * SS + c0 + lr], c);
        }
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int row = r0 + 8 * (v >> 2) + 4 * lh + (v & 3), col = c0 + lr;
            P[SO_W1 + row * SUR_H + col] = c[v];
        }
    }
    if (tid < SUR_H) {
        float sb1 = 0.0f;
        for (int q = 0; q < N; ++q) sb1 += Z2[q * SS + tid];
        P[SO_B1 + tid] = sb1;
    }
}

// grad[p] = mean over the batch of partial[b][p] (fixed summation order)
__global__ void sur_reduce_kernel(const float *partial, int B, float *grad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= SUR_P) return;
    double s = 0.0;
    for (int b = 0; b < B; ++b) s += (double)partial[(size_t)b * SUR_P + p];
    grad[p] = (float)(s / (double)B);
}

// Adam (Kingma & Ba 2015) with bias correction, step t >= 1
__global__ void sur_adam_kernel(float *w, const float *grad, float *m, float *v, int t, float lr)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= SUR_P) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / (1.0f - powf(b1, (float)t));
    const float vh = vv / (1.0f - powf(b2, (float)t));
    w[p] -= lr * mh / (sqrtf(vh) + eps);
}

size_t sur_train_lds_bytes()
{
    return sizeof(float) * ((size_t)NNP_NP * SUR_SS * 4 + (size_t)SUR_K * SUR_SS + (size_t)SUR_H * SUR_SS + 3 * SUR_H +
                            NNP_NP);
}

}  // namespace ptmc

namespace ptmc {

// host-side launchers used by the engine
hipError_t sur_launch_desc(const float *x, const int32_t *n_atoms, float *bd, float r_cut, int n, hipStream_t st)
{
    hipLaunchKernelGGL(sur_desc_kernel, dim3(n), dim3(SUR_TT), 0, st, x, n_atoms, bd, r_cut);
    return hipGetLastError();
}

hipError_t sur_launch_collect(const SweepArgs &a, int n_slots, float *bd, double *be, int32_t *bn, long long base,
                              int cap, hipStream_t st)
{
    hipLaunchKernelGGL(sur_collect_kernel, dim3(n_slots), dim3(SUR_TT), 0, st, a, bd, be, bn, base, cap);
    return hipGetLastError();
}

hipError_t sur_launch_train(const float *bd, const double *be, const int32_t *bn, const int32_t *idx, const float *w,
                            float *partial, double *pred, int grad, int batch, hipStream_t st)
{
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)sur_train_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)sur_train_lds_bytes());
        if (e != hipSuccess) return e;
        attr = true;
    }
    hipLaunchKernelGGL(sur_train_kernel, dim3(batch), dim3(SUR_TT), sur_train_lds_bytes(), st, bd, be, bn, idx, w,
                       partial, pred, grad);
    return hipGetLastError();
}

hipError_t sur_launch_reduce(const float *partial, int batch, float *grad, hipStream_t st)
{
    hipLaunchKernelGGL(sur_reduce_kernel, dim3((SUR_P + 255) / 256), dim3(256), 0, st, partial, batch, grad);
    return hipGetLastError();
}

hipError_t sur_launch_adam(float *w, const float *grad, float *m, float *v, int t, float lr, hipStream_t st)
{
    hipLaunchKernelGGL(sur_adam_kernel, dim3((SUR_P + 255) / 256), dim3(256), 0, st, w, grad, m, v, t, lr);
    return hipGetLastError();
}

}  // namespace ptmc
