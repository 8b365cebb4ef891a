// `forward`: drop-in for the RuNNer forward pass that the reference reaches
// through ccall((:forward, "librunnerjulia.so"), ...)
// (MachineLearningPotential/ForwardPass.jl:60-84, :121-145), computed on the
// GPU in FP64 from the network's definition: for every atom b of the batch,
// a_0 = input[:, b] (column-major n_in x batch), a_l = act_l(W_l^T a_{l-1} + b_l)
// with each layer's weights stored input-major (w[i * n_out + j]) and followed
// by its biases; eatom[b] = a_last[0].  Activation codes: 1 linear, 2 tanh
// (the code of layer 1, the input layer, is not used).  Returns
// eatom[batch - 1], or NaN (eatom untouched) for an invalid description or
// a device error.  Calls are serialised; any number of host threads may call.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>
#include <vector>

#include "../../include/ptmc.h"

namespace {

constexpr int FWD_THREADS = 128;
constexpr int FWD_MAXW = 1024;  // widest layer
constexpr int FWD_MAXL = 16;    // layers

struct FwdLayers {
    int n_layers;
    int nodes[FWD_MAXL];
    int act[FWD_MAXL];
};

__global__ void __launch_bounds__(FWD_THREADS) forward_kernel(const double *in, const double *params, FwdLayers L,
                                                              double *eatom)
{
    __shared__ double a[2][FWD_MAXW];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int n0 = L.nodes[0];
    for (int i = tid; i < n0; i += FWD_THREADS) a[0][i] = in[(size_t)b * n0 + i];
    __syncthreads();
    const double *p = params;
    int cur = 0;
    for (int l = 1; l < L.n_layers; ++l) {
        const int ni = L.nodes[l - 1], no = L.nodes[l];
        const double *w = p, *bias = p + (size_t)ni * no;
        for (int j = tid; j < no; j += FWD_THREADS) {
            double s = 0.0;
            for (int i = 0; i < ni; ++i) s += a[cur][i] * w[(size_t)i * no + j];
            s += bias[j];
            a[cur ^ 1][j] = L.act[l] == 2 ? tanh(s) : s;
        }
        __syncthreads();
        cur ^= 1;
        p = bias + no;
    }
    if (tid == 0) eatom[b] = a[cur][0];
}

std::mutex g_mutex;
double *g_in = nullptr, *g_par = nullptr, *g_out = nullptr;
size_t g_in_n = 0, g_par_n = 0, g_out_n = 0;

bool grow(double *&p, size_t &have, size_t need)
{
    if (need <= have) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    have = 0;
    if (hipMalloc((void **)&p, need * sizeof(double)) != hipSuccess) return false;
    have = need;
    return true;
}

}  // namespace

extern "C" double forward(const double *input, const int32_t *n_in, const int32_t *n_layers, const int32_t *nodes,
                          const int32_t *batch, const int32_t *n_params, const double *params,
                          const int32_t *activations, double *eatom)
{
    const double bad = std::nan("");
    if (!input || !n_in || !n_layers || !nodes || !batch || !n_params || !params || !activations || !eatom) return bad;
    const int nl = *n_layers, B = *batch;
    if (nl < 2 || nl > FWD_MAXL || B < 1 || *n_in != nodes[0] || nodes[nl - 1] != 1) return bad;
    FwdLayers L{};
    L.n_layers = nl;
    long long need = 0;
    for (int l = 0; l < nl; ++l) {
        if (nodes[l] < 1 || nodes[l] > FWD_MAXW) return bad;
        L.nodes[l] = nodes[l];
        L.act[l] = activations[l];
        if (l > 0) {
            if (activations[l] != 1 && activations[l] != 2) return bad;
            need += (long long)nodes[l - 1] * nodes[l] + nodes[l];
        }
    }
    if (need != *n_params) return bad;
    std::lock_guard<std::mutex> lock(g_mutex);
    const size_t n_inp = (size_t)nodes[0] * B;
    if (!grow(g_in, g_in_n, n_inp) || !grow(g_par, g_par_n, (size_t)need) || !grow(g_out, g_out_n, (size_t)B))
        return bad;
    if (hipMemcpy(g_in, input, n_inp * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g_par, params, (size_t)need * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return bad;
    hipLaunchKernelGGL(forward_kernel, dim3(B), dim3(FWD_THREADS), 0, 0, g_in, g_par, L, g_out);
    if (hipGetLastError() != hipSuccess) return bad;
    if (hipMemcpy(eatom, g_out, (size_t)B * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return bad;
    return eatom[B - 1];
}
