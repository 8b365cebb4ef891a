// Two-species RuNNer network potential (RuNNerPotential2Atom,
// src/EnergyEvaluation.jl:1287-1317, :1748-1768, :1822-1856;
// MachineLearningPotential/SymmFunc.jl:54-175, :313-547): FP64 energies of a
// batch of configurations on the device, one workgroup per configuration.
//
// Atoms 0..n1-1 are species 1, n1..N-1 species 2.  A radial function gives
// two feature rows per atom (neighbours of species 1, of species 2), an
// angular function three (neighbour pairs 11, 12, 22); rows are scaled per
// centre species (radial g * norm + offset, angular g * tpz + offset) and
// each species has its own network (88-20-20-1 in the reference, any shape
// of the RuNNer layout here).  The features are gathered per (function,
// atom): a thread owns one and sums over partners (radial) or partner pairs
// (angular) in a fixed order, so no atomics and the result does not depend
// on the schedule; the reference accumulates per pair / triple instead, so
// the two agree to rounding.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/ptmc.h"

namespace {

constexpr int N2_THREADS = 256;
constexpr int N2_MAXN = 64;    // atoms per configuration
constexpr int N2_MAXG = 128;   // feature rows
constexpr int N2_MAXW = 64;    // widest hidden layer
constexpr int N2_MAXL = 8;     // layers

struct N2Args {
    int n1, n2, n_rad, n_ang, n_layers;
    int nodes[N2_MAXL], act[N2_MAXL];
    double r_cut;
    const double *x;       // n_conf x N x 3
    const double *rad;     // n_rad x 9: eta, norm[4], offset[4]
    const double *ang;     // n_ang x 15: eta, lambda, zeta, tpz[6], offset[6]
    const double *p1, *p2; // networks of species 1 and 2
    double *energy;        // n_conf
    double *eatom;         // n_conf x N (may be null)
};

__device__ __forceinline__ double cutoff(double r, double rc)
{
    // MachineLearningPotential/Cutoff.jl:79-90
    return r < rc ? 0.5 * (cos(3.14159265358979323846 * (r / rc)) + 1.0) : 0.0;
}

__global__ void __launch_bounds__(N2_THREADS) nnp2_energy_kernel(N2Args a)
{
    __shared__ double sx[N2_MAXN * 3];
    __shared__ double sr2[N2_MAXN * N2_MAXN];
    __shared__ double sf[N2_MAXN * N2_MAXN];
    __shared__ double sg[N2_MAXN * N2_MAXG];
    __shared__ double act[2][N2_MAXW > N2_MAXG ? N2_MAXW : N2_MAXG];
    __shared__ double seat[N2_MAXN];
    const int conf = blockIdx.x, tid = threadIdx.x;
    const int n1 = a.n1, N = a.n1 + a.n2, Ng = 2 * a.n_rad + 3 * a.n_ang;
    const double *x = a.x + (size_t)conf * N * 3;
    for (int e = tid; e < 3 * N; e += N2_THREADS) sx[e] = x[e];
    __syncthreads();
    for (int e = tid; e < N * N; e += N2_THREADS) {
        const int p = e / N, q = e % N;
        const double dx = sx[3 * p] - sx[3 * q], dy = sx[3 * p + 1] - sx[3 * q + 1], dz = sx[3 * p + 2] - sx[3 * q + 2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        sr2[e] = r2;
        sf[e] = cutoff(sqrt(r2), a.r_cut);
    }
    __syncthreads();

    // radial rows (SymmFunc.jl:319-361): pair value fc exp(-eta r^2), zero at fc = 0 or 1
    for (int t = tid; t < a.n_rad * N; t += N2_THREADS) {
        const int q = t / N, i = t % N;
        const double *R = a.rad + (size_t)q * 9;
        const double eta = R[0];
        double s1 = 0.0, s2 = 0.0;
        for (int j = 0; j < N; ++j) {
            if (j == i) continue;
            const double fc = sf[i * N + j];
            const double v = (fc != 0.0 && fc != 1.0) ? fc * exp(-eta * sr2[i * N + j]) : 0.0;
            if (j < n1) s1 += v;
            else s2 += v;
        }
        const int sp = i < n1 ? 0 : 2;
        sg[i * N2_MAXG + 2 * q] = s1 * R[1 + sp] + R[5 + sp];
        sg[i * N2_MAXG + 2 * q + 1] = s2 * R[2 + sp] + R[6 + sp];
    }
    // angular rows (SymmFunc.jl:363-468): triple value exp(-eta (r_ij^2 + r_ik^2 + r_jk^2))
    // f_ij f_ik f_jk (1 + lambda cos theta_i)^zeta at centre i, row by the partners' species
    for (int t = tid; t < a.n_ang * N; t += N2_THREADS) {
        const int q = t / N, i = t % N;
        const double *Ag = a.ang + (size_t)q * 15;
        const double eta = Ag[0], lam = Ag[1], zeta = Ag[2];
        double s[3] = {0.0, 0.0, 0.0};
        const double xi = sx[3 * i], yi = sx[3 * i + 1], zi = sx[3 * i + 2];
        for (int j = 0; j < N; ++j) {
            if (j == i) continue;
            const double fij = sf[i * N + j];
            if (fij == 0.0) continue;
            const double rij = sr2[i * N + j], dij = sqrt(rij);
            const double ax = xi - sx[3 * j], ay = yi - sx[3 * j + 1], az = zi - sx[3 * j + 2];
            for (int k = j + 1; k < N; ++k) {
                if (k == i) continue;
                const double fik = sf[i * N + k], fjk = sf[j * N + k];
                if (fik == 0.0 || fjk == 0.0) continue;
                const double rik = sr2[i * N + k], rjk = sr2[j * N + k];
                const double bx = xi - sx[3 * k], by = yi - sx[3 * k + 1], bz = zi - sx[3 * k + 2];
                const double cth = (ax * bx + ay * by + az * bz) / (dij * sqrt(rik));
                const double e = exp(-eta * (rij + rik + rjk)) * fij * fik * fjk;
                const int row = (j < n1 ? 0 : 1) + (k < n1 ? 0 : 1);
                s[row] += e * pow(1.0 + lam * cth, zeta);
            }
        }
        const int sp = i < n1 ? 0 : 3;
        for (int r = 0; r < 3; ++r) sg[i * N2_MAXG + 2 * a.n_rad + 3 * q + r] = s[r] * Ag[3 + sp + r] + Ag[9 + sp + r];
    }
    __syncthreads();

    // networks (RuNNer layout: per layer input-major weights, then biases)
    for (int i = 0; i < N; ++i) {
        const double *p = i < n1 ? a.p1 : a.p2;
        for (int k = tid; k < Ng; k += N2_THREADS) act[0][k] = sg[i * N2_MAXG + k];
        __syncthreads();
        int cur = 0;
        for (int l = 1; l < a.n_layers; ++l) {
            const int ni = a.nodes[l - 1], no = a.nodes[l];
            const double *w = p, *bias = p + (size_t)ni * no;
            for (int j = tid; j < no; j += N2_THREADS) {
                double s = 0.0;
                for (int k = 0; k < ni; ++k) s += act[cur][k] * w[(size_t)k * no + j];
                s += bias[j];
                act[cur ^ 1][j] = a.act[l] == 2 ? tanh(s) : s;
            }
            __syncthreads();
            cur ^= 1;
            p = bias + no;
        }
        if (tid == 0) seat[i] = act[cur][0];
        __syncthreads();
    }
    if (tid == 0) {
        double tot = 0.0;
        for (int i = 0; i < N; ++i) tot += seat[i];
        a.energy[conf] = tot;
    }
    if (a.eatom)
        for (int i = tid; i < N; i += N2_THREADS) a.eatom[(size_t)conf * N + i] = seat[i];
}

template <typename T>
struct DBuf {
    T *p = nullptr;
    ~DBuf()
    {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" int ptmc_nnp2_energy(int32_t n_conf, int32_t n1, int32_t n2, const double *x, int32_t n_rad,
                                int32_t n_ang, const double *rad, const double *ang, double r_cut, int32_t n_layers,
                                const int32_t *nodes, const int32_t *activations, int32_t n_params,
                                const double *params1, const double *params2, double *energy, double *eatom,
                                int32_t device)
{
    const int N = n1 + n2, Ng = 2 * n_rad + 3 * n_ang;
    if (n_conf <= 0 || n1 < 0 || n2 < 0 || N < 1 || N > N2_MAXN || n_rad < 0 || n_ang < 0 || Ng < 1 ||
        Ng > N2_MAXG || !x || (n_rad && !rad) || (n_ang && !ang) || !(r_cut > 0) || n_layers < 2 ||
        n_layers > N2_MAXL || !nodes || !activations || !params1 || !params2 || !energy || nodes[0] != Ng ||
        nodes[n_layers - 1] != 1)
        return PTMC_E_ARG;
    int need = 0;
    for (int l = 1; l < n_layers; ++l) {
        if (nodes[l] < 1 || nodes[l] > N2_MAXW || (activations[l] != 1 && activations[l] != 2)) return PTMC_E_ARG;
        need += nodes[l - 1] * nodes[l] + nodes[l];
    }
    if (need != n_params) return PTMC_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PTMC_E_NO_DEVICE;
    if (device < 0 || device >= ndev || hipSetDevice(device) != hipSuccess) return PTMC_E_ARG;
    N2Args a{};
    a.n1 = n1;
    a.n2 = n2;
    a.n_rad = n_rad;
    a.n_ang = n_ang;
    a.n_layers = n_layers;
    for (int l = 0; l < n_layers; ++l) {
        a.nodes[l] = nodes[l];
        a.act[l] = activations[l];
    }
    a.r_cut = r_cut;
    DBuf<double> dx, drad, dang, dp1, dp2, de, dea;
#define N2CHK(e)                                     \
    do {                                             \
        if ((e) != hipSuccess) return PTMC_E_DEVICE; \
    } while (0)
    const size_t nx = (size_t)n_conf * N * 3;
    N2CHK(hipMalloc((void **)&dx.p, nx * 8));
    N2CHK(hipMemcpy(dx.p, x, nx * 8, hipMemcpyHostToDevice));
    N2CHK(hipMalloc((void **)&drad.p, (size_t)(n_rad ? n_rad : 1) * 9 * 8));
    if (n_rad) N2CHK(hipMemcpy(drad.p, rad, (size_t)n_rad * 9 * 8, hipMemcpyHostToDevice));
    N2CHK(hipMalloc((void **)&dang.p, (size_t)(n_ang ? n_ang : 1) * 15 * 8));
    if (n_ang) N2CHK(hipMemcpy(dang.p, ang, (size_t)n_ang * 15 * 8, hipMemcpyHostToDevice));
    N2CHK(hipMalloc((void **)&dp1.p, (size_t)n_params * 8));
    N2CHK(hipMemcpy(dp1.p, params1, (size_t)n_params * 8, hipMemcpyHostToDevice));
    N2CHK(hipMalloc((void **)&dp2.p, (size_t)n_params * 8));
    N2CHK(hipMemcpy(dp2.p, params2, (size_t)n_params * 8, hipMemcpyHostToDevice));
    N2CHK(hipMalloc((void **)&de.p, (size_t)n_conf * 8));
    if (eatom) N2CHK(hipMalloc((void **)&dea.p, (size_t)n_conf * N * 8));
    a.x = dx.p;
    a.rad = drad.p;
    a.ang = dang.p;
    a.p1 = dp1.p;
    a.p2 = dp2.p;
    a.energy = de.p;
    a.eatom = dea.p;
    hipLaunchKernelGGL(nnp2_energy_kernel, dim3(n_conf), dim3(N2_THREADS), 0, 0, a);
    N2CHK(hipGetLastError());
    N2CHK(hipDeviceSynchronize());
    N2CHK(hipMemcpy(energy, de.p, (size_t)n_conf * 8, hipMemcpyDeviceToHost));
    if (eatom) N2CHK(hipMemcpy(eatom, dea.p, (size_t)n_conf * N * 8, hipMemcpyDeviceToHost));
#undef N2CHK
    return PTMC_OK;
}
