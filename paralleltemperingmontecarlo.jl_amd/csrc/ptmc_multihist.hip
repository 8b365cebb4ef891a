// Multihistogram heat capacity (multihistogram_NVT, src/multihist_NVT.jl:11-221)
// in FP64 on the device, one workgroup per PT run.
//
// The reference iterates free energies f_i = log sum_m q(beta_i, m) with
//   q(b, m) = H_m / D_m * exp(-b E_m - off_m),
//   off_m   = max(-1e6, max_i(-beta_i E_m - f_i)) + log(1000),
//   D_m     = sum_i n_i exp(-beta_i E_m - f_i - off_m),  H_m = sum_i h_i(m).
// off_m and D_m do not depend on b, so each iteration computes them once per
// bin (same operations, same order as every call in the reference) and then
// the R sums over bins.  Heat capacities follow on a 10x denser temperature
// grid: Cv = (<E^2> - <E>^2) / (k T^2).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/ptmc.h"

namespace {

constexpr int MH_THREADS = 256;

struct MHArgs {
    int n_temps, n_bin, n_res, max_iter;
    double conv;
    const double *beta;      // runs x R
    const double *hist;      // runs x R x (n_bin + 2)
    const double *en_min;    // runs
    const double *en_max;    // runs
    const double *beta_res;  // runs x n_res
    const double *t_res;     // runs x n_res
    double k_boltzmann;
    double *cv;              // runs x n_res
    double *free_energy;     // runs x R
    int *iterations;         // runs
};

__global__ void __launch_bounds__(MH_THREADS) multihist_kernel(MHArgs a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const int R = a.n_temps, nb = a.n_bin, run = blockIdx.x, tid = threadIdx.x;
    double *fe = reinterpret_cast<double *>(lds);  // R
    double *nfe = fe + R;                          // R
    double *ncyc = nfe + R;                        // R
    double *beta = ncyc + R;                       // R
    double *off = beta + R;                        // nb
    double *ratio = off + nb;                      // nb
    double *hsum = ratio + nb;                     // nb
    int *flag = reinterpret_cast<int *>(hsum + nb);

    const double *H = a.hist + (size_t)run * R * (nb + 2);
    const double emin = a.en_min[run];
    const double de = (a.en_max[run] - emin) / (double)nb;
    for (int i = tid; i < R; i += MH_THREADS) {
        double s = 0.0;
        for (int m = 1; m <= nb; ++m) s = s + H[(size_t)i * (nb + 2) + m];
        ncyc[i] = s;
        fe[i] = 0.0;
        beta[i] = a.beta[(size_t)run * R + i];
    }
    __syncthreads();
    for (int m = tid; m < nb; m += MH_THREADS) {
        double s = 0.0;
        for (int i = 0; i < R; ++i) s = s + H[(size_t)i * (nb + 2) + m + 1];
        hsum[m] = s;
    }
    __syncthreads();

    auto bin_terms = [&]() {
        for (int m = tid; m < nb; m += MH_THREADS) {
            const double et = emin + ((double)(m + 1) - 0.5) * de;
            double o = -1.0e6;
            for (int i = 0; i < R; ++i) o = fmax(o, -beta[i] * et - fe[i]);
            o = o + log(1000.0);
            double d = 0.0;
            for (int i = 0; i < R; ++i) d = d + ncyc[i] * exp(-beta[i] * et - fe[i] - o);
            off[m] = o;
            ratio[m] = hsum[m] / d;
        }
    };

    int it = 0;
    for (it = 1; it <= a.max_iter; ++it) {
        bin_terms();
        __syncthreads();
        for (int i = tid; i < R; i += MH_THREADS) {
            double s = 0.0;
            for (int m = 0; m < nb; ++m) {
                const double et = emin + ((double)(m + 1) - 0.5) * de;
                s = s + ratio[m] * exp(-beta[i] * et - off[m]);
            }
            nfe[i] = log(s);
        }
        __syncthreads();
        if (tid == 0) {
            double delta = 0.0;
            for (int i = 0; i < R; ++i) {
                const double d = fabs(nfe[i] - fe[i]);
                delta = delta + d * d;
                fe[i] = nfe[i];
            }
            *flag = delta < a.conv;
        }
        __syncthreads();
        if (*flag) break;
    }
    if (it > a.max_iter) it = a.max_iter;
    bin_terms();
    __syncthreads();
    for (int j = tid; j < a.n_res; j += MH_THREADS) {
        const double bt = a.beta_res[(size_t)run * a.n_res + j];
        double nc = 0.0;
        for (int m = 0; m < nb; ++m) {
            const double et = emin + ((double)(m + 1) - 0.5) * de;
            nc = nc + ratio[m] * exp(-bt * et - off[m]);
        }
        double e1 = 0.0, e2 = 0.0;
        for (int m = 0; m < nb; ++m) {
            const double et = emin + ((double)(m + 1) - 0.5) * de;
            const double q = ratio[m] * exp(-bt * et - off[m]);
            e1 = e1 + q / nc * et;
            e2 = e2 + q / nc * (et * et);
        }
        const double t = a.t_res[(size_t)run * a.n_res + j];
        a.cv[(size_t)run * a.n_res + j] = (e2 - e1 * e1) / (a.k_boltzmann * (t * t));
    }
    for (int i = tid; i < R; i += MH_THREADS) a.free_energy[(size_t)run * R + i] = fe[i];
    if (tid == 0) a.iterations[run] = it;
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

extern "C" int ptmc_multihistogram_nvt(int32_t n_runs, int32_t n_temps, int32_t n_bin, const double *beta,
                                       const double *hist, const double *en_min, const double *en_max,
                                       int32_t n_result, const double *beta_result, const double *t_result,
                                       double k_boltzmann, double conv_threshold, int32_t max_iter, double *cv,
                                       double *free_energy, int32_t *iterations, int32_t device)
{
    if (n_runs <= 0 || n_temps <= 0 || n_bin <= 0 || n_result <= 0 || max_iter <= 0 || !beta || !hist ||
        !en_min || !en_max || !beta_result || !t_result || !cv)
        return PTMC_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PTMC_E_NO_DEVICE;
    if (device < 0 || device >= ndev || hipSetDevice(device) != hipSuccess) return PTMC_E_ARG;
    const size_t lds = sizeof(double) * (4 * (size_t)n_temps + 3 * (size_t)n_bin) + sizeof(int);
    int max_lds = 0;
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
        return PTMC_E_DEVICE;
    if (lds > (size_t)max_lds) return PTMC_E_UNSUPPORTED;
    const size_t nR = (size_t)n_runs * n_temps, nH = nR * (n_bin + 2), nX = (size_t)n_runs * n_result;
    DBuf<double> d_beta, d_hist, d_emin, d_emax, d_bres, d_tres, d_cv, d_fe;
    DBuf<int> d_it;
#define MHCHK(x)                                  \
    do {                                          \
        if ((x) != hipSuccess) return PTMC_E_DEVICE; \
    } while (0)
    MHCHK(hipMalloc((void **)&d_beta.p, nR * 8));
    MHCHK(hipMalloc((void **)&d_hist.p, nH * 8));
    MHCHK(hipMalloc((void **)&d_emin.p, n_runs * 8));
    MHCHK(hipMalloc((void **)&d_emax.p, n_runs * 8));
    MHCHK(hipMalloc((void **)&d_bres.p, nX * 8));
    MHCHK(hipMalloc((void **)&d_tres.p, nX * 8));
    MHCHK(hipMalloc((void **)&d_cv.p, nX * 8));
    MHCHK(hipMalloc((void **)&d_fe.p, nR * 8));
    MHCHK(hipMalloc((void **)&d_it.p, n_runs * sizeof(int)));
    MHCHK(hipMemcpy(d_beta.p, beta, nR * 8, hipMemcpyHostToDevice));
    MHCHK(hipMemcpy(d_hist.p, hist, nH * 8, hipMemcpyHostToDevice));
    MHCHK(hipMemcpy(d_emin.p, en_min, n_runs * 8, hipMemcpyHostToDevice));
    MHCHK(hipMemcpy(d_emax.p, en_max, n_runs * 8, hipMemcpyHostToDevice));
    MHCHK(hipMemcpy(d_bres.p, beta_result, nX * 8, hipMemcpyHostToDevice));
    MHCHK(hipMemcpy(d_tres.p, t_result, nX * 8, hipMemcpyHostToDevice));
    MHArgs a{};
    a.n_temps = n_temps;
    a.n_bin = n_bin;
    a.n_res = n_result;
    a.max_iter = max_iter;
    a.conv = conv_threshold;
    a.beta = d_beta.p;
    a.hist = d_hist.p;
    a.en_min = d_emin.p;
    a.en_max = d_emax.p;
    a.beta_res = d_bres.p;
    a.t_res = d_tres.p;
    a.k_boltzmann = k_boltzmann;
    a.cv = d_cv.p;
    a.free_energy = d_fe.p;
    a.iterations = d_it.p;
    hipLaunchKernelGGL(multihist_kernel, dim3(n_runs), dim3(MH_THREADS), lds, 0, a);
    MHCHK(hipGetLastError());
    MHCHK(hipDeviceSynchronize());
    MHCHK(hipMemcpy(cv, d_cv.p, nX * 8, hipMemcpyDeviceToHost));
    if (free_energy) MHCHK(hipMemcpy(free_energy, d_fe.p, nR * 8, hipMemcpyDeviceToHost));
    if (iterations) MHCHK(hipMemcpy(iterations, d_it.p, n_runs * sizeof(int), hipMemcpyDeviceToHost));
#undef MHCHK
    return PTMC_OK;
}
