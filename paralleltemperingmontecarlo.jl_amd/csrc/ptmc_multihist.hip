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
#include <mutex>
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

// ---- multihistogram_NPT (src/multihist_NPT.jl:17-340): the same fixed-point
// iteration over the (energy, volume) bins of the sampled 2D histograms, with
// the enthalpy h = E + pV in place of E, and heat capacities from the
// enthalpy fluctuation, Cp = (<h^2> - <h>^2) / (k T^2).  n_bin^2 bins do not
// fit LDS: the per-bin terms live in a global scratch block and each thread
// always owns the same bins, so no barrier separates writing and reading
// them.  Sums over bins are per-thread partials combined by a fixed-order
// block reduction.
constexpr int MHN_THREADS = 256;

struct MHNArgs {
    int n_temps, n_bin, n_res, max_iter;
    double conv;
    const double *beta;      // runs x R
    const double *hist;      // runs x R x (n_bin + 2)^2, rows energy, columns volume
    const double *en_min, *en_max, *v_min, *v_max, *pressure;  // runs
    const double *beta_res;  // runs x n_res
    const double *t_res;     // runs x n_res
    double k_boltzmann;
    double *cp;              // runs x n_res
    double *moments;         // runs x n_res x 3: <E>, <V>, <h>
    double *free_energy;     // runs x R
    int *iterations;         // runs
    double *scratch;         // runs x 3 x n_bin^2: histogram sums, offsets, ratios
};

// Sum over the block in a fixed order (xor butterfly in each wave, then the
// wave totals in wave order); every thread returns the total.
__device__ double block_sum(double v, double *red)
{
    for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < MHN_THREADS / 64; ++k) t = t + red[k];
    __syncthreads();
    return t;
}

__global__ void __launch_bounds__(MHN_THREADS) multihist_npt_kernel(MHNArgs a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const int R = a.n_temps, nb = a.n_bin, run = blockIdx.x, tid = threadIdx.x;
    const int B = nb * nb, W = nb + 2;
    double *fe = reinterpret_cast<double *>(lds);  // R
    double *nfe = fe + R;                          // R
    double *ncyc = nfe + R;                        // R
    double *beta = ncyc + R;                       // R
    double *red = beta + R;                        // MHN_THREADS / 64
    int *flag = reinterpret_cast<int *>(red + MHN_THREADS / 64);

    const double *H = a.hist + (size_t)run * R * W * W;
    double *hs = a.scratch + (size_t)run * 3 * B;
    double *off = hs + B;
    double *ratio = off + B;
    const double emin = a.en_min[run], vmin = a.v_min[run], p = a.pressure[run];
    const double de = (a.en_max[run] - emin) / (double)nb;
    const double dv = (a.v_max[run] - vmin) / (double)nb;
    auto energy = [&](int b) { return emin + ((double)(b / nb + 1) - 0.5) * de; };
    auto volume = [&](int b) { return vmin + ((double)(b % nb + 1) - 0.5) * dv; };
    auto cell = [&](int i, int b) { return H[(size_t)i * W * W + (size_t)(b / nb + 1) * W + (b % nb + 1)]; };

    for (int i = tid; i < R; i += MHN_THREADS) {
        fe[i] = 0.0;
        beta[i] = a.beta[(size_t)run * R + i];
    }
    for (int b = tid; b < B; b += MHN_THREADS) {
        double s = 0.0;
        for (int i = 0; i < R; ++i) s = s + cell(i, b);
        hs[b] = s;
    }
    for (int i = 0; i < R; ++i) {
        double s = 0.0;
        for (int b = tid; b < B; b += MHN_THREADS) s = s + cell(i, b);
        s = block_sum(s, red);
        if (tid == 0) ncyc[i] = s;
    }
    __syncthreads();

    auto bin_terms = [&]() {
        for (int b = tid; b < B; b += MHN_THREADS) {
            const double h = energy(b) + p * volume(b);
            double o = -1.0e6;
            for (int i = 0; i < R; ++i) o = fmax(o, -beta[i] * h - fe[i]);
            o = o + log(1000.0);
            double d = 0.0;
            for (int i = 0; i < R; ++i) d = d + ncyc[i] * exp(-beta[i] * h - fe[i] - o);
            off[b] = o;
            ratio[b] = hs[b] / d;
        }
    };

    int it = 0;
    for (it = 1; it <= a.max_iter; ++it) {
        bin_terms();
        for (int i = 0; i < R; ++i) {
            double s = 0.0;
            for (int b = tid; b < B; b += MHN_THREADS) {
                const double h = energy(b) + p * volume(b);
                s = s + ratio[b] * exp(-beta[i] * h - off[b]);
            }
            s = block_sum(s, red);
            if (tid == 0) nfe[i] = log(s);
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
    for (int j = 0; j < a.n_res; ++j) {
        const double bt = a.beta_res[(size_t)run * a.n_res + j];
        double s = 0.0;
        for (int b = tid; b < B; b += MHN_THREADS) {
            const double h = energy(b) + p * volume(b);
            s = s + ratio[b] * exp(-bt * h - off[b]);
        }
        const double nc = block_sum(s, red);
        double e1 = 0.0, v1 = 0.0, h1 = 0.0, h2 = 0.0;
        for (int b = tid; b < B; b += MHN_THREADS) {
            const double et = energy(b), vol = volume(b);
            const double h = et + p * vol;
            const double w = ratio[b] * exp(-bt * h - off[b]) / nc;
            e1 = e1 + w * et;
            v1 = v1 + w * vol;
            h1 = h1 + w * h;
            h2 = h2 + w * (h * h);
        }
        e1 = block_sum(e1, red);
        v1 = block_sum(v1, red);
        h1 = block_sum(h1, red);
        h2 = block_sum(h2, red);
        if (tid == 0) {
            const double t = a.t_res[(size_t)run * a.n_res + j];
            a.cp[(size_t)run * a.n_res + j] = (h2 - h1 * h1) / (a.k_boltzmann * (t * t));
            if (a.moments) {
                double *mo = a.moments + ((size_t)run * a.n_res + j) * 3;
                mo[0] = e1;
                mo[1] = v1;
                mo[2] = h1;
            }
        }
    }
    if (a.free_energy)
        for (int i = tid; i < R; i += MHN_THREADS) a.free_energy[(size_t)run * R + i] = fe[i];
    if (tid == 0 && a.iterations) a.iterations[run] = it;
}

// Device-resident scratch of the allocation-free entry point: the per-bin
// terms of every run of one call (3 n_bin^2 doubles per run), one copy per
// GPU for the life of the process.  Every element a call reads it has
// written first (histogram sums, then offsets and ratios per iteration), so
// nothing needs clearing between calls; calls that overlap on one GPU share
// it and race.
static __device__ __align__(256) double npt_scratch[PTMC_NOALLOC_NPT_SCRATCH_DOUBLES];

// device address of npt_scratch on each GPU, looked up once per device
constexpr int MAX_DEVICES = 64;
std::mutex npt_scratch_mutex;
double *npt_scratch_ptr[MAX_DEVICES] = {};

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

extern "C" int ptmc_multihistogram_npt(int32_t n_runs, int32_t n_temps, int32_t n_bin, const double *beta,
                                       const double *hist, const double *en_min, const double *en_max,
                                       const double *v_min, const double *v_max, const double *pressure,
                                       int32_t n_result, const double *beta_result, const double *t_result,
                                       double k_boltzmann, double conv_threshold, int32_t max_iter, double *cp,
                                       double *moments, double *free_energy, int32_t *iterations, int32_t device)
{
    if (n_runs <= 0 || n_temps <= 0 || n_bin <= 0 || n_result <= 0 || max_iter <= 0 || !beta || !hist ||
        !en_min || !en_max || !v_min || !v_max || !pressure || !beta_result || !t_result || !cp)
        return PTMC_E_ARG;
    if ((int64_t)n_bin * n_bin > (int64_t)1 << 26) return PTMC_E_UNSUPPORTED;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PTMC_E_NO_DEVICE;
    if (device < 0 || device >= ndev || hipSetDevice(device) != hipSuccess) return PTMC_E_ARG;
    const size_t lds = sizeof(double) * (4 * (size_t)n_temps + MHN_THREADS / 64) + sizeof(int);
    int max_lds = 0;
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
        return PTMC_E_DEVICE;
    if (lds > (size_t)max_lds) return PTMC_E_UNSUPPORTED;
    const size_t W = (size_t)n_bin + 2, B = (size_t)n_bin * n_bin;
    const size_t nR = (size_t)n_runs * n_temps, nH = nR * W * W, nX = (size_t)n_runs * n_result;
    DBuf<double> d_beta, d_hist, d_bounds, d_bres, d_tres, d_cp, d_mom, d_fe, d_scr;
    DBuf<int> d_it;
#define MHCHK(x)                                     \
    do {                                             \
        if ((x) != hipSuccess) return PTMC_E_DEVICE; \
    } while (0)
    MHCHK(hipMalloc((void **)&d_beta.p, nR * 8));
    MHCHK(hipMalloc((void **)&d_hist.p, nH * 8));
    MHCHK(hipMalloc((void **)&d_bounds.p, 5 * (size_t)n_runs * 8));
    MHCHK(hipMalloc((void **)&d_bres.p, nX * 8));
    MHCHK(hipMalloc((void **)&d_tres.p, nX * 8));
    MHCHK(hipMalloc((void **)&d_cp.p, nX * 8));
    MHCHK(hipMalloc((void **)&d_mom.p, 3 * nX * 8));
    MHCHK(hipMalloc((void **)&d_fe.p, nR * 8));
    MHCHK(hipMalloc((void **)&d_scr.p, 3 * B * (size_t)n_runs * 8));
    MHCHK(hipMalloc((void **)&d_it.p, n_runs * sizeof(int)));
    MHCHK(hipMemcpy(d_beta.p, beta, nR * 8, hipMemcpyHostToDevice));
    MHCHK(hipMemcpy(d_hist.p, hist, nH * 8, hipMemcpyHostToDevice));
    const double *bounds[5] = {en_min, en_max, v_min, v_max, pressure};
    for (int k = 0; k < 5; ++k)
        MHCHK(hipMemcpy(d_bounds.p + (size_t)k * n_runs, bounds[k], n_runs * 8, hipMemcpyHostToDevice));
    MHCHK(hipMemcpy(d_bres.p, beta_result, nX * 8, hipMemcpyHostToDevice));
    MHCHK(hipMemcpy(d_tres.p, t_result, nX * 8, hipMemcpyHostToDevice));
    MHNArgs a{};
    a.n_temps = n_temps;
    a.n_bin = n_bin;
    a.n_res = n_result;
    a.max_iter = max_iter;
    a.conv = conv_threshold;
    a.beta = d_beta.p;
    a.hist = d_hist.p;
    a.en_min = d_bounds.p;
    a.en_max = d_bounds.p + n_runs;
    a.v_min = d_bounds.p + 2 * (size_t)n_runs;
    a.v_max = d_bounds.p + 3 * (size_t)n_runs;
    a.pressure = d_bounds.p + 4 * (size_t)n_runs;
    a.beta_res = d_bres.p;
    a.t_res = d_tres.p;
    a.k_boltzmann = k_boltzmann;
    a.cp = d_cp.p;
    a.moments = d_mom.p;
    a.free_energy = d_fe.p;
    a.iterations = d_it.p;
    a.scratch = d_scr.p;
    hipLaunchKernelGGL(multihist_npt_kernel, dim3(n_runs), dim3(MHN_THREADS), lds, 0, a);
    MHCHK(hipGetLastError());
    MHCHK(hipDeviceSynchronize());
    MHCHK(hipMemcpy(cp, d_cp.p, nX * 8, hipMemcpyDeviceToHost));
    if (moments) MHCHK(hipMemcpy(moments, d_mom.p, 3 * nX * 8, hipMemcpyDeviceToHost));
    if (free_energy) MHCHK(hipMemcpy(free_energy, d_fe.p, nR * 8, hipMemcpyDeviceToHost));
    if (iterations) MHCHK(hipMemcpy(iterations, d_it.p, n_runs * sizeof(int), hipMemcpyDeviceToHost));
#undef MHCHK
    return PTMC_OK;
}

extern "C" int ptmc_multihistogram_npt_noalloc(int32_t n_runs, int32_t n_temps, int32_t n_bin, const double *beta,
                                               const double *hist, const double *en_min, const double *en_max,
                                               const double *v_min, const double *v_max, const double *pressure,
                                               int32_t n_result, const double *beta_result, const double *t_result,
                                               double k_boltzmann, double conv_threshold, int32_t max_iter,
                                               double *cp, double *moments, double *free_energy,
                                               int32_t *iterations, int32_t device, void *stream)
{
    if (n_runs <= 0 || n_temps <= 0 || n_bin <= 0 || n_result <= 0 || max_iter <= 0 || !beta || !hist ||
        !en_min || !en_max || !v_min || !v_max || !pressure || !beta_result || !t_result || !cp)
        return PTMC_E_ARG;
    if ((int64_t)n_runs * 3 * n_bin * n_bin > (int64_t)PTMC_NOALLOC_NPT_SCRATCH_DOUBLES) return PTMC_E_UNSUPPORTED;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PTMC_E_NO_DEVICE;
    if (device < 0 || device >= ndev || device >= MAX_DEVICES || hipSetDevice(device) != hipSuccess)
        return PTMC_E_ARG;
    const size_t lds = sizeof(double) * (4 * (size_t)n_temps + MHN_THREADS / 64) + sizeof(int);
    int max_lds = 0;
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
        return PTMC_E_DEVICE;
    if (lds > (size_t)max_lds) return PTMC_E_UNSUPPORTED;
    double *scr = nullptr;
    {
        std::lock_guard<std::mutex> lock(npt_scratch_mutex);
        if (!npt_scratch_ptr[device]) {
            void *p = nullptr;
            if (hipGetSymbolAddress(&p, HIP_SYMBOL(npt_scratch)) != hipSuccess || !p) return PTMC_E_DEVICE;
            npt_scratch_ptr[device] = static_cast<double *>(p);
        }
        scr = npt_scratch_ptr[device];
    }
    MHNArgs a{};
    a.n_temps = n_temps;
    a.n_bin = n_bin;
    a.n_res = n_result;
    a.max_iter = max_iter;
    a.conv = conv_threshold;
    a.beta = beta;
    a.hist = hist;
    a.en_min = en_min;
    a.en_max = en_max;
    a.v_min = v_min;
    a.v_max = v_max;
    a.pressure = pressure;
    a.beta_res = beta_result;
    a.t_res = t_result;
    a.k_boltzmann = k_boltzmann;
    a.cp = cp;
    a.moments = moments;
    a.free_energy = free_energy;
    a.iterations = iterations;
    a.scratch = scr;
    hipLaunchKernelGGL(multihist_npt_kernel, dim3(n_runs), dim3(MHN_THREADS), lds, static_cast<hipStream_t>(stream),
                       a);
    if (hipGetLastError() != hipSuccess) return PTMC_E_DEVICE;
    return PTMC_OK;
}
