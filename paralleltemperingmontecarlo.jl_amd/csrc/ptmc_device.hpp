// Device-side building blocks of the PT engine: counter-based random numbers,
// deterministic exp/log, group (lane-cluster) reductions.
//
// Random stream contract (identical on host-side checkers):
//   Philox4x32-10, key = (seed lo, seed hi),
//   counter = (c0, sweep, (type << 28) | (sub << 20) | temperature index, run)
// The reference draws its moves from Julia's global RNG
// (src/MCRun.jl:145, src/MCMoves.jl:34-38, src/MCRun.jl:123, :203,
// src/Exchange.jl:135, :145); the order and the meaning of the draws follow
// those lines, the bits come from Philox.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ptmc {

enum : uint32_t { RNG_MOVE = 0, RNG_EXCH = 1, RNG_INIT = 2, RNG_PARAM = 3 };

struct u32x4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        u32x4 n;
        n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
        n.y = (uint32_t)p1;
        n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        n.w = (uint32_t)p0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

__device__ __forceinline__ u32x4 draw(uint64_t seed, uint32_t run, uint32_t type, uint32_t sub,
                                      uint32_t temp, uint32_t sweep, uint32_t c0)
{
    u32x4 c;
    c.x = c0;
    c.y = sweep;
    c.z = (type << 28) | ((sub & 0xffu) << 20) | (temp & 0xfffffu);
    c.w = run;
    return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

__device__ __forceinline__ double u53(uint32_t hi, uint32_t lo)
{
    uint64_t m = ((uint64_t)hi << 21) | (uint64_t)(lo >> 11);
    return (double)m * 0x1.0p-53;
}

__device__ __forceinline__ float u24(uint32_t w) { return (float)(w >> 8) * 0x1.0p-24f; }

__device__ __forceinline__ uint32_t below(uint32_t w, uint32_t n) { return __umulhi(w, n); }

// exp(x): Cody-Waite reduction, degree-12 Taylor/Horner with fma, 2^k scale.
__device__ __forceinline__ double det_exp(double x)
{
    if (!(x == x)) return x;
    if (x > 709.782712893384) return __longlong_as_double(0x7ff0000000000000ll);
    if (x < -745.1332191019412) return 0.0;
    const double log2e = 1.4426950408889634;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    double kd = rint(x * log2e);
    double r = fma(-kd, ln2_hi, x);
    r = fma(-kd, ln2_lo, r);
    double p = 2.08767569878680989792e-09;
    p = fma(p, r, 2.50521083854417187751e-08);
    p = fma(p, r, 2.75573192239858906526e-07);
    p = fma(p, r, 2.75573192239858906526e-06);
    p = fma(p, r, 2.48015873015873015873e-05);
    p = fma(p, r, 1.98412698412698412698e-04);
    p = fma(p, r, 1.38888888888888888889e-03);
    p = fma(p, r, 8.33333333333333333333e-03);
    p = fma(p, r, 4.16666666666666666667e-02);
    p = fma(p, r, 1.66666666666666666667e-01);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    int k = (int)kd;
    if (k >= -1021) {
        if (k > 1023) return p * 2.0 * __longlong_as_double((long long)(1023 + k - 1) << 52);
        return p * __longlong_as_double((long long)(1023 + k) << 52);
    }
    return (p * __longlong_as_double((long long)(1023 + k + 64) << 52)) * 0x1.0p-64;
}

// log(x), x > 0: 2 atanh((m-1)/(m+1)) series to s^21 plus e ln2.
__device__ __forceinline__ double det_log(double x)
{
    if (!(x > 0.0)) return x == 0.0 ? -__longlong_as_double(0x7ff0000000000000ll)
                                    : __longlong_as_double(0x7ff8000000000000ll);
    if (x == __longlong_as_double(0x7ff0000000000000ll)) return x;
    int eadj = 0;
    if (x < 0x1.0p-1022) {
        x *= 0x1.0p54;
        eadj = -54;
    }
    uint64_t b = (uint64_t)__double_as_longlong(x);
    int e = (int)((b >> 52) & 0x7ff) - 1023 + eadj;
    double m = __longlong_as_double((long long)((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull));
    if (m > 1.4142135623730951) {
        m *= 0.5;
        e += 1;
    }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double q = 1.0 / 21.0;
    q = fma(q, z, 1.0 / 19.0);
    q = fma(q, z, 1.0 / 17.0);
    q = fma(q, z, 1.0 / 15.0);
    q = fma(q, z, 1.0 / 13.0);
    q = fma(q, z, 1.0 / 11.0);
    q = fma(q, z, 1.0 / 9.0);
    q = fma(q, z, 1.0 / 7.0);
    q = fma(q, z, 1.0 / 5.0);
    q = fma(q, z, 1.0 / 3.0);
    double two_s = 2.0 * s;
    double lm = fma(two_s * z, q, two_s);
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    double ed = (double)e;
    return fma(ed, ln2_hi, fma(ed, ln2_lo, lm));
}

// Sum over the KL lanes of a lane cluster with an xor butterfly.  Every lane
// ends with the same bits because each step adds the same two operands.
template <int KL, typename T>
__device__ __forceinline__ T group_sum(T v)
{
#pragma unroll
    for (int m = 1; m < KL; m <<= 1) v = v + __shfl_xor(v, m, 64);
    return v;
}

}  // namespace ptmc
