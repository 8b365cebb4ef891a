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

// DPP lane permutations (VALU, no LDS round trip).
template <int CTRL>
__device__ __forceinline__ float dpp_perm(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_perm(double v)
{
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
    int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum over the KL lanes of a lane cluster (KL <= 16: one DPP row).  Step k
// adds the partner whose index differs in bit k after steps 0..k-1 have made
// every lane of a 2^k block hold the block sum: quad_perm [1,0,3,2] (xor 1),
// quad_perm [2,3,0,1] (xor 2), row_half_mirror and row_mirror pair blocks of
// 4 and 8.  Each step adds the same two operands on every lane, so all lanes
// end with identical bits, equal to an xor butterfly.
template <int KL, typename T>
__device__ __forceinline__ T group_sum(T v)
{
    static_assert(KL == 1 || KL == 2 || KL == 4 || KL == 8 || KL == 16, "cluster width");
    if constexpr (KL >= 2) v = v + dpp_perm<0xB1>(v);
    if constexpr (KL >= 4) v = v + dpp_perm<0x4E>(v);
    if constexpr (KL >= 8) v = v + dpp_perm<0x141>(v);
    if constexpr (KL >= 16) v = v + dpp_perm<0x140>(v);
    return v;
}

// Value of lane `src` (absolute lane index within the wave) of a 32-bit word.
__device__ __forceinline__ uint32_t lane_get(uint32_t v, int src) { return (uint32_t)__shfl((int)v, src, 64); }

}  // namespace ptmc
