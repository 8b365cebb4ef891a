/*
 * ORACLE — test infrastructure only.  Nothing in the shipped library links
 * this file; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg load the oracle, and only as the checker.
 *
 * Counter-based random numbers (Philox, Salmon et al., "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11) and deterministic exp/log built from
 * IEEE-754 basic operations and fma only, so that a device implementation of
 * the same algorithm returns bit-identical results.
 *
 * The reference draws from Julia's task-local Xoshiro stream (`rand()` calls
 * in src/MCRun.jl:145, src/MCMoves.jl:34-38, src/MCRun.jl:123,203,
 * src/Exchange.jl:135,145).  That stream cannot be reproduced without Julia,
 * so both this oracle and the device engine draw from the Philox stream laid
 * out in orc_draw_* below; the order and meaning of the draws follow the
 * reference.
 */
#ifndef ORC_MATH_H
#define ORC_MATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

/* ---------------------------------------------------------------- Philox */

typedef struct { uint32_t v[4]; } orc_u32x4;
typedef struct { uint64_t v[4]; } orc_u64x4;

static inline orc_u32x4 orc_philox4x32_10(orc_u32x4 c, uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c.v[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c.v[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        orc_u32x4 n;
        n.v[0] = hi1 ^ c.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ c.v[3] ^ k1;
        n.v[3] = lo0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

/* 4x64 variant: only used to cross-check the round structure against
 * numpy.random.Philox (which is Philox4x64-10). */
static inline orc_u64x4 orc_philox4x64_10(orc_u64x4 c, uint64_t k0, uint64_t k1)
{
    for (int r = 0; r < 10; ++r) {
        __uint128_t p0 = (__uint128_t)0xD2E7470EE14C6C93ull * c.v[0];
        __uint128_t p1 = (__uint128_t)0xCA5A826395121157ull * c.v[2];
        uint64_t hi0 = (uint64_t)(p0 >> 64), lo0 = (uint64_t)p0;
        uint64_t hi1 = (uint64_t)(p1 >> 64), lo1 = (uint64_t)p1;
        orc_u64x4 n;
        n.v[0] = hi1 ^ c.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ c.v[3] ^ k1;
        n.v[3] = lo0;
        c = n;
        k0 += 0x9E3779B97F4A7C15ull;
        k1 += 0xBB67AE8584CAA73Bull;
    }
    return c;
}

/* Stream layout (shared contract with the device engine):
 *   key     = (seed low 32 bits, seed high 32 bits)
 *   counter = (c0, c1 = sweep, c2 = (type << 28) | (sub << 20) | replica, c3 = run)
 * type 0 = atom/volume move draws, 1 = exchange, 2 = start configurations,
 * 3 = potential parameters.  For moves c0 is the move index inside the sweep
 * and sub = 2*attempt + half (half selects the second block of an FP64
 * move); for exchange c0 = 0 is the attempt block and c0 = 1 + n the block
 * of pair (n, n+1). */
enum { ORC_T_MOVE = 0, ORC_T_EXCH = 1, ORC_T_INIT = 2, ORC_T_PARAM = 3 };

static inline orc_u32x4 orc_draw(uint64_t seed, uint32_t run, uint32_t type,
                                 uint32_t sub, uint32_t replica, uint32_t sweep,
                                 uint32_t c0)
{
    orc_u32x4 c;
    c.v[0] = c0;
    c.v[1] = sweep;
    c.v[2] = (type << 28) | ((sub & 0xffu) << 20) | (replica & 0xfffffu);
    c.v[3] = run;
    return orc_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

/* uniform in [0,1) with 53 random bits */
static inline double orc_u53(uint32_t hi, uint32_t lo)
{
    uint64_t m = ((uint64_t)hi << 21) | (uint64_t)(lo >> 11);
    return (double)m * 0x1.0p-53;
}

/* uniform in [0,1) with 24 random bits (float) */
static inline float orc_u24(uint32_t w) { return (float)(w >> 8) * 0x1.0p-24f; }

/* integer in [0, n) by multiply-high (Lemire) */
static inline uint32_t orc_below(uint32_t w, uint32_t n)
{
    return (uint32_t)(((uint64_t)w * (uint64_t)n) >> 32);
}

/* ------------------------------------------------ deterministic exp / log */

static inline double orc_as_double(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static inline uint64_t orc_as_u64(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

/* exp(x): Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, degree-12 Taylor
 * polynomial evaluated by Horner with fma, scale by 2^k (two steps in the
 * subnormal range).  Relative error below 2 ulp over the normal range. */
static inline double orc_exp(double x)
{
    if (!(x == x)) return x;
    if (x > 709.782712893384) return INFINITY;
    if (x < -745.1332191019412) return 0.0;
    const double log2e = 1.4426950408889634;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    double kd = rint(x * log2e);
    double r = fma(-kd, ln2_hi, x);
    r = fma(-kd, ln2_lo, r);
    double p = 2.08767569878680989792e-09;           /* 1/12! */
    p = fma(p, r, 2.50521083854417187751e-08);       /* 1/11! */
    p = fma(p, r, 2.75573192239858906526e-07);       /* 1/10! */
    p = fma(p, r, 2.75573192239858906526e-06);       /* 1/9!  */
    p = fma(p, r, 2.48015873015873015873e-05);       /* 1/8!  */
    p = fma(p, r, 1.98412698412698412698e-04);       /* 1/7!  */
    p = fma(p, r, 1.38888888888888888889e-03);       /* 1/6!  */
    p = fma(p, r, 8.33333333333333333333e-03);       /* 1/5!  */
    p = fma(p, r, 4.16666666666666666667e-02);       /* 1/4!  */
    p = fma(p, r, 1.66666666666666666667e-01);       /* 1/3!  */
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    int k = (int)kd;
    if (k >= -1021) {
        if (k > 1023) return p * 2.0 * orc_as_double((uint64_t)(1023 + k - 1) << 52);
        return p * orc_as_double((uint64_t)(1023 + k) << 52);
    }
    /* result in (or near) the subnormal range */
    return (p * orc_as_double((uint64_t)(1023 + k + 64) << 52)) * 0x1.0p-64;
}

/* log(x) for x > 0: x = 2^e m, m in [sqrt(1/2), sqrt(2)], log m = 2 atanh(s),
 * s = (m-1)/(m+1), odd series to s^21. */
static inline double orc_log(double x)
{
    if (!(x > 0.0)) return x == 0.0 ? -INFINITY : NAN;
    if (x == INFINITY) return x;
    int eadj = 0;
    if (x < 0x1.0p-1022) { x *= 0x1.0p54; eadj = -54; }
    uint64_t b = orc_as_u64(x);
    int e = (int)((b >> 52) & 0x7ff) - 1023 + eadj;
    double m = orc_as_double((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 1.4142135623730951) { m *= 0.5; e += 1; }
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
    /* log m = 2s + 2s z q  (=  f - f^2/2 ... ) */
    double two_s = 2.0 * s;
    double lm = fma(two_s * z, q, two_s);
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    double ed = (double)e;
    return fma(ed, ln2_hi, fma(ed, ln2_lo, lm));
}

#endif
