// Persistent PT sweep kernel for pair potentials (extended Lennard-Jones).
//
// Layout: one workgroup holds whole temperature ladders of runs that share
// the atom count N.  A configuration ("slot") is owned by a cluster of KL
// lanes; lane q of the cluster keeps atoms j = s*KL + q (s < n_slab) in LDS,
// laid out [xyz][s][lane] so the partner loop reads one bank per lane.
// Exchange swaps temperature labels between slots, not coordinates; the
// per-temperature data (beta, step sizes, counters, energy sums, histograms)
// stay with the temperature exactly as the reference keeps them in its
// MCState vector (src/Exchange.jl:115-124 swaps config, r^2, energy and
// potential variables only).
//
// One cycle (src/MCRun.jl:192-253):
//   every slot makes N (NVT) or N+1 (NPT) moves           src/MCRun.jl:169-182
//     move = draw index, trial, dE, Metropolis, commit      src/MCRun.jl:142-162
//   exchange attempt(s)                                     src/MCRun.jl:203-205
//   step-size adaptation every n_adjust cycles              src/MCRun.jl:206-212
//   sampling every mc_sample cycles (production)            src/MCRun.jl:248-250
#pragma once
#include "ptmc_device.hpp"
#include "ptmc_types.hpp"

namespace ptmc {

enum PairKind : int { PK_EVEN_H = 0, PK_ODD_H = 1, PK_EVEN_R = 2, PK_ODD_R = 3, PK_LJ = 4, PK_ELJB = 5, PK_LUT = 6 };

template <typename Real>
__device__ __forceinline__ Real rcp_(Real x);
template <>
__device__ __forceinline__ float rcp_<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <>
__device__ __forceinline__ double rcp_<double>(double x) { return 1.0 / x; }

template <typename Real>
__device__ __forceinline__ Real fma_(Real a, Real b, Real c);
template <>
__device__ __forceinline__ float fma_<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <>
__device__ __forceinline__ double fma_<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Pair energy e(r^2).  Coefficient array c has KC entries (zero padded:
// trailing zeros leave Horner and the reference loop bit-identical).
template <typename Real, int PK, int KC>
__device__ __forceinline__ Real pair_energy(Real r2, const Real (&c)[KC])
{
    if constexpr (PK == PK_LJ) {
        Real u = rcp_<Real>(r2);
        Real u3 = u * u * u;
        return u3 * fma_<Real>(c[1], u3, c[0]);
    } else if constexpr (PK == PK_EVEN_H) {
        Real u = rcp_<Real>(r2);
        Real q = c[KC - 1];
#pragma unroll
        for (int k = KC - 2; k >= 0; --k) q = fma_<Real>(q, u, c[k]);
        Real u3 = u * u * u;
        return q * u3;
    } else if constexpr (PK == PK_ODD_H) {
        Real v;
        if constexpr (sizeof(Real) == 4) v = __builtin_amdgcn_rsqf(r2);
        else v = 1.0 / __builtin_sqrt(r2);
        Real q = c[KC - 1];
#pragma unroll
        for (int k = KC - 2; k >= 0; --k) q = fma_<Real>(q, v, c[k]);
        Real u = rcp_<Real>(r2);
        Real u3 = u * u * u;
        return q * u3;
    } else if constexpr (PK == PK_EVEN_R) {
        // src/EnergyEvaluation.jl:934-942
        Real r6inv = Real(1) / (r2 * r2 * r2);
        Real sum1 = 0;
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            sum1 += c[k] * r6inv;
            r6inv /= r2;
        }
        return sum1;
    } else {
        // PK_ODD_R, src/EnergyEvaluation.jl:923-932
        Real r = __builtin_sqrt(r2);
        Real r6inv = Real(1) / (r2 * r2 * r2);
        Real sum1 = 0;
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            sum1 += c[k] * r6inv;
            r6inv /= r;
        }
        return sum1;
    }
}

// Pair energy from u = 1/r^2 (binary32 even-power forms).
template <typename Real, int PK, int KC>
__device__ __forceinline__ Real pair_energy_u(Real u, const Real (&c)[KC])
{
    if constexpr (PK == PK_LJ) {
        Real u3 = u * u * u;
        return u3 * fma_<Real>(c[1], u3, c[0]);
    } else {
        Real q = c[KC - 1];
#pragma unroll
        for (int k = KC - 2; k >= 0; --k) q = fma_<Real>(q, u, c[k]);
        Real u3 = u * u * u;
        return q * u3;
    }
}

// FP64 pair energy used by full recomputes and volume moves.
template <int PK, int KC>
__device__ __forceinline__ double pair_energy64(double r2, const double (&c)[KC])
{
    if constexpr (PK == PK_EVEN_R || PK == PK_ODD_R || PK == PK_ODD_H)
        return pair_energy<double, PK, KC>(r2, c);
    else
        return pair_energy<double, PK_EVEN_H, KC>(r2, c);
}

// The same for the volume-move energies of the binary32 paths (the wave pair
// sums): 1/r^2 from the hardware reciprocal and two Newton steps, within an
// ulp of the division and without its scaling and fix-up passes (these
// paths' running totals carry binary32 errors; their FP64 recomputes use
// pair_energy64).
template <int PK, int KC>
__device__ __forceinline__ double pair_energy64_vol(double r2, const double (&c)[KC])
{
    if constexpr (PK == PK_EVEN_R || PK == PK_ODD_R || PK == PK_ODD_H) {
        return pair_energy<double, PK, KC>(r2, c);
    } else {
        double u = __builtin_amdgcn_rcp(r2);
        double e = __builtin_fma(-r2, u, 1.0);
        u = __builtin_fma(u, e, u);
        e = __builtin_fma(-r2, u, 1.0);
        u = __builtin_fma(u, e, u);
        return pair_energy_u<double, PK_EVEN_H, KC>(u, c);
    }
}

// ELJPotentialB pair energy (src/EnergyEvaluation.jl:943-958) from r^2 and the
// pair vector's squared xy extent dxy2 and z extent dz: c_k r^-p_k (1 + a_k
// cos 2t + b_k cos 4t), p = 6, 8, 9, 10, ..., t the angle to z; 0.1 inside
// r^2 < 5.30.  FP64 follows the reference's steps (tan t = |sqrt(dxy2) / dz|,
// cos 2t = 2 / (tan^2 + 1) - 1, the powers by division); binary32 takes
// cos 2t = (dz^2 - dxy2) / r^2 and the powers from one reciprocal and one
// reciprocal square root.  Zero-padded coefficients add exact zeros.
template <typename Real, int KC>
__device__ __forceinline__ Real pair_energy_b(Real r2, Real dxy2, Real dz, const Real (&c)[KC],
                                              const Real (&ca)[KC], const Real (&cb)[KC])
{
    if constexpr (sizeof(Real) == 8) {
        if (!(r2 >= 5.30)) return 0.1;
        const double tz = __builtin_fabs(__builtin_sqrt(dxy2) / dz);
        double r6inv = 1.0 / (r2 * r2 * r2);
        const double t2 = 2.0 / (tz * tz + 1.0) - 1.0;
        const double t4 = 2.0 * (t2 * t2) - 1.0;
        double sum1 = c[0] * r6inv * (1.0 + ca[0] * t2 + cb[0] * t4);
        r6inv /= r2;
#pragma unroll
        for (int i = 1; i < KC; ++i) {
            sum1 += c[i] * r6inv * (1.0 + ca[i] * t2 + cb[i] * t4);
            r6inv /= __builtin_sqrt(r2);
        }
        return sum1;
    } else {
        const float u = __builtin_amdgcn_rcpf(r2);
        const float rs = __builtin_amdgcn_rsqf(r2);
        const float t2 = (dz * dz - dxy2) * u;
        const float t4 = __builtin_fmaf(2.0f * t2, t2, -1.0f);
        float p = u * u * u;
        float sum1 = c[0] * p * (1.0f + ca[0] * t2 + cb[0] * t4);
        p = p * u;
#pragma unroll
        for (int i = 1; i < KC; ++i) {
            sum1 += c[i] * p * (1.0f + ca[i] * t2 + cb[i] * t4);
            p = p * rs;
        }
        return r2 >= 5.30f ? sum1 : 0.1f;
    }
}

// LookupTablePotential pair energy (src/EnergyEvaluation.jl:273-295): the
// entry at the nearest tabulated angle to z (index 1 below tan 0.5 degrees,
// 91 above tan 89.5 degrees) and the nearest tabulated distance, the first
// column inside the first half step, c6 / r^6 beyond the table.  FP64
// follows the reference's steps; binary32 the same with binary32 atan and
// square root and the binary32 copy of the table.  The indices are clamped
// to the table (the engine admits only tables where that changes nothing).
template <typename Real>
__device__ __forceinline__ Real pair_energy_lut(Real r2, Real dxy2, Real dz, const PotParams &P)
{
    if constexpr (sizeof(Real) == 8) {
        const double at = __builtin_fabs(__builtin_sqrt(dxy2) / dz);
        int ai = 1;
        if (at > 0.00872687153 && at <= 114.592845357)
            ai = (int)__builtin_rint(atan(at) / P.lut_dangle * 180.0 / 3.14159265358979323846 + 1.0);
        else if (at > 114.592845357)
            ai = 91;
        ai = ai < 1 ? 1 : (ai > P.lut_nangle ? P.lut_nangle : ai);
        const double lo = P.lut_start + 0.5 * P.lut_d;
        const double hi = P.lut_start + (double)P.lut_ndist * P.lut_d;
        const double *row = P.lut64 + (size_t)(ai - 1) * P.lut_ndist;
        if (r2 <= lo * lo) return row[0];
        if (r2 <= hi * hi) {
            int di = (int)__builtin_rint((__builtin_sqrt(r2) - P.lut_start) / P.lut_d);
            di = di < 1 ? 1 : (di > P.lut_ndist ? P.lut_ndist : di);
            return row[di - 1];
        }
        return P.lut_c6 / (r2 * r2 * r2);
    } else {
        const float at = __builtin_fabsf(__builtin_sqrtf(dxy2) / dz);
        int ai = 1;
        if (at > 0.00872687153f && at <= 114.592845357f)
            ai = (int)__builtin_rintf(atanf(at) * (float)(180.0 / 3.14159265358979323846 / P.lut_dangle) + 1.0f);
        else if (at > 114.592845357f)
            ai = 91;
        ai = ai < 1 ? 1 : (ai > P.lut_nangle ? P.lut_nangle : ai);
        const float lo = (float)(P.lut_start + 0.5 * P.lut_d);
        const float hi = (float)(P.lut_start + (double)P.lut_ndist * P.lut_d);
        const float *row = P.lut32 + (size_t)(ai - 1) * P.lut_ndist;
        if (r2 <= lo * lo) return row[0];
        if (r2 <= hi * hi) {
            int di = (int)__builtin_rintf((__builtin_sqrtf(r2) - (float)P.lut_start) * (float)(1.0 / P.lut_d));
            di = di < 1 ? 1 : (di > P.lut_ndist ? P.lut_ndist : di);
            return row[di - 1];
        }
        const float u = __builtin_amdgcn_rcpf(r2);
        return (float)P.lut_c6 * (u * u * u);
    }
}

// Squared distance between a and b (src/Configurations.jl:101-109).
// Cubic minimum image: FP64 moves b by L*round((a-b)/L) like the reference
// (division for the reference pair form, multiplication by 1/L otherwise);
// binary32 folds the image into the difference.
template <typename Real, int BC, bool DIV>
__device__ __forceinline__ Real dist2(Real ax, Real ay, Real az, Real bx, Real by, Real bz, Real L,
                                      Real invL)
{
    if constexpr (sizeof(Real) == 8) {
        if constexpr (BC == BC_CUBIC) {
            Real qx = DIV ? (ax - bx) / L : (ax - bx) * invL;
            Real qy = DIV ? (ay - by) / L : (ay - by) * invL;
            Real qz = DIV ? (az - bz) / L : (az - bz) * invL;
            bx = bx + L * __builtin_rint(qx);
            by = by + L * __builtin_rint(qy);
            bz = bz + L * __builtin_rint(qz);
        }
        Real dx = ax - bx, dy = ay - by, dz = az - bz;
        return dx * dx + dy * dy + dz * dz;
    } else {
        Real dx = ax - bx, dy = ay - by, dz = az - bz;
        if constexpr (BC == BC_CUBIC) {
            dx = __builtin_fmaf(-L, __builtin_rintf(dx * invL), dx);
            dy = __builtin_fmaf(-L, __builtin_rintf(dy * invL), dy);
            dz = __builtin_fmaf(-L, __builtin_rintf(dz * invL), dz);
        }
        return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    }
}

__device__ __forceinline__ double dist2_64(double ax, double ay, double az, double bx, double by,
                                           double bz, int bc, double L, double invL, bool div)
{
    if (bc == BC_CUBIC) {
        double qx = div ? (ax - bx) / L : (ax - bx) * invL;
        double qy = div ? (ay - by) / L : (ay - by) * invL;
        double qz = div ? (az - bz) / L : (az - bz) * invL;
        bx = bx + L * __builtin_rint(qx);
        by = by + L * __builtin_rint(qy);
        bz = bz + L * __builtin_rint(qz);
    }
    double dx = ax - bx, dy = ay - by, dz = az - bz;
    return dx * dx + dy * dy + dz * dz;
}

// Long-range correction, src/EnergyEvaluation.jl:968-982.
template <int KC>
__device__ __forceinline__ double lrc64(int n_atoms, double r_cut, const double (&c)[KC])
{
    if (r_cut <= 50.0) return 0.0;
    double rcs = __builtin_sqrt(r_cut);
    double rc3 = r_cut * rcs;
    double e = 0.0;
#pragma unroll
    for (int i = 1; i <= KC; ++i) {
        e += c[i - 1] / rc3 / (double)(2 * i + 1);
        rc3 *= r_cut;
    }
    double n2 = (double)((long long)n_atoms * n_atoms);
    e *= 3.14159265358979323846 * n2 / 4.0 / (rcs * rcs * rcs);
    return e;
}

// ---- periodic boxes: cubic (L), rectangular (L, L, H) and the rhombic prism
// (60 degree rhombus of side L, height H).  Minimum image
// src/Configurations.jl:101-126, trial wrap src/MCMoves.jl:51-98, cut-off
// src/Ensembles.jl:118-155, volume src/Configurations.jl:246-256, tail
// correction src/EnergyEvaluation.jl:420-474.
constexpr double SQRT3 = 1.7320508075688772;  // 3^0.5
constexpr double INV_SQRT3 = 1.0 / 1.7320508075688772;

// Box constants in one precision, each rounded from its FP64 value.
template <typename T>
struct BoxC {
    T L, iL, H, iH;  // length, height and reciprocals
    T Ly, iLy;       // rhombic row spacing of the distance rule, (3^0.5 / 2) L
    T W, iW, qo;     // rhombic wrap: L 3^0.5 / 2 and L 3^0.5 / 4
    T hL, hH;        // L / 2, H / 2
};

template <typename T>
__device__ __forceinline__ BoxC<T> make_box(double l, double h)
{
    BoxC<T> b;
    const double ly = SQRT3 / 2.0 * l, w = l * SQRT3 / 2.0;
    b.L = (T)l;
    b.iL = (T)(1.0 / l);
    b.H = (T)h;
    b.iH = (T)(1.0 / h);
    b.Ly = (T)ly;
    b.iLy = (T)(1.0 / ly);
    b.W = (T)w;
    b.iW = (T)(1.0 / w);
    b.qo = (T)(l * SQRT3 / 4.0);
    b.hL = (T)(l / 2.0);
    b.hH = (T)(h / 2.0);
    return b;
}

// Squared distance from a to the image of b nearest it.  FP64 moves b as the
// reference does (DIV: its divisions; else multiplications by the box
// reciprocals); binary32 folds the image into the difference.  The rhombic
// rule is the reference's: y by the row spacing, then x along the row after
// removing the skew of the y difference, z by the height.
template <typename T, int BC, bool DIV>
__device__ __forceinline__ T box_d2(T ax, T ay, T az, T bx, T by, T bz, const BoxC<T> &b)
{
    if constexpr (sizeof(T) == 8) {
        if constexpr (BC == BC_CUBIC) {
            bx = bx + b.L * __builtin_rint(DIV ? (ax - bx) / b.L : (ax - bx) * b.iL);
            by = by + b.L * __builtin_rint(DIV ? (ay - by) / b.L : (ay - by) * b.iL);
            bz = bz + b.H * __builtin_rint(DIV ? (az - bz) / b.H : (az - bz) * b.iH);
        } else if constexpr (BC == BC_RHOMB) {
            const T by2 = by + b.Ly * __builtin_rint(DIV ? (ay - by) / b.Ly : (ay - by) * b.iLy);
            const T sk = (ax - bx) - INV_SQRT3 * (ay - by);
            const T kx = __builtin_rint(DIV ? sk / b.L : sk * b.iL);
            bx = (DIV ? bx - by / SQRT3 : bx - by * INV_SQRT3) + b.L * kx + INV_SQRT3 * by2;
            by = by2;
            bz = bz + b.H * __builtin_rint(DIV ? (az - bz) / b.H : (az - bz) * b.iH);
        }
        const T dx = ax - bx, dy = ay - by, dz = az - bz;
        return dx * dx + dy * dy + dz * dz;
    } else {
        T dx = ax - bx, dy = ay - by, dz = az - bz;
        if constexpr (BC == BC_CUBIC) {
            dx = __builtin_fmaf(-b.L, __builtin_rintf(dx * b.iL), dx);
            dy = __builtin_fmaf(-b.L, __builtin_rintf(dy * b.iL), dy);
            dz = __builtin_fmaf(-b.H, __builtin_rintf(dz * b.iH), dz);
        } else if constexpr (BC == BC_RHOMB) {
            const float k2 = __builtin_rintf(dy * b.iLy);
            const float k1 = __builtin_rintf(__builtin_fmaf(-(float)INV_SQRT3, dy, dx) * b.iL);
            dx = __builtin_fmaf(-b.hL, k2, __builtin_fmaf(-b.L, k1, dx));
            dy = __builtin_fmaf(-b.Ly, k2, dy);
            dz = __builtin_fmaf(-b.H, __builtin_rintf(dz * b.iH), dz);
        }
        return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    }
}

// box_d2 that also returns the squared xy extent and the z extent of the
// pair vector (the angle to z of get_tan, src/Configurations.jl:172-218).
template <typename T, int BC, bool DIV>
__device__ __forceinline__ T box_d2c(T ax, T ay, T az, T bx, T by, T bz, const BoxC<T> &b, T &dxy2, T &dzo)
{
    if constexpr (sizeof(T) == 8) {
        if constexpr (BC == BC_CUBIC) {
            bx = bx + b.L * __builtin_rint(DIV ? (ax - bx) / b.L : (ax - bx) * b.iL);
            by = by + b.L * __builtin_rint(DIV ? (ay - by) / b.L : (ay - by) * b.iL);
            bz = bz + b.H * __builtin_rint(DIV ? (az - bz) / b.H : (az - bz) * b.iH);
        } else if constexpr (BC == BC_RHOMB) {
            const T by2 = by + b.Ly * __builtin_rint(DIV ? (ay - by) / b.Ly : (ay - by) * b.iLy);
            const T sk = (ax - bx) - INV_SQRT3 * (ay - by);
            const T kx = __builtin_rint(DIV ? sk / b.L : sk * b.iL);
            bx = (DIV ? bx - by / SQRT3 : bx - by * INV_SQRT3) + b.L * kx + INV_SQRT3 * by2;
            by = by2;
            bz = bz + b.H * __builtin_rint(DIV ? (az - bz) / b.H : (az - bz) * b.iH);
        }
        const T dx = ax - bx, dy = ay - by, dz = az - bz;
        dxy2 = dx * dx + dy * dy;
        dzo = dz;
        return dxy2 + dz * dz;
    } else {
        T dx = ax - bx, dy = ay - by, dz = az - bz;
        if constexpr (BC == BC_CUBIC) {
            dx = __builtin_fmaf(-b.L, __builtin_rintf(dx * b.iL), dx);
            dy = __builtin_fmaf(-b.L, __builtin_rintf(dy * b.iL), dy);
            dz = __builtin_fmaf(-b.H, __builtin_rintf(dz * b.iH), dz);
        } else if constexpr (BC == BC_RHOMB) {
            const float k2 = __builtin_rintf(dy * b.iLy);
            const float k1 = __builtin_rintf(__builtin_fmaf(-(float)INV_SQRT3, dy, dx) * b.iL);
            dx = __builtin_fmaf(-b.hL, k2, __builtin_fmaf(-b.L, k1, dx));
            dy = __builtin_fmaf(-b.Ly, k2, dy);
            dz = __builtin_fmaf(-b.H, __builtin_rintf(dz * b.iH), dz);
        }
        dxy2 = __builtin_fmaf(dy, dy, dx * dx);
        dzo = dz;
        return __builtin_fmaf(dz, dz, dxy2);
    }
}

// r_cut^2 of a box (src/Ensembles.jl:118-155); kind BC_CUBIC, BC_RECT or BC_RHOMB
__device__ __forceinline__ double box_rcut(int kind, double l, double h)
{
    if (kind == BC_RECT) return __builtin_fmin(l * l / 4.0, h * h / 4.0);
    if (kind == BC_RHOMB) return __builtin_fmin(l * l * 3.0 / 16.0, h * h / 4.0);
    return l * l / 4.0;
}

// get_volume (src/Configurations.jl:246-256)
__device__ __forceinline__ double box_volume(int kind, double l, double h)
{
    if (kind == BC_RECT) return l * l * h;
    if (kind == BC_RHOMB) return l * l * h * SQRT3 / 2.0;
    return l * l * l;
}

// the volume the NPT histograms bin (find_hist_index, src/Sampling.jl:72-125;
// the rhombic form there evaluates L^2 * 3 H^0.5 / 2)
__device__ __forceinline__ double box_hist_volume(int kind, double l, double h)
{
    if (kind == BC_RECT) return l * l * h;
    if (kind == BC_RHOMB) return l * l * (3.0 * __builtin_sqrt(h)) / 2.0;
    return l * l * l;
}

// lrc(ELJPotentialB), src/EnergyEvaluation.jl:983-998: four fixed even-power
// coefficients, nothing inside r_cut <= 16
__device__ __forceinline__ double lrc_b64(int n_atoms, double r_cut)
{
    const double cb[4] = {-0.1279111890228638, -1.328138539967966, 12.260941135261255, 41.12212408251662};
    if (r_cut <= 16.0) return 0.0;
    double rcs = __builtin_sqrt(r_cut);
    double rc3 = r_cut * rcs;
    double e = 0.0;
#pragma unroll
    for (int i = 1; i <= 4; ++i) {
        e += cb[i - 1] / rc3 / (double)(2 * i + 1);
        rc3 *= r_cut;
    }
    double n2 = (double)((long long)n_atoms * n_atoms);
    e *= 3.14159265358979323846 * n2 / 4.0 / (rcs * rcs * rcs);
    return e;
}

// tail correction of a box (src/EnergyEvaluation.jl:420, :447-448, :469-473)
template <int PK, int KC>
__device__ __forceinline__ double box_lrc(int kind, int n_atoms, double r_cut, double l, double h,
                                          const double (&c)[KC])
{
    double e;
    if constexpr (PK == PK_ELJB) {
        e = lrc_b64(n_atoms, r_cut);
    } else if constexpr (PK == PK_LUT) {
        // lrc(LookupTablePotential), src/EnergyEvaluation.jl:999-1008 (1.0 for r_cut <= 10); c[0] holds c6
        if (r_cut <= 10.0) {
            e = 1.0;
        } else {
            const double rcs = __builtin_sqrt(r_cut);
            const double rc3 = r_cut * rcs;
            e = c[0] / rc3 / 3.0;
            e *= 3.14159265358979323846 * (double)((long long)n_atoms * n_atoms) / 4.0 / (rcs * rcs * rcs);
        }
    } else {
        e = lrc64<KC>(n_atoms, r_cut, c);
    }
    if (kind == BC_RHOMB) return e * 3.0 / 4.0 * l / h;
    if (kind == BC_RECT) return l < h ? e * l / h : e * (h * h) / (l * l);
    return e;
}

// binary32 squared distances of a partner x_j to the old position o of the
// moved atom and to its trial o + d: r2o = |v|^2 with v = x_j - o, and
// r2n = |v - d|^2 = r2o + |d|^2 - 2 v.d (m2d = -2 d, dd = |d|^2 per move):
// ten operations for the pair of distances instead of twelve.  |d| is at
// most the step size, so r2n is never small against r2o and the rounding
// stays a few binary32 ulps of r2n.
__device__ __forceinline__ void dist2_pair_f32(float xj, float yj, float zj, float ox, float oy, float oz,
                                               float m2dx, float m2dy, float m2dz, float dd, float &r2n, float &r2o)
{
    const float vx = xj - ox, vy = yj - oy, vz = zj - oz;
    r2o = __builtin_fmaf(vz, vz, __builtin_fmaf(vy, vy, vx * vx));
    r2n = __builtin_fmaf(m2dz, vz, __builtin_fmaf(m2dy, vy, __builtin_fmaf(m2dx, vx, r2o + dd)));
}

template <typename Real>
struct MoveDraw {
    uint32_t index;
    Real ux, uy, uz;
    Real umet;
};

// Decode the words of one move attempt.  FP64 spends two Philox blocks per
// attempt (53-bit uniforms, the atom index assembled from their spare low
// bits); binary32 one block (24-bit uniforms, the Metropolis uniform built
// from the three spare low bytes).
template <typename Real>
__device__ __forceinline__ MoveDraw<Real> decode_move(const u32x4 &A, const u32x4 &B, uint32_t nsteps)
{
    MoveDraw<Real> d;
    if constexpr (sizeof(Real) == 8) {
        d.ux = u53(A.x, A.y);
        d.uy = u53(A.z, A.w);
        d.uz = u53(B.x, B.y);
        d.umet = u53(B.z, B.w);
        const uint32_t w = (A.y & 0x7ffu) | ((A.w & 0x7ffu) << 11) | ((B.y & 0x3ffu) << 22);
        d.index = below(w, nsteps);
    } else {
        (void)B;
        d.index = below(A.x, nsteps);
        d.ux = u24(A.y);
        d.uy = u24(A.z);
        d.uz = u24(A.w);
        const uint32_t w = ((A.y & 0xffu) << 16) | ((A.z & 0xffu) << 8) | (A.w & 0xffu);
        d.umet = (float)w * 0x1.0p-24f;
    }
    return d;
}

template <typename Real>
__device__ __forceinline__ MoveDraw<Real> move_draw(uint64_t seed, uint32_t run, uint32_t temp, uint32_t sweep,
                                                    uint32_t m, uint32_t attempt, uint32_t nsteps)
{
    if constexpr (sizeof(Real) == 8) {
        const u32x4 A = draw(seed, run, RNG_MOVE, 2u * attempt, temp, sweep, m);
        const u32x4 B = draw(seed, run, RNG_MOVE, 2u * attempt + 1u, temp, sweep, m);
        return decode_move<Real>(A, B, nsteps);
    } else {
        const u32x4 A = draw(seed, run, RNG_MOVE, attempt, temp, sweep, m);
        return decode_move<Real>(A, A, nsteps);
    }
}

template <typename Real, int BC, bool DIV>
__device__ __forceinline__ void make_trial(Real xi, Real yi, Real zi, const MoveDraw<Real> &d, Real delta,
                                           Real L, Real invL, Real &tx, Real &ty, Real &tz)
{
    if constexpr (sizeof(Real) == 8) {
        tx = xi + (d.ux - 0.5) * delta;
        ty = yi + (d.uy - 0.5) * delta;
        tz = zi + (d.uz - 0.5) * delta;
        if constexpr (BC == BC_CUBIC) {
            tx = tx - L * __builtin_rint(DIV ? tx / L : tx * invL);
            ty = ty - L * __builtin_rint(DIV ? ty / L : ty * invL);
            tz = tz - L * __builtin_rint(DIV ? tz / L : tz * invL);
        }
    } else {
        tx = xi + (d.ux - 0.5f) * delta;
        ty = yi + (d.uy - 0.5f) * delta;
        tz = zi + (d.uz - 0.5f) * delta;
        if constexpr (BC == BC_CUBIC) {
            float q;
            q = __builtin_rintf(tx * invL);
            tx = tx - L * q;
            q = __builtin_rintf(ty * invL);
            ty = ty - L * q;
            q = __builtin_rintf(tz * invL);
            tz = tz - L * q;
        }
    }
}

// Trial position in a periodic box (src/MCMoves.jl:51-98): displacement,
// then the wrap back into the box; binary32 repeats the FP64 steps with the
// box constants rounded to binary32.
template <typename Real, int BC, bool DIV>
__device__ __forceinline__ void box_trial(Real xi, Real yi, Real zi, const MoveDraw<Real> &d, Real delta,
                                          const BoxC<Real> &b, Real &tx, Real &ty, Real &tz)
{
    const Real half = (Real)0.5;
    tx = xi + (d.ux - half) * delta;
    ty = yi + (d.uy - half) * delta;
    tz = zi + (d.uz - half) * delta;
    if constexpr (sizeof(Real) == 8) {
        if constexpr (BC == BC_CUBIC) {
            tx = tx - b.L * __builtin_rint(DIV ? tx / b.L : tx * b.iL);
            ty = ty - b.L * __builtin_rint(DIV ? ty / b.L : ty * b.iL);
            tz = tz - b.H * __builtin_rint(DIV ? tz / b.H : tz * b.iH);
        } else if constexpr (BC == BC_RHOMB) {
            const Real k2 = __builtin_rint(DIV ? (ty - b.qo) / b.W : (ty - b.qo) * b.iW);
            const Real k1 = __builtin_rint(DIV ? (tx - ty / SQRT3 - b.hL) / b.L : (tx - ty * INV_SQRT3 - b.hL) * b.iL);
            const Real k3 = __builtin_rint(DIV ? (tz - b.hH) / b.H : (tz - b.hH) * b.iH);
            const Real d0 = b.L * k1 + b.hL * k2;
            tx = tx - d0;
            ty = ty - b.W * k2;
            tz = tz - b.H * k3;
        }
    } else {
        if constexpr (BC == BC_CUBIC) {
            float q;
            q = __builtin_rintf(tx * b.iL);
            tx = tx - b.L * q;
            q = __builtin_rintf(ty * b.iL);
            ty = ty - b.L * q;
            q = __builtin_rintf(tz * b.iH);
            tz = tz - b.H * q;
        } else if constexpr (BC == BC_RHOMB) {
            const float k2 = __builtin_rintf((ty - b.qo) * b.iW);
            const float k1 = __builtin_rintf((tx - ty * (float)INV_SQRT3 - b.hL) * b.iL);
            const float k3 = __builtin_rintf((tz - b.hH) * b.iH);
            const float d0 = b.L * k1 + b.hL * k2;
            tx = tx - d0;
            ty = ty - b.W * k2;
            tz = tz - b.H * k3;
        }
    }
}

template <typename Real>
__device__ __forceinline__ bool outside_sphere(Real x, Real y, Real z, Real r2)
{
    if constexpr (sizeof(Real) == 8) {
        return x * x + y * y + z * z > r2;
    } else {
        float s = x * x + y * y;
        s = s + z * z;
        return s > r2;
    }
}

// LDS / global layout of a tile: element (slab s, component c, lane) at
// (s * 3 + c) * LANES + lane.  The compile-time lane stride lets one base
// register and immediate offsets address x, y, z of consecutive slabs.
template <int LANES>
__device__ __forceinline__ int xoff(int s, int c, int lane) { return (s * 3 + c) * LANES + lane; }

// Parking place of padding atoms (j >= N) and of the moved atom during the
// partner loop in a spherical container: every pair energy with it is
// exactly zero in both precisions (1/r^6 underflows), so no per-pair mask
// is needed.
template <typename Real>
__device__ __forceinline__ Real far_point();
template <>
__device__ __forceinline__ float far_point<float>() { return 1.0e18f; }
template <>
__device__ __forceinline__ double far_point<double>() { return 1.0e100; }

// Trial moves leaving the spherical container are redrawn (src/MCMoves.jl:115-135):
// at the 50th failure the configuration is recentred (sequential mean,
// src/Configurations.jl:79-85) without a new draw, the 100th failure is an
// error.  Returns false on that error.  On return xi/yi/zi hold the (possibly
// recentred) old position of the moved atom and tx/ty/tz the trial.
template <typename Real, int KL, int LANES>
__device__ bool sphere_redraw(Real *X, int tid, int q, int N, int A, int si, int li, Real radius2, Real delta_r,
                              uint64_t seed, uint32_t run_key, uint32_t temp, uint32_t sweep, uint32_t m,
                              uint32_t nsteps, Real &xi, Real &yi, Real &zi, Real &tx, Real &ty, Real &tz)
{
    int count = 0;
    uint32_t attempt = 0;
    bool ok = true;
    while (outside_sphere<Real>(tx, ty, tz, radius2)) {
        count += 1;
        if (count == 50) {
            // every lane of the cluster forms the same sequential mean, atom by atom
            const int gb = tid - q;
            double sx = 0.0, sy = 0.0, sz = 0.0;
            for (int j = 0; j < N; ++j) {
                const int sj = j / KL, lj = gb + (j % KL);
                sx += (double)X[xoff<LANES>(sj, 0, lj)];
                sy += (double)X[xoff<LANES>(sj, 1, lj)];
                sz += (double)X[xoff<LANES>(sj, 2, lj)];
            }
            sx /= (double)N;
            sy /= (double)N;
            sz /= (double)N;
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            for (int s = 0; s < A; ++s) {
                if (s * KL + q >= N) break;
                X[xoff<LANES>(s, 0, tid)] = (Real)((double)X[xoff<LANES>(s, 0, tid)] - sx);
                X[xoff<LANES>(s, 1, tid)] = (Real)((double)X[xoff<LANES>(s, 1, tid)] - sy);
                X[xoff<LANES>(s, 2, tid)] = (Real)((double)X[xoff<LANES>(s, 2, tid)] - sz);
            }
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
        } else {
            attempt += 1;
            MoveDraw<Real> dd = move_draw<Real>(seed, run_key, temp, sweep, m, attempt, nsteps);
            xi = X[xoff<LANES>(si, 0, li)];
            yi = X[xoff<LANES>(si, 1, li)];
            zi = X[xoff<LANES>(si, 2, li)];
            make_trial<Real, BC_SPHERE, false>(xi, yi, zi, dd, delta_r, Real(0), Real(0), tx, ty, tz);
            if (count == 100) {
                ok = false;
                break;
            }
        }
    }
    xi = X[xoff<LANES>(si, 0, li)];
    yi = X[xoff<LANES>(si, 1, li)];
    zi = X[xoff<LANES>(si, 2, li)];
    return ok;
}

// FP64 pair-energy sum of a slot's configuration (src/EnergyEvaluation.jl:378-474):
// rows j of this lane, partners l > j, summed by the cluster butterfly.
// Positions enter scaled by (px, px, pz) and the squared distances, taken in
// box bx, are multiplied by m2: (1, 1, 1) for the configuration as it is,
// (1, 1, s^2) for a uniformly scaled trial (the reference scales its r^2
// matrix), (sx, sz, 1) with the trial box for a separated volume move (the
// reference recomputes r^2 from the scaled positions).  NPT keeps pairs with
// r^2 <= rc; the caller adds the tail correction.
template <typename Real, int PK, int KC, int BC, int ENS, int KL, int LANES>
__device__ double pair_sum64(const Real *X, int gb, int q, int N, const BoxC<double> &bx, double px, double pz,
                             double m2, double rc, const double (&c64)[KC], const double (&ca64)[KC],
                             const double (&cb64)[KC], const PotParams &pp)
{
    constexpr bool DIV = (PK == PK_EVEN_R || PK == PK_ODD_R);
    const int A = (N + KL - 1) / KL;
    double part = 0.0;
    for (int s = 0; s < A; ++s) {
        const int j = s * KL + q;
        if (j >= N) break;
        const double xj = (double)X[xoff<LANES>(s, 0, gb + q)] * px;
        const double yj = (double)X[xoff<LANES>(s, 1, gb + q)] * px;
        const double zj = (double)X[xoff<LANES>(s, 2, gb + q)] * pz;
        for (int l = j + 1; l < N; ++l) {
            const int sl = l / KL, ll = gb + (l % KL);
            const double xl = (double)X[xoff<LANES>(sl, 0, ll)] * px;
            const double yl = (double)X[xoff<LANES>(sl, 1, ll)] * px;
            const double zl = (double)X[xoff<LANES>(sl, 2, ll)] * pz;
            if constexpr (PK == PK_ELJB || PK == PK_LUT) {
                // the angle to z of the unscaled pair vector (a uniform scaling keeps it)
                double dxy2, dz;
                double r2 = box_d2c<double, BC, DIV>(xj, yj, zj, xl, yl, zl, bx, dxy2, dz);
                r2 = r2 * m2;
                if (ENS != ENS_NPT || r2 <= rc) {
                    if constexpr (PK == PK_ELJB) part += pair_energy_b<double, KC>(r2, dxy2, dz, c64, ca64, cb64);
                    else part += pair_energy_lut<double>(r2, dxy2, dz, pp);
                }
            } else {
                double r2 = box_d2<double, BC, DIV>(xj, yj, zj, xl, yl, zl, bx);
                r2 = r2 * m2;
                if (ENS == ENS_NPT) {
                    if (r2 <= rc) part += pair_energy64<PK, KC>(r2, c64);
                } else {
                    part += pair_energy64<PK, KC>(r2, c64);
                }
            }
        }
    }
    return group_sum<KL>(part);
}

__device__ __forceinline__ double readlane_d(double v, int src)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// The pair walk of the wave sums: the N(N-1)/2 pairs as (j, j + d mod N)
// for d = 1..N/2 (at d = N/2 of an even N only j < N/2), pair slot
// p = (d - 1) N + j dealt round-robin over the nl live lanes (this lane's
// rank first).  Four slots per step, each into its own partial sum, so the
// four pair chains (positions, minimum image, division, polynomial) overlap
// instead of running one after another; the partial sums are added as
// (s0 + s1) + (s2 + s3).  `pos(j, x, y, z)` reads atom j, `energy` the pair
// energy (0 beyond the cut-off).
template <typename Pos, typename En>
__device__ __forceinline__ double wave_pair_walk(int N, int rank, int nl, Pos &&pos, En &&energy)
{
    const int h = N / 2;
    const bool even = (N & 1) == 0;
    const int P = h * N;
    const int step = 4 * nl;
    // slot advance of a step as (d, j) increments
    const int dj = step % N, dd = step / N;
    int j[4], d[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int p = rank + u * nl;
        j[u] = p % N;
        d[u] = p / N + 1;
    }
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int p0 = rank; p0 < P; p0 += step) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int ju = j[u], du = d[u];
            if (p0 + u * nl < P && !(even && du == h && ju >= h)) {
                int l = ju + du;
                if (l >= N) l -= N;
                double xj, yj, zj, xl, yl, zl;
                pos(ju, xj, yj, zj);
                pos(l, xl, yl, zl);
                s[u] += energy(xj, yj, zj, xl, yl, zl);
            }
            int jn = ju + dj, dn = du + dd;
            if (jn >= N) {
                jn -= N;
                dn += 1;
            }
            j[u] = jn;
            d[u] = dn;
        }
    }
    return (s[0] + s[1]) + (s[2] + s[3]);
}

// The pair sum of pair_sum64 (NPT form, cut-off rc) for the configuration
// held by the cluster whose first lane is wave lane `cl`, computed by every
// executing lane of the wave together (volume-move energies of the binary32
// path).  The N(N-1)/2 pairs are enumerated as (j, j + d mod N) for
// d = 1..N/2 (at d = N/2 of an even N only j < N/2) and dealt round-robin
// over the `live` lanes; cluster butterflies, then the cluster totals of the
// `leaders` (first lanes) in lane order.  Every lane returns the same value.
// Pairs and cut-off test are those of pair_sum64; only the summation order
// differs (wave_pair_walk).
template <int PK, int KC, int BC, int KL, int LANES, typename Real>
__device__ double wave_pair_sum64(const Real *X, int cl, int N, const BoxC<double> &bx, double px, double pz,
                                  double m2, double rc, const double (&c64)[KC], const double (&ca64)[KC],
                                  const double (&cb64)[KC], const PotParams &pp, uint64_t live, uint64_t leaders)
{
    constexpr bool DIV = (PK == PK_EVEN_R || PK == PK_ODD_R);
    const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(live >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)live, 0u));
    const int nl = __popcll(live);
    const int cb = (threadIdx.x & ~63) + cl;  // block lane of the cluster's first atom column
    auto pos = [&](int j, double &x, double &y, double &z) {
        const int sj = j / KL, lj = cb + (j % KL);
        x = (double)X[xoff<LANES>(sj, 0, lj)] * px;
        y = (double)X[xoff<LANES>(sj, 1, lj)] * px;
        z = (double)X[xoff<LANES>(sj, 2, lj)] * pz;
    };
    auto energy = [&](double xj, double yj, double zj, double xl, double yl, double zl) -> double {
        if constexpr (PK == PK_ELJB || PK == PK_LUT) {
            double dxy2, dz;
            const double r2 = box_d2c<double, BC, DIV>(xj, yj, zj, xl, yl, zl, bx, dxy2, dz) * m2;
            if (r2 > rc) return 0.0;
            if constexpr (PK == PK_ELJB) return pair_energy_b<double, KC>(r2, dxy2, dz, c64, ca64, cb64);
            else return pair_energy_lut<double>(r2, dxy2, dz, pp);
        } else {
            const double r2 = box_d2<double, BC, DIV>(xj, yj, zj, xl, yl, zl, bx) * m2;
            return r2 <= rc ? pair_energy64_vol<PK, KC>(r2, c64) : 0.0;
        }
    };
    double part = wave_pair_walk(N, rank, nl, pos, energy);
    part = group_sum<KL>(part);
    double tot = 0.0;
    for (uint64_t b = leaders; b != 0; b &= b - 1) tot += readlane_d(part, __builtin_ctzll(b));
    return tot;
}

// Philox words of move mm for this lane (block A; block B only in FP64).
template <typename Real>
__device__ __forceinline__ void move_words(uint64_t seed, uint32_t run, uint32_t temp, uint32_t sweep, uint32_t mm,
                                           u32x4 &A, u32x4 &B)
{
    A = draw(seed, run, RNG_MOVE, 0u, temp, sweep, mm);
    if constexpr (sizeof(Real) == 8) B = draw(seed, run, RNG_MOVE, 1u, temp, sweep, mm);
}

// FP64 Metropolis threshold of a move (src/Exchange.jl:53-57, src/MCRun.jl:
// 122-132): min(1, e^{-beta dE}) >= u  <=>  dE <= -ln(u) / beta (u < 1, beta
// > 0; u = 0 gives +inf, always accepted).  The lane that draws a move's
// words also forms its threshold, so the move itself only compares.
__device__ __forceinline__ double metro_thr(double umet, double beta) { return -det_log(umet) / beta; }

// value of lane `src` of a double
__device__ __forceinline__ double lane_get_d(double v, int src)
{
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = lane_get((uint32_t)b, src), hi = lane_get((uint32_t)(b >> 32), src);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

template <typename Real>
__device__ __forceinline__ void fetch_words(const u32x4 &mA, const u32x4 &mB, int src, u32x4 &A, u32x4 &B)
{
    A.x = lane_get(mA.x, src);
    A.y = lane_get(mA.y, src);
    A.z = lane_get(mA.z, src);
    A.w = lane_get(mA.w, src);
    if constexpr (sizeof(Real) == 8) {
        B.x = lane_get(mB.x, src);
        B.y = lane_get(mB.y, src);
        B.z = lane_get(mB.z, src);
        B.w = lane_get(mB.w, src);
    }
}

template <typename Real, int PK, int KC, int BC, int ENS, int KL, int LANES>
__global__ void __launch_bounds__(LANES) sweep_kernel(SweepArgs a)
{
    constexpr bool DIV = (PK == PK_EVEN_R || PK == PK_ODD_R);
    constexpr bool F32 = sizeof(Real) == 4;
    constexpr int NSLOT = LANES / KL;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const TileInfo tile = a.tiles[blockIdx.x];
    const int tid = threadIdx.x;
    const int lg = tid / KL;
    const int q = tid % KL;
    const int gb = lg * KL;
    const bool active = lg < tile.n_slots;
    const int N = tile.n_atoms;
    const int A = tile.n_slab;
    const uint32_t nsteps = (uint32_t)(ENS == ENS_NPT ? N + 1 : N);

    Real *X = reinterpret_cast<Real *>(lds_raw);
    double *exE = reinterpret_cast<double *>(lds_raw + sizeof(Real) * 3 * A * LANES);
    double *exV = exE + NSLOT;
    int *exT = reinterpret_cast<int *>(exV + NSLOT);
    int *exS = exT + NSLOT;
    // FP64: per-temperature state of the tile's ladders for the launch,
    // indexed by the local temperature slot lb + t (each temperature belongs
    // to one tile): read at the start, updated here, written back at the end,
    // so the per-cycle bookkeeping waits on no global read-modify-write (the
    // latency-bound single ladders).  Binary32 tiles, whose LDS is the tight
    // resource, keep it in global memory.
    constexpr bool RS = !F32;
    long long *rsC = reinterpret_cast<long long *>(exS + NSLOT);  // atom moves (total, window), exchanges (attempted, accepted)
    double *rsH = reinterpret_cast<double *>(rsC + 4 * NSLOT);    // sums of E (NPT: H) and its square
    double *rsD = rsH + 2 * NSLOT;                                 // atom step size
    double *rsB = rsD + NSLOT;                                     // beta
    // k: 0, 1 atom moves (total, window), 2, 3 exchanges (attempted, accepted)
    auto count = [&](int lt, int T, int k) -> long long & {
        if constexpr (RS) return rsC[4 * lt + k];
        else return k < 2 ? a.t_count_atom[2 * T + k] : a.t_count_exc[2 * T + k - 2];
    };
    auto ham = [&](int lt, int T, int k) -> double & {
        if constexpr (RS) return rsH[2 * lt + k];
        else return a.t_ham[2 * T + k];
    };
    auto displ = [&](int lt, int T) -> double & {
        if constexpr (RS) return rsD[lt];
        else return a.t_displ[2 * T];
    };
    auto beta_of = [&](int lt, int T) -> double {
        if constexpr (RS) return rsB[lt];
        else return a.t_beta[T];
    };

    // coefficients: Real copy for the move path, FP64 copy for totals
    Real c[KC];
    double c64[KC], ca64[KC], cb64[KC];
    Real ca[KC], cb[KC];  // ELJPotentialB anisotropy coefficients
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        double v = k < a.pot.ncoeff ? a.pot.coeff[k] : 0.0;
        c64[k] = v;
        c[k] = (Real)v;
        const double va = (PK == PK_ELJB && k < a.pot.ncoeff && k < 8) ? a.pot.coeff_a[k] : 0.0;
        const double vb = (PK == PK_ELJB && k < a.pot.ncoeff && k < 8) ? a.pot.coeff_b[k] : 0.0;
        ca64[k] = va;
        cb64[k] = vb;
        ca[k] = (Real)va;
        cb[k] = (Real)vb;
    }
    if constexpr (PK == PK_LJ) {
        c[0] = (Real)a.pot.coeff[0];
        c[1] = (Real)a.pot.coeff[3];
    }

    Real *gX = reinterpret_cast<Real *>(a.coords) + tile.coord_off;
    for (int e = tid; e < 3 * A * LANES; e += LANES) X[e] = gX[e];

    const int slot = tile.slot0 + lg;
    int t = 0, R = 1, tb = 0, lb = 0;
    uint32_t run_key = 0;
    constexpr bool PBC = BC != BC_SPHERE;
    // box kind of the geometry formulas: the cubic kernels serve cubic and rectangular boxes
    const int kind = BC == BC_RHOMB ? (int)BC_RHOMB : (BC == BC_CUBIC ? a.box_kind : (int)BC_SPHERE);
    double E = 0.0, box = 1.0, boxh = 1.0;
    Real radius2 = 0;  // container radius^2 in the move precision
    if (active) {
        const RunInfo &ri = a.runs[a.slot_run[slot]];
        R = ri.n_temps;
        tb = ri.temp_base;
        lb = ri.local_base;
        run_key = ri.run_key;
        radius2 = (Real)ri.radius2;
        t = a.slot_temp[slot];
        E = a.slot_en[slot];
        if constexpr (PBC) {
            box = a.slot_box[slot];
            boxh = a.slot_boxh[slot];
        }
        if (q == 0) {
            exT[lg] = t;
            exS[lb + t] = lg;
            const int lt = lb + t, T0 = tb + t;
            if constexpr (RS) {
            rsC[4 * lt] = a.t_count_atom[2 * T0];
            rsC[4 * lt + 1] = a.t_count_atom[2 * T0 + 1];
            rsC[4 * lt + 2] = a.t_count_exc[2 * T0];
            rsC[4 * lt + 3] = a.t_count_exc[2 * T0 + 1];
            rsH[2 * lt] = a.t_ham[2 * T0];
            rsH[2 * lt + 1] = a.t_ham[2 * T0 + 1];
            rsD[lt] = a.t_displ[2 * T0];
            rsB[lt] = a.t_beta[T0];
            }
        }
    }
    // the run's histogram bounds and pressure for the sampling
    double en_min = 0.0, delta_en = 1.0, v_min = 0.0, delta_v = 1.0, pressure = 0.0;
    if (active) {
        const RunInfo &ri = a.runs[a.slot_run[slot]];
        en_min = ri.en_min;
        delta_en = ri.delta_en;
        v_min = ri.v_min;
        delta_v = ri.delta_v;
        pressure = ri.pressure;
    }
    __syncthreads();

    for (int cyc = 0; cyc < a.n_cycles; ++cyc) {
        const uint32_t sweep = (uint32_t)(a.sweep0 + cyc);
        const long long pidx = a.cycle0 + cyc + 1;
        if (active) {
            const double beta = beta_of(lb + t, tb + t);
            const double delta = displ(lb + t, tb + t);
            const Real delta_r = (Real)delta;
            const float beta_l2e = (float)beta * 1.4426950408889634f;
            BoxC<Real> bxr{};
            Real rc = 0;
            if constexpr (PBC) {
                bxr = make_box<Real>(box, boxh);
                rc = (Real)box_rcut(kind, box, boxh);
            }
            int nacc = 0, nvacc = 0, nvxy = 0, nvz = 0;
            // Shared Philox words: lane q of the cluster computes the block(s)
            // of move m0 + q; the words of the next move are fetched while
            // the current move runs.
            u32x4 mA = {0, 0, 0, 0}, mB = {0, 0, 0, 0};
            u32x4 nA = {0, 0, 0, 0}, nB = {0, 0, 0, 0};
            double mthr = 0.0, nthr = 0.0;  // FP64: Metropolis thresholds of the lane's and the next move
            // FP64 in the spherical container: the lane that draws a move's words
            // also decodes them (LDS offset and lane of the atom, displacement
            // (u - 1/2) delta), so the cluster fetches ready values instead of
            // every lane decoding every move on the move's serial chain
            constexpr bool PRE = !F32 && BC == BC_SPHERE && KL > 1;
            int m_off = 0, m_li = 0, n_off = 0, n_li = 0;
            double m_dx = 0.0, m_dy = 0.0, m_dz = 0.0, n_dx = 0.0, n_dy = 0.0, n_dz = 0.0;
            auto pre_decode = [&]() {
                const MoveDraw<Real> dw = decode_move<Real>(mA, mB, nsteps);
                m_li = gb + (int)(dw.index % KL);
                m_off = xoff<LANES>((int)(dw.index / KL), 0, m_li);
                m_dx = (dw.ux - 0.5) * delta_r;
                m_dy = (dw.uy - 0.5) * delta_r;
                m_dz = (dw.uz - 0.5) * delta_r;
            };
            auto pre_fetch = [&](int src) {
                n_off = (int)lane_get((uint32_t)m_off, src);
                n_li = (int)lane_get((uint32_t)m_li, src);
                n_dx = lane_get_d(m_dx, src);
                n_dy = lane_get_d(m_dy, src);
                n_dz = lane_get_d(m_dz, src);
            };
            if constexpr (KL > 1) {
                if ((uint32_t)q < nsteps) {
                    move_words<Real>(a.seed, run_key, (uint32_t)t, sweep, (uint32_t)q, mA, mB);
                    if constexpr (!F32) mthr = metro_thr(u53(mB.z, mB.w), beta);
                    if constexpr (PRE) pre_decode();
                }
                if constexpr (PRE) pre_fetch(gb);
                else fetch_words<Real>(mA, mB, gb, nA, nB);
                if constexpr (!F32) nthr = lane_get_d(mthr, gb);
            }
            for (uint32_t m = 0; m < nsteps; ++m) {
                // ---- draws of this move (src/MCRun.jl:145, src/MCMoves.jl:34-38, src/MCRun.jl:123)
                MoveDraw<Real> d0{};
                double thr;
                int xo = 0, li = 0;  // LDS offset of the moved atom's x, the lane holding it
                double pdx = 0.0, pdy = 0.0, pdz = 0.0;
                if constexpr (KL == 1) {
                    d0 = move_draw<Real>(a.seed, run_key, (uint32_t)t, sweep, m, 0u, nsteps);
                    thr = F32 ? 0.0 : metro_thr((double)d0.umet, beta);
                } else {
                    if constexpr (PRE) {
                        xo = n_off;
                        li = n_li;
                        pdx = n_dx;
                        pdy = n_dy;
                        pdz = n_dz;
                    } else {
                        d0 = decode_move<Real>(nA, nB, nsteps);
                    }
                    thr = nthr;
                    const uint32_t m1 = m + 1;
                    if (m1 < nsteps) {
                        const uint32_t k1 = m1 & (KL - 1);
                        if (k1 == 0 && m1 + q < nsteps) {
                            move_words<Real>(a.seed, run_key, (uint32_t)t, sweep, m1 + q, mA, mB);
                            if constexpr (!F32) mthr = metro_thr(u53(mB.z, mB.w), beta);
                            if constexpr (PRE) pre_decode();
                        }
                        if constexpr (PRE) pre_fetch(gb + (int)k1);
                        else fetch_words<Real>(mA, mB, gb + (int)k1, nA, nB);
                        if constexpr (!F32) nthr = lane_get_d(mthr, gb + (int)k1);
                    }
                }
                // NVT in the sphere: every index is an atom; NPT: index N is the volume move
                const int i = PRE ? -1 : (int)d0.index;
                if constexpr (!PRE) {
                    if (i < N) {
                        li = gb + (i % KL);
                        xo = xoff<LANES>(i / KL, 0, li);
                    }
                }
                bool accept = false;
                double dE_used = 0.0;
                if (PRE || i < N) {
                    Real xi = X[xo];
                    Real yi = X[xo + LANES];
                    Real zi = X[xo + 2 * LANES];
                    Real tx, ty, tz;
                    if constexpr (PBC) {
                        box_trial<Real, BC, DIV>(xi, yi, zi, d0, delta_r, bxr, tx, ty, tz);
                    } else if constexpr (PRE) {
                        tx = xi + pdx;
                        ty = yi + pdy;
                        tz = zi + pdz;
                    } else {
                        make_trial<Real, BC, DIV>(xi, yi, zi, d0, delta_r, Real(1), Real(1), tx, ty, tz);
                    }
                    if constexpr (BC == BC_SPHERE) {
                        // src/MCMoves.jl:115-135: redraw until inside, recentre at 50, error at 100
                        if (outside_sphere<Real>(tx, ty, tz, radius2)) {
                            const int si = xo / (3 * LANES);
                            if (!sphere_redraw<Real, KL, LANES>(X, tid, q, N, A, si, li, radius2, delta_r, a.seed, run_key,
                                                                (uint32_t)t, sweep, m, nsteps, xi, yi, zi, tx, ty, tz) &&
                                q == 0)
                                a.run_error[a.slot_run[slot]] = 1;
                        }
                        // park the moved atom so its self pair contributes zero
                        if (tid == li) {
                            X[xo] = far_point<Real>();
                            X[xo + LANES] = far_point<Real>();
                            X[xo + 2 * LANES] = far_point<Real>();
                        }
                        __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    }
                    // dE = sum_j e(r'_ij) - sum_j e(r_ij), src/EnergyEvaluation.jl:804-821
                    Real sn = 0, so = 0;  // FP64: the two sums; binary32: sum of differences in sn
                    const Real *Xl = X + tid;
                    // binary32 sphere: trial distances from the old ones (dist2_pair_f32)
                    float m2dx = 0.0f, m2dy = 0.0f, m2dz = 0.0f, ddd = 0.0f;
                    if constexpr (F32 && BC == BC_SPHERE) {
                        const float dx = (float)tx - (float)xi, dy = (float)ty - (float)yi, dz = (float)tz - (float)zi;
                        m2dx = -2.0f * dx;
                        m2dy = -2.0f * dy;
                        m2dz = -2.0f * dz;
                        ddd = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                    }
                    constexpr int UNROLL = LANES > 256 ? 2 : 4;
#pragma unroll UNROLL
                    for (int s = 0; s < A; ++s) {
                        const Real xj = Xl[xoff<LANES>(s, 0, 0)];
                        const Real yj = Xl[xoff<LANES>(s, 1, 0)];
                        const Real zj = Xl[xoff<LANES>(s, 2, 0)];
                        Real r2n, r2o, en, eo;
                        if constexpr (PK == PK_ELJB || PK == PK_LUT) {
                            Real xyn, zn, xyo, zo;
                            r2n = box_d2c<Real, BC, DIV>(tx, ty, tz, xj, yj, zj, bxr, xyn, zn);
                            r2o = box_d2c<Real, BC, DIV>(xi, yi, zi, xj, yj, zj, bxr, xyo, zo);
                            if constexpr (PK == PK_ELJB) {
                                en = pair_energy_b<Real, KC>(r2n, xyn, zn, c, ca, cb);
                                eo = pair_energy_b<Real, KC>(r2o, xyo, zo, c, ca, cb);
                            } else {
                                en = pair_energy_lut<Real>(r2n, xyn, zn, a.pot);
                                eo = pair_energy_lut<Real>(r2o, xyo, zo, a.pot);
                            }
                        } else {
                        if constexpr (F32 && BC == BC_SPHERE) {
                            float fn, fo;
                            dist2_pair_f32((float)xj, (float)yj, (float)zj, (float)xi, (float)yi, (float)zi, m2dx, m2dy,
                                           m2dz, ddd, fn, fo);
                            r2n = (Real)fn;
                            r2o = (Real)fo;
                        } else {
                            r2n = box_d2<Real, BC, DIV>(tx, ty, tz, xj, yj, zj, bxr);
                            r2o = box_d2<Real, BC, DIV>(xi, yi, zi, xj, yj, zj, bxr);
                        }
                        if constexpr (F32 && PK == PK_LJ && BC == BC_SPHERE) {
                            // e(u) = u^3 (c12 u^3 + c6): e_new - e_old = (a - b)(c12 (a + b) + c6)
                            // with a = u_new^3, b = u_old^3 (accumulated below); one
                            // reciprocal serves both pairs
                            const float ip = __builtin_amdgcn_rcpf(r2n * r2o);
                            const float un = ip * r2o, uo = ip * r2n;
                            const float an = un * un * un, bo = uo * uo * uo;
                            en = an - bo;
                            eo = __builtin_fmaf(c[1], an + bo, c[0]);
                        } else if constexpr (F32 && (PK == PK_LJ || PK == PK_EVEN_H)) {
                            // one reciprocal serves the new and the old pair
                            const float ip = __builtin_amdgcn_rcpf(r2n * r2o);
                            en = pair_energy_u<Real, PK, KC>(ip * r2o, c);
                            eo = pair_energy_u<Real, PK, KC>(ip * r2n, c);
                        } else {
                            en = pair_energy<Real, PK, KC>(r2n, c);
                            eo = pair_energy<Real, PK, KC>(r2o, c);
                        }
                        }
                        if constexpr (BC == BC_SPHERE) {
                            if constexpr (F32 && PK == PK_LJ) sn = __builtin_fmaf(en, eo, sn);
                            else if constexpr (F32) sn += en - eo;
                            else {
                                sn += en;
                                so += eo;
                            }
                        } else {
                            const int j = s * KL + q;
                            bool ok = (j < N) & (j != i);
                            bool okn = ok, oko = ok;
                            if constexpr (ENS == ENS_NPT) {
                                okn = ok & (r2n <= rc);
                                oko = ok & (r2o <= rc);
                            }
                            if constexpr (F32) {
                                sn += (okn ? en : Real(0)) - (oko ? eo : Real(0));
                            } else {
                                sn += okn ? en : Real(0);
                                so += oko ? eo : Real(0);
                            }
                        }
                    }
                    if constexpr (F32) {
                        // Metropolis in binary32 (src/Exchange.jl:53-57, src/MCRun.jl:122-132)
                        const float dEf = group_sum<KL>((float)sn);
                        dE_used = (double)dEf;
                        if constexpr (!PBC) {
                            // exp2(-beta log2e dE) >= u  <=>  dE <= -log2(u) / (beta log2 e),
                            // the rule (and the rounding) of the spread layout's decoded draws
                            accept = dEf <= -__builtin_amdgcn_logf(d0.umet) * (1.0f / beta_l2e);
                        } else {
                            const float pr = __builtin_amdgcn_exp2f(-dEf * beta_l2e);
                            accept = (dEf <= 0.0f) || (pr >= d0.umet);
                        }
                        if (accept) {
                            E = E + (double)dEf;
                            nacc += 1;
                        }
                    } else {
                        sn = group_sum<KL>(sn);
                        so = group_sum<KL>(so);
                        const double delta_en = sn - so;
                        const double new_en = delta_en + E;
                        const double dd = new_en - E;
                        dE_used = dd;
                        accept = dd <= thr;
                        if (accept) {
                            E = new_en;
                            nacc += 1;
                        }
                    }
                    // commit (src/swap_config.jl:40-51) or restore the parked atom
                    if (tid == li) {
                        X[xo] = accept ? tx : xi;
                        X[xo + LANES] = accept ? ty : yi;
                        X[xo + 2 * LANES] = accept ? tz : zi;
                    }
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                }
                if constexpr (ENS == ENS_NPT) {
                    // volume move: uniform (src/MCMoves.jl:169-177, :277-287) or separated,
                    // z 1/6, xy 1/3, xyz 1/2 of the moves (src/MCMoves.jl:235-343); energy
                    // src/EnergyEvaluation.jl:397-474; acceptance src/Exchange.jl:58-71
                    const bool vm = i >= N;
                    int vkind = 0;  // 0: xyz, 1: xy, 2: z
                    double sx = 1.0, sz = 1.0;
                    if (vm) {
                        const int T = tb + t;
                        if (!a.separated) {
                            const double y = ((double)d0.ux - 0.5) * a.t_displ[2 * T + 1];
                            double scale = det_exp(y / 3.0);
                            if (box >= a.t_maxbox[T] && scale > 1.0) scale = 1.0;
                            sx = sz = scale;
                        } else {
                            const int ra = 1 + (int)((double)d0.uy * 6.0);
                            const double lh = box / boxh, lr = a.t_lh[T];
                            if (ra == 1) {
                                vkind = 2;
                                const double y = ((double)d0.ux - 0.5) * a.t_displ_sep[2 * T + 1];
                                double scale = det_exp(y);
                                if (lh <= lr * 1.1 && scale > 1.0) scale = 1.0 / scale;
                                else if (lh >= lr * 0.909 && scale < 1.0) scale = 1.0 / scale;
                                if (boxh >= a.t_maxboxh[T] && scale > 1.0) scale = 1.0 / scale;
                                sz = scale;
                            } else if (ra <= 3) {
                                vkind = 1;
                                const double y = ((double)d0.ux - 0.5) * a.t_displ_sep[2 * T];
                                double scale = det_exp(y / 2.0);
                                if (lh >= lr * 1.1 && scale > 1.0) scale = 1.0 / scale;
                                else if (lh <= lr * 0.909 && scale < 1.0) scale = 1.0 / scale;
                                if (box >= a.t_maxbox[T] && scale > 1.0) scale = 1.0 / scale;
                                sx = scale;
                            } else {
                                const double y = ((double)d0.ux - 0.5) * a.t_displ[2 * T + 1];
                                double scale = det_exp(y / 3.0);
                                if (box >= a.t_maxbox[T] && scale > 1.0) scale = 1.0;
                                sx = sz = scale;
                            }
                        }
                    }
                    const double nbox = box * sx, nh = boxh * sz;
                    const double nrc = box_rcut(kind, nbox, nh);
                    // uniform: the current positions in the current box, r^2 scaled by s^2;
                    // separated: the scaled positions in the trial box
                    const bool sep = a.separated != 0;
                    const double px = sep ? sx : 1.0, pz = sep ? sz : 1.0, m2 = sep ? 1.0 : sx * sx;
                    const double eb = sep ? nbox : box, ebh = sep ? nh : boxh;
                    double new_en = 0.0;
                    if constexpr (F32) {
                        // the clusters of a wave evaluate each pending volume move
                        // together instead of leaving the wave to one cluster's
                        // O(N^2) sum, once per cluster
                        __atomic_signal_fence(__ATOMIC_SEQ_CST);
                        const uint64_t live = __ballot(1);
                        const uint64_t leaders = __ballot(q == 0);
                        uint64_t todo = __ballot(vm && q == 0);
                        while (todo != 0) {
                            const int cl = __builtin_ctzll(todo);
                            todo &= todo - 1;
                            const double e = wave_pair_sum64<PK, KC, BC, KL, LANES>(
                                X, cl, N, make_box<double>(readlane_d(eb, cl), readlane_d(ebh, cl)), readlane_d(px, cl),
                                readlane_d(pz, cl), readlane_d(m2, cl), readlane_d(nrc, cl), c64, ca64, cb64, a.pot,
                                live, leaders);
                            if ((gb & 63) == cl) new_en = e;
                        }
                    } else {
                        if (vm)
                            new_en = pair_sum64<Real, PK, KC, BC, ENS, KL, LANES>(X, gb, q, N, make_box<double>(eb, ebh),
                                                                                  px, pz, m2, nrc, c64, ca64, cb64,
                                                                                  a.pot);
                    }
                    if (vm) {
                        new_en = new_en + box_lrc<PK, KC>(kind, N, nrc, nbox, nh, c64);
                        const double pressure = a.runs[a.slot_run[slot]].pressure;
                        const double dd = new_en - E;
                        dE_used = dd;
                        const double vn = box_volume(kind, nbox, nh), vo = box_volume(kind, box, boxh);
                        const double dh = dd + pressure * (vn - vo);
                        const double arg = -dh * beta + (double)N * det_log(vn / vo);
                        double prob = det_exp(arg);
                        if (prob > 1.0) prob = 1.0;
                        accept = prob >= (double)d0.umet;
                        if (accept) {
                            for (int s = 0; s < A; ++s) {
                                Real *pxp = &X[xoff<LANES>(s, 0, tid)];
                                Real *pyp = &X[xoff<LANES>(s, 1, tid)];
                                Real *pzp = &X[xoff<LANES>(s, 2, tid)];
                                *pxp = (Real)((double)*pxp * sx);
                                *pyp = (Real)((double)*pyp * sx);
                                *pzp = (Real)((double)*pzp * sz);
                            }
                            __atomic_signal_fence(__ATOMIC_SEQ_CST);
                            box = nbox;
                            boxh = nh;
                            E = new_en;
                            if (vkind == 0) nvacc += 1;
                            else if (vkind == 1) nvxy += 1;
                            else nvz += 1;
                        }
                    }
                }
                if (a.trace_de != nullptr && q == 0 && cyc < a.trace_cycles) {
                    const size_t ti = ((size_t)(tb + t) * a.trace_cycles + cyc) * nsteps + m;
                    a.trace_de[ti] = dE_used;
                    a.trace_acc[ti] = (int8_t)accept;
                    a.trace_idx[ti] = PRE ? (xo / (3 * LANES)) * KL + (li - gb) : i;
                }
                if constexpr (ENS == ENS_NPT) {
                    // an accepted volume move changes the box for the remaining moves
                    if (i >= N) {
                        bxr = make_box<Real>(box, boxh);
                        rc = (Real)box_rcut(kind, box, boxh);
                    }
                }
            }
            if (q == 0) {
                count(lb + t, tb + t, 0) += nacc;
                count(lb + t, tb + t, 1) += nacc;
                if (ENS == ENS_NPT) {
                    a.t_count_vol[2 * (tb + t)] += nvacc;
                    a.t_count_vol[2 * (tb + t) + 1] += nvacc;
                    if (a.separated) {
                        a.t_count_volxy[2 * (tb + t)] += nvxy;
                        a.t_count_volxy[2 * (tb + t) + 1] += nvxy;
                        a.t_count_volz[2 * (tb + t)] += nvz;
                        a.t_count_volz[2 * (tb + t) + 1] += nvz;
                    }
                }
                exE[lg] = E;
                if constexpr (ENS == ENS_NPT) exV[lg] = box_volume(kind, box, boxh);
            }
        }
        __syncthreads();

        // replica exchange, src/Exchange.jl:103-186 and src/MCRun.jl:203-205,
        // attempted every exchange_every cycles
        if (active && q == 0 && R > 1 && (uint32_t)(sweep + 1u) % (uint32_t)a.exchange_every == 0u) {
            int n = -1;
            if (a.exchange_mode == EXC_EVEN_ODD) {
                if ((t & 1) == (int)(sweep & 1u) && t + 1 < R) n = t;
            } else if (t == 0) {
                u32x4 X0 = draw(a.seed, run_key, RNG_EXCH, 0u, 0u, sweep, 0u);
                double u = u53(X0.x, X0.y);
                if (a.exchange_mode == EXC_EVERY || u < 0.1) n = (int)below(X0.z, (uint32_t)(R - 1));
            }
            if (n >= 0) {
                u32x4 X1 = draw(a.seed, run_key, RNG_EXCH, 0u, 0u, sweep, 1u + (uint32_t)n);
                const double u = u53(X1.x, X1.y);
                const int sa = exS[lb + n], sb = exS[lb + n + 1];
                double e1 = exE[sa], e2 = exE[sb];
                if (a.ensemble == ENS_NPT) {
                    const double pressure = a.runs[a.slot_run[slot]].pressure;
                    e1 = e1 + pressure * exV[sa];
                    e2 = e2 + pressure * exV[sb];
                }
                const double de = e1 - e2;
                const double db = beta_of(lb + n, tb + n) - beta_of(lb + n + 1, tb + n + 1);
                double acc = det_exp(db * de);
                if (acc > 1.0) acc = 1.0;
                count(lb + n, tb + n, 2) += 1;
                count(lb + n + 1, tb + n + 1, 2) += 1;
                if (acc > u) {
                    count(lb + n, tb + n, 3) += 1;
                    count(lb + n + 1, tb + n + 1, 3) += 1;
                    exT[sa] = n + 1;
                    exT[sb] = n;
                    exS[lb + n] = sb;
                    exS[lb + n + 1] = sa;
                }
            }
        }
        __syncthreads();

        if (active) {
            t = exT[lg];
            if (q == 0) {
                const int T = tb + t, lt = lb + t;
                // update_max_stepsize!, src/Exchange.jl:198-263
                if (pidx % a.n_adjust == 0) {
                    const double ar = (double)count(lt, T, 1) / (double)((long long)a.n_adjust * N);
                    if (ar < a.min_acc) displ(lt, T) *= 0.9;
                    else if (ar > a.max_acc) displ(lt, T) *= 1.1;
                    count(lt, T, 1) = 0;
                    if (ENS == ENS_NPT && !a.separated) {
                        const double av = (double)a.t_count_vol[2 * T + 1] / (double)a.n_adjust;
                        if (av < a.min_acc) a.t_displ[2 * T + 1] *= 0.9;
                        else if (av > a.max_acc) a.t_displ[2 * T + 1] *= 1.1;
                        a.t_count_vol[2 * T + 1] = 0;
                    } else if (ENS == ENS_NPT) {
                        // separated moves: expected shares 1/2 (xyz), 1/3 (xy), 1/6 (z)
                        double av = (double)a.t_count_vol[2 * T + 1] / ((double)a.n_adjust / 2.0);
                        if (av < a.min_acc) a.t_displ[2 * T + 1] *= 0.9;
                        else if (av > a.max_acc) a.t_displ[2 * T + 1] *= 1.1;
                        a.t_count_vol[2 * T + 1] = 0;
                        av = (double)a.t_count_volxy[2 * T + 1] / ((double)a.n_adjust / 3.0);
                        if (av < a.min_acc) a.t_displ_sep[2 * T] *= 0.9;
                        else if (av > a.max_acc) a.t_displ_sep[2 * T] *= 1.1;
                        a.t_count_volxy[2 * T + 1] = 0;
                        av = (double)a.t_count_volz[2 * T + 1] / ((double)a.n_adjust / 6.0);
                        if (av < a.min_acc) a.t_displ_sep[2 * T + 1] *= 0.9;
                        else if (av > a.max_acc) a.t_displ_sep[2 * T + 1] *= 1.1;
                        a.t_count_volz[2 * T + 1] = 0;
                    }
                }
                // sampling_step!, src/Sampling.jl:29-66, :258-280, :330-371
                // (histogram counts by integer atomics, which the wave does not wait for)
                if (a.sampling && pidx % a.mc_sample == 0) {
                    if (ENS == ENS_NPT) {
                        const double v = box_volume(kind, box, boxh);
                        const double h = E + pressure * v;
                        ham(lt, T, 0) += h;
                        ham(lt, T, 1) += h * h;
                        const double ie = __builtin_floor((E - en_min) / delta_en) + 1.0;
                        const double iv = __builtin_floor((box_hist_volume(kind, box, boxh) - v_min) / delta_v) + 1.0;
                        const int he = ie < 1.0 ? 1 : (ie > (double)a.n_bin ? a.n_bin + 2 : (int)__builtin_floor(ie) + 1);
                        const int hv = iv < 1.0 ? 1 : (iv > (double)a.n_bin ? a.n_bin + 2 : (int)__builtin_floor(iv) + 1);
                        atomicAdd(&a.t_hist[(size_t)T * a.hist_stride + (size_t)(he - 1) * (a.n_bin + 2) + (hv - 1)], 1u);
                    } else {
                        ham(lt, T, 0) += E;
                        ham(lt, T, 1) += E * E;
                        const double hi = (E - en_min) / delta_en + 1.0;
                        const int h = hi < 1.0 ? 1 : (hi > (double)a.n_bin ? a.n_bin + 2 : (int)__builtin_floor(hi) + 1);
                        atomicAdd(&a.t_hist[(size_t)T * a.hist_stride + (h - 1)], 1u);
                    }
                }
            }
            // periodic FP64 recompute bounding binary32 drift
            if (a.recompute_every > 0 && ((long long)sweep + 1) % a.recompute_every == 0) {
                const double rc64 = box_rcut(kind, box, boxh);
                double ef = pair_sum64<Real, PK, KC, BC, ENS, KL, LANES>(X, gb, q, N, make_box<double>(box, boxh), 1.0,
                                                                         1.0, 1.0, rc64, c64, ca64, cb64, a.pot);
                if constexpr (ENS == ENS_NPT) ef = ef + box_lrc<PK, KC>(kind, N, rc64, box, boxh, c64);
                const double dr = __builtin_fabs(E - ef);
                if (q == 0 && dr > a.slot_drift[slot]) a.slot_drift[slot] = dr;
                E = ef;
            }
        }
    }

    __syncthreads();
    for (int e = tid; e < 3 * A * LANES; e += LANES) gX[e] = X[e];
    if (active && q == 0) {
        a.slot_temp[slot] = t;
        a.slot_en[slot] = E;
        if constexpr (PBC) {
            a.slot_box[slot] = box;
            a.slot_boxh[slot] = boxh;
        }
        // the per-temperature state of the temperature this slot holds now
        const int lt = lb + t, T = tb + t;
        if constexpr (RS) {
        a.t_count_atom[2 * T] = rsC[4 * lt];
        a.t_count_atom[2 * T + 1] = rsC[4 * lt + 1];
        a.t_count_exc[2 * T] = rsC[4 * lt + 2];
        a.t_count_exc[2 * T + 1] = rsC[4 * lt + 3];
        a.t_ham[2 * T] = rsH[2 * lt];
        a.t_ham[2 * T + 1] = rsH[2 * lt + 1];
        a.t_displ[2 * T] = rsD[lt];
        }
    }
}


// initialise_energy (src/EnergyEvaluation.jl:1662-1692): FP64 total energy of
// each slot of the tile from its positions (all runs, or the one init_run).
template <typename Real, int PK, int KC, int BC, int ENS, int KL, int LANES>
__global__ void __launch_bounds__(LANES) init_kernel(SweepArgs a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const TileInfo tile = a.tiles[blockIdx.x];
    const int tid = threadIdx.x;
    const int lg = tid / KL;
    const int q = tid % KL;
    const int N = tile.n_atoms;
    const int A = tile.n_slab;
    Real *X = reinterpret_cast<Real *>(lds_raw);
    const Real *gX = reinterpret_cast<const Real *>(a.coords) + tile.coord_off;
    for (int e = tid; e < 3 * A * LANES; e += LANES) X[e] = gX[e];
    __syncthreads();
    if (lg >= tile.n_slots) return;  // whole clusters leave together
    const int slot = tile.slot0 + lg;
    if (a.init_run >= 0 && a.slot_run[slot] != a.init_run) return;
    double c64[KC], ca64[KC], cb64[KC];
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        c64[k] = k < a.pot.ncoeff ? a.pot.coeff[k] : 0.0;
        ca64[k] = (PK == PK_ELJB && k < a.pot.ncoeff && k < 8) ? a.pot.coeff_a[k] : 0.0;
        cb64[k] = (PK == PK_ELJB && k < a.pot.ncoeff && k < 8) ? a.pot.coeff_b[k] : 0.0;
    }
    const bool pbc = BC != BC_SPHERE;
    const int kind = BC == BC_RHOMB ? (int)BC_RHOMB : (BC == BC_CUBIC ? a.box_kind : (int)BC_SPHERE);
    const double box = pbc ? a.slot_box[slot] : 1.0, boxh = pbc ? a.slot_boxh[slot] : 1.0;
    const double rc64 = pbc ? box_rcut(kind, box, boxh) : 0.0;
    double E = pair_sum64<Real, PK, KC, BC, ENS, KL, LANES>(X, lg * KL, q, N, make_box<double>(box, boxh), 1.0, 1.0,
                                                            1.0, rc64, c64, ca64, cb64, a.pot);
    if constexpr (ENS == ENS_NPT) E = E + box_lrc<PK, KC>(kind, N, rc64, box, boxh, c64);
    if (q == 0) a.slot_en[slot] = E;
}

}  // namespace ptmc
