// Sine/cosine and reciprocal building blocks for the waveform kernels.
//
// All sine/cosine routines reduce x = y + q*pi/2 with y in [-pi/4, pi/4] and
// evaluate Taylor series of sin and cos in y (coefficients 1/n!).  The first
// omitted terms bound the truncation error:
//   FP64: sin through y^15, cos through y^16 -> |y|^17/17! < 4.7e-17 (below
//         half an ulp of |sin y| >= 0.64 |y| and of cos y >= 0.7);
//   FP32: sin through y^9,  cos through y^10 -> |y|^11/11! < 1.8e-9.
// Reductions:
//   sincos_f64(x): k = rint(2x/pi), y = x - k*P1 - k*P2 with two FMAs, where
//     P1 = fl(pi/2) and P2 = fl(pi/2 - P1); absolute error ~1e-16 for
//     |k| < 2^31 (|x| < 3e9 rad), far beyond the phases met here.
//   sincos_f32(x): the same in FP32 for |x| < ~1e4;
//   sincos_rev(r): r in revolutions held in FP64: the quarter-turn split
//     r - k/4 is exact in FP64, the remainder goes to the FP32 polynomial.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace gbgpu {

// ------------------------------------------------------------------ FP64 ---
__device__ __forceinline__ void sincos_poly_f64(double y, int q, double &s, double &c) {
    const double y2 = y * y;
    double ps = -7.647163731819816e-13;        // -1/15!
    ps = fma(ps, y2, 1.6059043836821613e-10);  // 1/13!
    ps = fma(ps, y2, -2.505210838544172e-08);  // -1/11!
    ps = fma(ps, y2, 2.7557319223985893e-06);  // 1/9!
    ps = fma(ps, y2, -1.984126984126984e-04);  // -1/7!
    ps = fma(ps, y2, 8.333333333333333e-03);   // 1/5!
    ps = fma(ps, y2, -0.16666666666666666);    // -1/3!
    const double sy = fma(y * y2, ps, y);
    double pc = -4.779477332387385e-14;        // -1/16!
    pc = fma(pc, y2, 1.1470745597729725e-11);  // 1/14!
    pc = fma(pc, y2, -2.08767569878681e-09);   // -1/12!
    pc = fma(pc, y2, 2.755731922398589e-07);   // 1/10!
    pc = fma(pc, y2, -2.48015873015873e-05);   // -1/8!
    pc = fma(pc, y2, 1.388888888888889e-03);   // 1/6!
    pc = fma(pc, y2, -0.041666666666666664);   // -1/4!
    pc = fma(pc, y2, 0.5);                     // 1/2!
    const double cy = fma(-y2, pc, 1.0);
    const bool swap = q & 1;
    const double s0 = swap ? cy : sy;
    const double c0 = swap ? sy : cy;
    s = (q & 2) ? -s0 : s0;
    c = ((q + 1) & 2) ? -c0 : c0;
}

__device__ __forceinline__ void sincos_f64(double x, double &s, double &c) {
    constexpr double TWO_OVER_PI = 0.63661977236758134308;
    constexpr double P1 = 1.5707963267948966;      // fl(pi/2)
    constexpr double P2 = 6.123233995736766e-17;   // fl(pi/2 - P1)
    const double k = rint(x * TWO_OVER_PI);
    double y = fma(-k, P1, x);
    y = fma(-k, P2, y);
    // |k| < 2^31 for every argument here, so a 32-bit conversion is exact
    sincos_poly_f64(y, (int)k & 3, s, c);
}

// sin and cos of x + d for a large x and a small d (|d| < 2): the quadrant is
// taken from x + d, the reduction is applied to x alone, and d is added to
// the reduced value, so |y| <= pi/4 (up to rounding) and the rounding of the
// large sum x + d never reaches the result.
__device__ __forceinline__ void sincos_f64_shift(double x, double d, double &s, double &c) {
    constexpr double TWO_OVER_PI = 0.63661977236758134308;
    constexpr double P1 = 1.5707963267948966;
    constexpr double P2 = 6.123233995736766e-17;
    const double k = rint((x + d) * TWO_OVER_PI);
    double y = fma(-k, P1, x);
    y = fma(-k, P2, y) + d;
    sincos_poly_f64(y, (int)k & 3, s, c);
}

// 1/x to ~1 ulp: hardware estimate refined by two Newton steps.
__device__ __forceinline__ double rcp_f64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// ------------------------------------------------------------------ FP32 ---
__device__ __forceinline__ void sincos_poly(float y, int q, float &s, float &c) {
    const float y2 = y * y;
    float ps = 1.0f / 362880.0f;
    ps = ps * y2 - 1.0f / 5040.0f;
    ps = ps * y2 + 1.0f / 120.0f;
    ps = ps * y2 - 1.0f / 6.0f;
    const float sy = y + y * y2 * ps;
    float pc = -1.0f / 3628800.0f;
    pc = pc * y2 + 1.0f / 40320.0f;
    pc = pc * y2 - 1.0f / 720.0f;
    pc = pc * y2 + 1.0f / 24.0f;
    pc = pc * y2 - 0.5f;
    const float cy = 1.0f + y2 * pc;
    // rotate by q quarter turns
    const bool swap = q & 1;
    const float s0 = swap ? cy : sy;
    const float c0 = swap ? sy : cy;
    s = (q & 2) ? -s0 : s0;
    c = ((q + 1) & 2) ? -c0 : c0;
}

__device__ __forceinline__ void sincos_f32(float x, float &s, float &c) {
    constexpr float TWO_OVER_PI = 0.636619772367581343f;
    constexpr float PIO2_HI = 1.57079637050628662109375f;  // fl32(pi/2)
    constexpr float PIO2_LO = -4.37113900018624283e-8f;    // pi/2 - PIO2_HI
    const float k = rintf(x * TWO_OVER_PI);
    const float y = fmaf(-k, PIO2_LO, fmaf(-k, PIO2_HI, x));
    sincos_poly(y, ((int)k) & 3, s, c);
}

__device__ __forceinline__ void sincos_rev(double r, float &s, float &c) {
    const double k = rint(4.0 * r);
    const float y = (float)(r - 0.25 * k) * 6.28318530717958647692f;
    sincos_poly(y, ((int)k) & 3, s, c);  // |k| < 2^31: exact 32-bit conversion
}

// ------------------------------------------------- reduction-free series ---
// Coefficients of the two even series
//   Rs(v) = sum_k (-1)^k v^k / (2k+1)!   (sin x = x Rs(x^2), sin u / u = Rs(u^2))
//   Is(v) = sum_k (-1)^k v^k / (2k+2)!   (cos x = 1 - x^2 Is(x^2),
//                                         (1 - cos u)/u = u Is(u^2))
struct SeriesCoef {
    static constexpr double R[13] = {1.0,
                                     -0.16666666666666666,
                                     0.008333333333333333,
                                     -0.0001984126984126984,
                                     2.7557319223985893e-06,
                                     -2.505210838544172e-08,
                                     1.6059043836821613e-10,
                                     -7.647163731819816e-13,
                                     2.8114572543455206e-15,
                                     -8.22063524662433e-18,
                                     1.9572941063391263e-20,
                                     -3.868170170630684e-23,
                                     6.446950284384474e-26};
    static constexpr double I[13] = {0.5,
                                     -0.041666666666666664,
                                     0.001388888888888889,
                                     -2.48015873015873e-05,
                                     2.755731922398589e-07,
                                     -2.08767569878681e-09,
                                     1.1470745597729725e-11,
                                     -4.779477332387385e-14,
                                     1.5619206968586225e-16,
                                     -4.110317623312165e-19,
                                     8.896791392450574e-22,
                                     -1.6117375710961184e-24,
                                     2.4795962632247976e-27};
};

template <int K, class R>
__device__ __forceinline__ R horner_R(R v) {
    R p = (R)SeriesCoef::R[K - 1];
#pragma unroll
    for (int k = K - 2; k >= 0; k--) p = p * v + (R)SeriesCoef::R[k];
    return p;
}

template <int K, class R>
__device__ __forceinline__ R horner_I(R v) {
    R p = (R)SeriesCoef::I[K - 1];
#pragma unroll
    for (int k = K - 2; k >= 0; k--) p = p * v + (R)SeriesCoef::I[k];
    return p;
}

// sin/cos of a small argument with K terms of each series.
template <int K, class R>
__device__ __forceinline__ void sincos_series(R x, R &s, R &c) {
    const R x2 = x * x;
    s = x * horner_R<K>(x2);
    c = (R)1 - x2 * horner_I<K>(x2);
}

// Link transfer factor sinc(u/2) e^{-iu/2} = sin(u)/u - i (1 - cos u)/u
// (gbgpu.py:518-525 with u = 2a), KR and KI terms.
template <int KR, int KI, class R>
__device__ __forceinline__ void transfer_series(R u, R &re, R &im) {
    const R u2 = u * u;
    re = horner_R<KR>(u2);
    im = -u * horner_I<KI>(u2);
}

// Series lengths per frequency tier.  Tier t applies when
// u_max = 2 max|f| / f* (light-time and Doppler margin included) is at most
// U_MAX[t]; the lengths are the shortest whose first omitted term is below
// 1e-16 (FP64) or 3e-9 (FP32), relative to a result of order one, for the
// transfer factor (|u| <= U_MAX: real part v^KR/(2KR+1)!, imaginary part
// u v^KI/(2KI+2)!, v = U_MAX^2), the sine/cosine of omL = f/f*
// (<= U_MAX/2) and of the spacecraft phase offsets (<= 0.29 U_MAX rad).
// Tier NTIER falls back to range-reduced sine/cosine.
constexpr int NTIER = 4;
constexpr double TIER_UMAX[NTIER] = {0.4, 0.75, 1.5, 2.25};

constexpr int tier_pick(int t, int a, int b, int c, int d) { return t == 0 ? a : t == 1 ? b : t == 2 ? c : d; }

template <int TIER>
struct TierF64 {
    static constexpr int KR = tier_pick(TIER, 7, 8, 10, 12);
    static constexpr int KI = tier_pick(TIER, 6, 8, 10, 12);
    static constexpr int KS = tier_pick(TIER, 6, 7, 8, 9);  // sincos of omL
    static constexpr int KD = tier_pick(TIER, 5, 6, 7, 8);  // sincos of phase offsets
};
template <int TIER>
struct TierF32 {
    static constexpr int KR = tier_pick(TIER, 4, 5, 7, 8);
    static constexpr int KI = tier_pick(TIER, 4, 5, 6, 8);
    static constexpr int KS = tier_pick(TIER, 3, 4, 5, 6);  // sincos of omL
    static constexpr int KD = tier_pick(TIER, 3, 4, 5, 5);  // sincos of phase offsets
};

// Calls f(std::integral_constant<int, TIER>) for a binary's tier (0 ..
// NTIER - 1: series; NTIER: reduced sine/cosine).  One call per binary, not
// per sample, so that each tier's loop is its own straight-line code.
template <class F>
__device__ __forceinline__ void with_tier(int tier, F &&f) {
    static_assert(NTIER == 4, "with_tier lists every tier");
    switch (tier) {
        case 0: f(std::integral_constant<int, 0>{}); break;
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 3: f(std::integral_constant<int, 3>{}); break;
        default: f(std::integral_constant<int, NTIER>{}); break;
    }
}

}  // namespace gbgpu
