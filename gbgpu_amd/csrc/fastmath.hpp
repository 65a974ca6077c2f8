// FP32 sine/cosine for the mixed-precision slow part.
//
// Both routines reduce to y in [-pi/4, pi/4] plus a quadrant and evaluate
// the Taylor series of sin (through y^9) and cos (through y^10); the first
// omitted terms are |y|^11/11! <= 1.8e-9 and |y|^12/12! <= 1.2e-10, below
// FP32 rounding.
//   sincos_f32(x):         x in radians, |x| < ~1e4 (Cody-Waite pi/2 split);
//   sincos_rev(r):         r in revolutions held in FP64 (any magnitude): the
//                          quadrant split r - k/4 is done in FP64 (exact for
//                          |r| < 2^50), so a large FP64 phase loses nothing
//                          before the FP32 polynomial.
#pragma once

#include "common.hpp"

namespace gbgpu {

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
    constexpr float PIO2_HI = 1.57079637050628662109375f;     // float(pi/2)
    constexpr float PIO2_LO = -4.37113900018624283e-8f;       // pi/2 - PIO2_HI
    const float k = rintf(x * TWO_OVER_PI);
    const float y = fmaf(-k, PIO2_LO, fmaf(-k, PIO2_HI, x));
    sincos_poly(y, ((int)k) & 3, s, c);
}

__device__ __forceinline__ void sincos_rev(double r, float &s, float &c) {
    const double k = rint(4.0 * r);
    const float y = (float)(r - 0.25 * k) * 6.28318530717958647692f;
    sincos_poly(y, ((long long)k) & 3, s, c);
}

}  // namespace gbgpu
