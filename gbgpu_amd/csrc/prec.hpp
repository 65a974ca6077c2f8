// Precision policies for the waveform kernels.
//
//   F64   ("plain"): every operation in FP64, sin/cos from the device math
//         library; spectra kept as double2.
//   Mixed: FP64 for t_n, spacecraft positions, k.P/c, xi, the carrier phase
//         argS and Phi = om k.P/c - argS (reduced to a quarter turn in FP64),
//         the start bin and (in the epilogues) the likelihood accumulation;
//         FP32 for arm vectors, antenna patterns, fonfs, the six transfer
//         functions, the TDI combination, all sines/cosines (fastmath.hpp)
//         and the LDS transform.  Amplitudes are carried relative to amp in
//         FP32 and amp is restored in FP64 after the transform.
//
// Each policy provides: CT (LDS complex type), Aux (extra per-binary
// constants), make_aux(), sample<THIRD, TIER>() (time-domain X, Y, Z at
// sample n and their common factor F = fctr2 * 0.5 T / N, applied by the
// caller once per output channel), fac() (complex factor applied to every
// output bin in FP64: TDI2 factor, and amp for Mixed) and twiddles().
#pragma once

#include "common.hpp"
#include "fastmath.hpp"
#include "source.hpp"

namespace gbgpu {

struct F64 {
    using CT = double2;
    struct Aux {};
    static __device__ __forceinline__ Aux make_aux(const Src &) { return Aux{}; }

    // TIER: the binary's series tier (Src::tier), chosen once per binary by
    // the caller so that each tier's sample loop is its own straight-line code
    template <bool THIRD, int TIER>
    static __device__ __forceinline__ void sample(const Src &s, const Aux &, const double *__restrict__ tab, int N,
                                                  int n, double fstar_inv, double scale, const float *v3,
                                                  const double *t3, CT &X, CT &Y, CT &Z, CT &F) {
        const SampleGeom g = load_geom(s, tab, N, n);
        double v3d[3] = {0.0, 0.0, 0.0};
        if (THIRD) {
            v3d[0] = v3[0];
            v3d[1] = v3[1];
            v3d[2] = v3[2];
        }
        slow_xyz_f64<THIRD, TIER>(s, g, fstar_inv, scale, v3d, t3, X, Y, Z, F);
    }

    // Derived waveforms: f_i and argS of the three spacecraft at sample n
    // given (fi3, argS3), as the class's hooks returned them.
    template <int TIER>
    static __device__ __forceinline__ void sample_ext(const Src &s, const double *__restrict__ tab, int N, int n,
                                                      double fstar_inv, double scale, const double *fi3,
                                                      const double *argS3, CT &X, CT &Y, CT &Z, CT &F) {
        const SampleGeom g = load_geom(s, tab, N, n);
        slow_xyz_f64<false, TIER, true>(s, g, fstar_inv, scale, fi3, argS3, X, Y, Z, F);
    }

    static __device__ __forceinline__ double2 fac(const Src &s) { return make_double2(s.tdr, s.tdi); }
    static __device__ __forceinline__ double2 up(CT v) { return v; }
    static const void *twiddles(int N) { return twiddles_f64(N); }
};

// FP32 copies of the per-binary constants used by the FP32 part.
struct SrcF {
    float k[3], u[3], v[3];
    float2 DP, DC;  // 0.25 * (DP, DC) / amp
    float f0, fdot, fddot;
};

// The same factor for any u, through sine/cosine with range reduction.
__device__ __forceinline__ float2 transfer_general(float u) {
    const float a = 0.5f * u;
    float sa, ca;
    sincos_f32(a, sa, ca);
    const float sinc = (a != 0.0f) ? sa * __builtin_amdgcn_rcpf(a) : 1.0f;
    return make_float2(sinc * ca, -sinc * sa);
}

// Mixed-precision slow part.  The light-time delay of spacecraft s is split
// as k.P_s/c = k.Pc/c + d_s: the guiding-centre delay k.Pc/c (~500 s) and the
// centre phase Phi_c = om k.Pc/c - argS(t - k.Pc/c) are formed and reduced to
// a quarter turn in FP64; the offsets d_s = k.dP_s/c (|d_s| < 5 s) and
//   Phi_s - Phi_c = 2 pi f(xi_c) d_s - pi (fdot + fddot xi_c) d_s^2
//                   + (pi/3) fddot d_s^3          (exact for the cubic argS)
// are FP32.  The third-body variant keeps one FP64 phase per spacecraft
// because its phase integral differs per spacecraft.  As in slow_xyz_f64,
// the circular path returns X, Y, Z without the links' common e^{-i Phi_c},
// which F carries: only F X, F Y, F Z are defined.
template <bool THIRD, int TIER>
__device__ __forceinline__ void slow_xyz_mixed(const Src &s, const SrcF &sf, const double *__restrict__ tab, int N,
                                               int n, double fstar_inv_d, float scale, const float *v3,
                                               const double *term3, float2 &X, float2 &Y, float2 &Z,
                                               float2 &F) {
    const float fstar_inv = (float)fstar_inv_d;
    constexpr double INV_C = 1.0 / C_SI;
    constexpr float INV_CF = (float)(1.0 / C_SI);
    constexpr float PIF = 3.14159265358979323846f;
    const TabRd T(tab, N);
    const double t = T(ROW_T, n);
    float r[9], dP[9];
#pragma unroll
    for (int i = 0; i < 9; i++) r[i] = (float)T(ROW_R + i, n);
#pragma unroll
    for (int i = 0; i < 9; i++) dP[i] = (float)T(ROW_DP + i, n);
    const double kPc = (s.k[0] * T(ROW_PC, n) + s.k[1] * T(ROW_PC + 1, n) + s.k[2] * T(ROW_PC + 2, n)) * INV_C;
    float d[3];
#pragma unroll
    for (int sc = 0; sc < 3; sc++)
        d[sc] = (sf.k[0] * dP[3 * sc] + sf.k[1] * dP[3 * sc + 1] + sf.k[2] * dP[3 * sc + 2]) * INV_CF;
    const double xic = t - kPc;
    const float xcf = (float)xic;
    const float fdc = sf.fdot + sf.fddot * xcf;               // f'(xi_c)
    const float fc = sf.f0 + sf.fdot * xcf + 0.5f * sf.fddot * xcf * xcf;
    float fonfs[3], cphi[3], sphi[3];
    if constexpr (!THIRD) {
        // The links' common e^{-i Phi_c} is left out of X, Y, Z and carried by
        // F, joined with the TDI factor's e^{-i omL} in one FP64-reduced
        // sine/cosine of Phi_c + omL (formed after the links, so that no
        // FP64 phase stays live through them); the links keep
        // e^{-i (Phi_s - Phi_c)}.
        const float w1 = 2.0f * PIF * fc, w2 = PIF * fdc, w3 = (PIF / 3.0f) * sf.fddot;
#pragma unroll
        for (int sc = 0; sc < 3; sc++) {
            const float dd = d[sc];
            const float dphi = dd * (w1 - dd * (w2 - dd * w3));
            if constexpr (TIER < NTIER)
                sincos_series<TierF32<TIER>::KD>(dphi, sphi[sc], cphi[sc]);
            else
                sincos_f32(dphi, sphi[sc], cphi[sc]);
            fonfs[sc] = (fc - dd * fdc) * fstar_inv;  // f(xi_c - d), dropping fddot d^2/2 (< 1e-22 Hz)
        }
    } else {
#pragma unroll
        for (int sc = 0; sc < 3; sc++) {
            const double xi = xic - (double)d[sc];
            const float xf = (float)xi;
            float fi = sf.f0 + sf.fdot * xf + 0.5f * sf.fddot * xf * xf;
            fi *= 1.0f + v3[sc] * INV_CF;
            fonfs[sc] = fi * fstar_inv;
            double argS = s.phi0 + s.omq * t + PI * s.fdot * (xi * xi) + (1.0 / 3.0) * PI * s.fddot * (xi * xi * xi);
            argS += term3[sc];
            const double Phi = s.om * (kPc + (double)d[sc]) - argS;
            sincos_rev(Phi * INV_TWO_PI, sphi[sc], cphi[sc]);
        }
    }
    const float kr12 = sf.k[0] * r[0] + sf.k[1] * r[1] + sf.k[2] * r[2];
    const float kr13 = sf.k[0] * r[3] + sf.k[1] * r[4] + sf.k[2] * r[5];
    const float kr23 = sf.k[0] * r[6] + sf.k[1] * r[7] + sf.k[2] * r[8];
    const float ur12 = sf.u[0] * r[0] + sf.u[1] * r[1] + sf.u[2] * r[2];
    const float vr12 = sf.v[0] * r[0] + sf.v[1] * r[1];  // v_z = 0
    const float ur13 = sf.u[0] * r[3] + sf.u[1] * r[4] + sf.u[2] * r[5];
    const float vr13 = sf.v[0] * r[3] + sf.v[1] * r[4];
    const float ur23 = sf.u[0] * r[6] + sf.u[1] * r[7] + sf.u[2] * r[8];
    const float vr23 = sf.v[0] * r[6] + sf.v[1] * r[7];
    const float ep12 = vr12 * vr12 - ur12 * ur12, ex12 = 2.0f * ur12 * vr12;
    const float ep23 = vr23 * vr23 - ur23 * ur23, ex23 = 2.0f * ur23 * vr23;
    const float ep31 = vr13 * vr13 - ur13 * ur13, ex31 = 2.0f * ur13 * vr13;
    const float2 A12 = make_float2(ep12 * sf.DP.x + ex12 * sf.DC.x, ep12 * sf.DP.y + ex12 * sf.DC.y);
    const float2 A23 = make_float2(ep23 * sf.DP.x + ex23 * sf.DC.x, ep23 * sf.DP.y + ex23 * sf.DC.y);
    const float2 A31 = make_float2(ep31 * sf.DP.x + ex31 * sf.DC.x, ep31 * sf.DP.y + ex31 * sf.DC.y);
    // e^{-i Phi_s} = (cphi, -sphi)
    const float2 E0 = make_float2(cphi[0], -sphi[0]), E1 = make_float2(cphi[1], -sphi[1]),
                 E2 = make_float2(cphi[2], -sphi[2]);
    auto Tf = [](float u) {
        if constexpr (TIER < NTIER) {
            float re, im;
            transfer_series<TierF32<TIER>::KR, TierF32<TIER>::KI>(u, re, im);
            return make_float2(re, im);
        } else {
            return transfer_general(u);
        }
    };
    // links (ij, arm, delayed spacecraft, k.r_ij)  (gbgpu.py:509-516)
    const float2 G12 = cmul(Tf(fonfs[0] * (1.0f + kr12)), cmul(E0, A12));
    const float2 G23 = cmul(Tf(fonfs[1] * (1.0f + kr23)), cmul(E1, A23));
    const float2 G31 = cmul(Tf(fonfs[2] * (1.0f - kr13)), cmul(E2, A31));
    const float2 G21 = cmul(Tf(fonfs[1] * (1.0f - kr12)), cmul(E1, A12));
    const float2 G32 = cmul(Tf(fonfs[2] * (1.0f - kr23)), cmul(E2, A23));
    const float2 G13 = cmul(Tf(fonfs[0] * (1.0f + kr13)), cmul(E0, A31));
    // TDI factor: omL = f(t)/f*
    float omL, sL, cL;
    double omLd;  // the same omL in FP64, for the centre phase
    if constexpr (TIER < NTIER) {
        const float tf = (float)t;
        const float f = sf.f0 + sf.fdot * tf + 0.5f * sf.fddot * tf * tf;
        omL = f * fstar_inv;
        omLd = (double)omL;
        sincos_series<TierF32<TIER>::KS>(omL, sL, cL);
    } else {
        // omL can sit near a null of sin(omL) (f ~ pi f*), where an FP32
        // argument would cost relative accuracy: reduce it in FP64.
        const double fd = s.f0 + s.fdot * t + 0.5 * s.fddot * t * t;
        omLd = fd * fstar_inv_d;
        omL = (float)omLd;
        sincos_rev(omLd * INV_TWO_PI, sL, cL);
    }
    const float m = 4.0f * omL * sL * scale;
    if constexpr (!THIRD) {
        const double kPc2 = (s.k[0] * T(ROW_PC, n) + s.k[1] * T(ROW_PC + 1, n) + s.k[2] * T(ROW_PC + 2, n)) * INV_C;
        const double xc = t - kPc2;
        const double argSc = s.phi0 + s.omq * t + PI * s.fdot * (xc * xc) + (1.0 / 3.0) * PI * s.fddot * (xc * xc * xc);
        float sC, cC;
        sincos_rev((s.om * kPc2 - argSc + omLd) * INV_TWO_PI, sC, cC);
        F = make_float2(m * cC, -m * sC);  // fctr2 e^{-i Phi_c}
    } else {
        F = make_float2(m * cL, -m * sL);  // fctr2
    }
    const float2 fctr = make_float2(cL, -sL);
    X = cadd(csub(G21, G31), cmul(csub(G12, G13), fctr));
    Y = cadd(csub(G32, G12), cmul(csub(G23, G21), fctr));
    Z = cadd(csub(G13, G23), cmul(csub(G31, G32), fctr));
}

struct Mixed {
    using CT = float2;
    using Aux = SrcF;
    static __device__ __forceinline__ Aux make_aux(const Src &s) {
        SrcF f;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            f.k[i] = (float)s.k[i];
            f.u[i] = (float)s.u[i];
            f.v[i] = (float)s.v[i];
        }
        const double ia = (s.amp != 0.0) ? 0.25 / s.amp : 0.0;  // link prefactor 1/4 folded in
        f.DP = make_float2((float)(s.DPr * ia), (float)(s.DPi * ia));
        f.DC = make_float2((float)(s.DCr * ia), (float)(s.DCi * ia));
        f.f0 = (float)s.f0;
        f.fdot = (float)s.fdot;
        f.fddot = (float)s.fddot;
        return f;
    }

    template <bool THIRD, int TIER>
    static __device__ __forceinline__ void sample(const Src &s, const Aux &sf, const double *__restrict__ tab, int N,
                                                  int n, double fstar_inv, double scale, const float *v3,
                                                  const double *t3, CT &X, CT &Y, CT &Z, CT &F) {
        slow_xyz_mixed<THIRD, TIER>(s, sf, tab, N, n, fstar_inv, (float)scale, v3, t3, X, Y, Z, F);
    }

    static __device__ __forceinline__ double2 fac(const Src &s) { return make_double2(s.tdr * s.amp, s.tdi * s.amp); }
    static __device__ __forceinline__ double2 up(CT v) { return make_double2((double)v.x, (double)v.y); }
    static const void *twiddles(int N) { return twiddles_f32(N); }
};

}  // namespace gbgpu
