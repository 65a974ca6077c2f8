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
// constants), make_aux(), sample<THIRD>() (time-domain X, Y, Z at sample n,
// already scaled by fctr2 * 0.5 T / N), fac() (complex factor applied to
// every output bin in FP64: TDI2 factor, and amp for Mixed) and twiddles().
#pragma once

#include "common.hpp"
#include "fastmath.hpp"
#include "source.hpp"

namespace gbgpu {

struct F64 {
    using CT = double2;
    struct Aux {};
    static __device__ __forceinline__ Aux make_aux(const Src &) { return Aux{}; }

    template <bool THIRD>
    static __device__ __forceinline__ void sample(const Src &s, const Aux &, const double *__restrict__ tab, int N,
                                                  int n, double fstar_inv, double scale, const float *v3,
                                                  const double *t3, CT &X, CT &Y, CT &Z) {
        const SampleGeom g = load_geom(s, tab, N, n);
        double v3d[3] = {0.0, 0.0, 0.0};
        if (THIRD) {
            v3d[0] = v3[0];
            v3d[1] = v3[1];
            v3d[2] = v3[2];
        }
        slow_xyz_f64<THIRD>(s, g, fstar_inv, scale, v3d, t3, X, Y, Z);
    }

    static __device__ __forceinline__ double2 fac(const Src &s) { return make_double2(s.tdr, s.tdi); }
    static __device__ __forceinline__ double2 up(CT v) { return v; }
    static const void *twiddles(int N) { return twiddles_f64(N); }
};

// FP32 copies of the per-binary constants used by the FP32 part.
struct SrcF {
    float k[3], u[3], v[3];
    float2 DP, DC;  // relative to amp
    float f0, fdot, fddot;
};

__device__ __forceinline__ float2 link_f32(float fonfs, float kr, float cphi, float sphi, float2 Aarm) {
    const float a = 0.5f * fonfs * (1.0f + kr);
    float sa, ca;
    sincos_f32(a, sa, ca);
    const float sinc = (a != 0.0f) ? 0.25f * sa * __builtin_amdgcn_rcpf(a) : 0.25f;
    const float er = ca * cphi - sa * sphi;
    const float ei = -(sa * cphi + ca * sphi);
    return cmul(make_float2(sinc * er, sinc * ei), Aarm);
}

template <bool THIRD>
__device__ __forceinline__ void slow_xyz_mixed(const Src &s, const SrcF &sf, const double *__restrict__ tab, int N,
                                               int n, float fstar_inv, float scale, const float *v3,
                                               const double *term3, float2 &X, float2 &Y, float2 &Z) {
    const double t = tab[n];
    float r[9];
#pragma unroll
    for (int i = 0; i < 9; i++) r[i] = (float)tab[(10 + i) * N + n];
    float fonfs[3], cphi[3], sphi[3];
    constexpr double INV_C = 1.0 / C_SI;
#pragma unroll
    for (int sc = 0; sc < 3; sc++) {
        const double P0 = tab[(1 + 3 * sc) * N + n], P1 = tab[(2 + 3 * sc) * N + n], P2 = tab[(3 + 3 * sc) * N + n];
        const double kP = (s.k[0] * P0 + s.k[1] * P1 + s.k[2] * P2) * INV_C;
        const double xi = t - kP;
        const float xf = (float)xi;
        float fi = sf.f0 + sf.fdot * xf + 0.5f * sf.fddot * xf * xf;
        if (THIRD) fi *= 1.0f + v3[sc] * (float)(1.0 / C_SI);
        fonfs[sc] = fi * fstar_inv;
        double argS = s.phi0 + s.omq * t + PI * s.fdot * (xi * xi) + (1.0 / 3.0) * PI * s.fddot * (xi * xi * xi);
        if (THIRD) argS += term3[sc];
        const double Phi = s.om * kP - argS;
        sincos_rev(Phi * INV_TWO_PI, sphi[sc], cphi[sc]);
    }
    const float kr12 = sf.k[0] * r[0] + sf.k[1] * r[1] + sf.k[2] * r[2];
    const float kr13 = sf.k[0] * r[3] + sf.k[1] * r[4] + sf.k[2] * r[5];
    const float kr23 = sf.k[0] * r[6] + sf.k[1] * r[7] + sf.k[2] * r[8];
    const float ur12 = sf.u[0] * r[0] + sf.u[1] * r[1] + sf.u[2] * r[2];
    const float vr12 = sf.v[0] * r[0] + sf.v[1] * r[1] + sf.v[2] * r[2];
    const float ur13 = sf.u[0] * r[3] + sf.u[1] * r[4] + sf.u[2] * r[5];
    const float vr13 = sf.v[0] * r[3] + sf.v[1] * r[4] + sf.v[2] * r[5];
    const float ur23 = sf.u[0] * r[6] + sf.u[1] * r[7] + sf.u[2] * r[8];
    const float vr23 = sf.v[0] * r[6] + sf.v[1] * r[7] + sf.v[2] * r[8];
    const float ep12 = vr12 * vr12 - ur12 * ur12, ex12 = 2.0f * ur12 * vr12;
    const float ep23 = vr23 * vr23 - ur23 * ur23, ex23 = 2.0f * ur23 * vr23;
    const float ep31 = vr13 * vr13 - ur13 * ur13, ex31 = 2.0f * ur13 * vr13;
    const float2 A12 = make_float2(ep12 * sf.DP.x + ex12 * sf.DC.x, ep12 * sf.DP.y + ex12 * sf.DC.y);
    const float2 A23 = make_float2(ep23 * sf.DP.x + ex23 * sf.DC.x, ep23 * sf.DP.y + ex23 * sf.DC.y);
    const float2 A31 = make_float2(ep31 * sf.DP.x + ex31 * sf.DC.x, ep31 * sf.DP.y + ex31 * sf.DC.y);
    const float2 G12 = link_f32(fonfs[0], kr12, cphi[0], sphi[0], A12);
    const float2 G23 = link_f32(fonfs[1], kr23, cphi[1], sphi[1], A23);
    const float2 G31 = link_f32(fonfs[2], -kr13, cphi[2], sphi[2], A31);
    const float2 G21 = link_f32(fonfs[1], -kr12, cphi[1], sphi[1], A12);
    const float2 G32 = link_f32(fonfs[2], -kr23, cphi[2], sphi[2], A23);
    const float2 G13 = link_f32(fonfs[0], kr13, cphi[0], sphi[0], A31);
    const float tf = (float)t;
    const float f = sf.f0 + sf.fdot * tf + 0.5f * sf.fddot * tf * tf;
    const float omL = f * fstar_inv;
    float sL, cL;
    sincos_f32(omL, sL, cL);
    const float2 fctr = make_float2(cL, -sL);
    const float m = 4.0f * omL * sL * scale;
    const float2 fctr2 = make_float2(m * cL, -m * sL);
    X = cmul(fctr2, cadd(csub(G21, G31), cmul(csub(G12, G13), fctr)));
    Y = cmul(fctr2, cadd(csub(G32, G12), cmul(csub(G23, G21), fctr)));
    Z = cmul(fctr2, cadd(csub(G13, G23), cmul(csub(G31, G32), fctr)));
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
        const double ia = (s.amp != 0.0) ? 1.0 / s.amp : 0.0;
        f.DP = make_float2((float)(s.DPr * ia), (float)(s.DPi * ia));
        f.DC = make_float2((float)(s.DCr * ia), (float)(s.DCi * ia));
        f.f0 = (float)s.f0;
        f.fdot = (float)s.fdot;
        f.fddot = (float)s.fddot;
        return f;
    }

    template <bool THIRD>
    static __device__ __forceinline__ void sample(const Src &s, const Aux &sf, const double *__restrict__ tab, int N,
                                                  int n, double fstar_inv, double scale, const float *v3,
                                                  const double *t3, CT &X, CT &Y, CT &Z) {
        slow_xyz_mixed<THIRD>(s, sf, tab, N, n, (float)fstar_inv, (float)scale, v3, t3, X, Y, Z);
    }

    static __device__ __forceinline__ double2 fac(const Src &s) { return make_double2(s.tdr * s.amp, s.tdi * s.amp); }
    static __device__ __forceinline__ double2 up(CT v) { return make_double2((double)v.x, (double)v.y); }
    static const void *twiddles(int N) { return twiddles_f32(N); }
};

}  // namespace gbgpu
