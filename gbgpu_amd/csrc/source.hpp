// Per-binary constants and the per-sample FastGB slow part (device code).
//
// Equations restate src/gbgpu/gbgpu.py:243-316 (per-binary set-up),
// :399-540 (slow part) and :348-388 (TDI X, Y, Z), with the third-body hooks
// of src/gbgpu/thirdbody.py:155-226 and :228-428.
#pragma once

#include "common.hpp"
#include "fastmath.hpp"

namespace gbgpu {

// Per-binary constants, computed once per binary and broadcast through LDS.
struct Src {
    double amp, f0, fdot, fddot, phi0;
    double om;   // 2 pi f0                          (gbgpu.py:464)
    double omq;  // om - 2 pi q / T                  (gbgpu.py:462-463, :497)
    double k[3], u[3], v[3];                      // (gbgpu.py:254-258)
    double DPr, DPi, DCr, DCi;                    // (gbgpu.py:250-251)
    double A2, varpi, e2, n2, T2;                 // third body (thirdbody.py:39-80)
    double tdr, tdi;                              // TDI2 factor, (1, 0) for TDI1
    long long start;                              // rint(f0 T) - N/2 (gbgpu.py:296, :387, :462)
    int tier;   // series tier (fastmath.hpp TIER_UMAX) from u_max = 2 max_t |f(t)| / f*; NTIER = none
};

__device__ __forceinline__ double fmax_(double a, double b) { return a > b ? a : b; }

__device__ inline Src make_src(const double *__restrict__ params, int num_params, long long B, long long b,
                               int N, double T, int tdi2, double armlength) {
    Src s;
    const double amp = params[0 * B + b];
    const double f0 = params[1 * B + b];
    const double fdot = params[2 * B + b];
    const double fddot = params[3 * B + b];
    const double phi0 = params[4 * B + b];
    const double iota = params[5 * B + b];
    const double psi = params[6 * B + b];
    const double lam = params[7 * B + b];
    const double beta = params[8 * B + b];
    s.amp = amp;
    s.f0 = f0;
    s.fdot = fdot;
    s.fddot = fddot;
    s.phi0 = phi0;
    const double theta = PI / 2 - beta;
    const double cosiota = cos(iota);
    double sinps, cosps;
    sincos(2.0 * psi, &sinps, &cosps);
    const double Aplus = amp * (1.0 + cosiota * cosiota);
    const double Across = -2.0 * amp * cosiota;
    // DP = Aplus cos2psi - i Across sin2psi ; DC = -Aplus sin2psi - i Across cos2psi
    s.DPr = Aplus * cosps;
    s.DPi = -Across * sinps;
    s.DCr = -Aplus * sinps;
    s.DCi = -Across * cosps;
    double sinth, costh, sinph, cosph;
    sincos(theta, &sinth, &costh);
    sincos(lam, &sinph, &cosph);
    s.u[0] = costh * cosph;
    s.u[1] = costh * sinph;
    s.u[2] = -sinth;
    s.v[0] = sinph;
    s.v[1] = -cosph;
    s.v[2] = 0.0;
    s.k[0] = -sinth * cosph;
    s.k[1] = -sinth * sinph;
    s.k[2] = -costh;
    const double q = rint(f0 * T);  // round half to even, as numpy.rint
    s.om = 2.0 * PI * f0;
    s.omq = s.om - 2.0 * PI * (q / T);
    s.start = (long long)q - (long long)(N / 2);
    if (num_params >= GBGPU_NPARAMS_THIRD) {
        s.A2 = params[9 * B + b];
        s.varpi = params[10 * B + b];
        s.e2 = params[11 * B + b];
        s.n2 = params[12 * B + b];
        s.T2 = params[13 * B + b];
    } else {
        s.A2 = s.varpi = s.e2 = s.n2 = s.T2 = 0.0;
    }
    {
        // largest |f(t)| over [0, T]: ends and the stationary point of the quadratic
        const double fT = f0 + fdot * T + 0.5 * fddot * T * T;
        double fmax = fmax_(fabs(f0), fabs(fT));
        if (fddot != 0.0) {
            const double ts = -fdot / fddot;
            if (ts > 0.0 && ts < T) fmax = fmax_(fmax, fabs(f0 + fdot * ts + 0.5 * fddot * ts * ts));
        }
        const double fstar = C_SI / (armlength * 2 * PI);
        // 1e-3 margin covers the light-time offsets (|xi - t| < 600 s) and any
        // third-body Doppler factor (|v|/c < 1e-3)
        const double umax = 2.0 * fmax * 1.001 / fstar;
        s.tier = NTIER;
        for (int t = NTIER - 1; t >= 0; t--)
            if (umax <= TIER_UMAX[t]) s.tier = t;
    }
    if (tdi2) {
        // 2i sin(2 wL) e^{-2 i wL}, wL = 2 pi f0 L / c   (gbgpu.py:301-304)
        const double omegaL = 2 * PI * f0 * (armlength / C_SI);
        double s2, c2;
        sincos(2 * omegaL, &s2, &c2);
        const double sn = sin(2 * omegaL);
        // 2i * sn * (c2 - i s2) = 2 sn s2 + 2i sn c2
        s.tdr = 2.0 * sn * s2;
        s.tdi = 2.0 * sn * c2;
    } else {
        s.tdr = 1.0;
        s.tdi = 0.0;
    }
    return s;
}

// ------------------------------------------------------------ third body ---
// Mikkola inverse of Kepler's equation, restating src/gbgpu/thirdbody.py:228-320
// element by element (including its branch structure).
__device__ inline double kepler_u(double l, double e) {
    const bool neg = l < 0;
    l = neg ? -l : l;
    const bool over2pi = l > 2 * PI;
    double mult = 0.0;
    if (over2pi) {
        mult = floor(l / (2 * PI));
        l -= mult * 2.0 * PI;
    }
    const bool overpi = l > PI;
    if (overpi) l = 2.0 * PI - l;
    const double alpha = (1.0 - e) / (4.0 * e + 0.5);
    const double beta = 0.5 * l / (4.0 * e + 0.5);
    double z = sqrt(beta * beta + alpha * alpha * alpha);
    z = neg ? beta - z : beta + z;
    z = (z < 0) ? -pow(-z, 1.0 / 3.0) : pow(z, 1.0 / 3.0);
    const double s = z - alpha / z;
    const double w = s - 0.078 * s * s * s * s * s / (1.0 + e);
    const double u0 = l + e * (3.0 * w - 4.0 * w * w * w);
    double su0, cu0;
    sincos_f64(u0, su0, cu0);
    const double f = u0 - e * su0 - l;
    const double f1 = 1.0 - e * cu0;
    double f2 = u0 - f - l;
    double f3 = 1.0 - f1;
    double f4 = -f2;
    f2 *= 0.5;
    f3 *= 0.166666666666667;
    f4 *= 0.0416666666666667;
    const double u1 = -f / f1;
    const double u2 = -f / (f1 + f2 * u1);
    const double u3 = -f / (f1 + f2 * u2 + f3 * u2 * u2);
    const double u4 = -f / (f1 + f2 * u3 + f3 * u3 * u3 + f4 * u3 * u3 * u3);
    double u = u0 + u4;
    if (overpi) u = 2.0 * PI - u;
    if (over2pi) u = 2.0 * PI * mult + u;
    if (neg) u = -u;
    return u;
}

// Line-of-sight velocity (m/s), src/gbgpu/thirdbody.py:322-380.  A circular
// orbit (e2 = 0) needs no Kepler solve: the eccentric anomaly is the mean
// anomaly, which the solver would return up to rounding.
__device__ inline double third_vlos(const Src &s, double t) {
    if (s.e2 == 0.0) {
        double sn, cs;
        sincos_f64(s.n2 * (t - s.T2) + s.varpi, sn, cs);
        return s.A2 * sn;
    }
    double u = kepler_u(s.n2 * (t - s.T2), s.e2);
    if (s.e2 > 1e-6) {
        const double bb = (1.0 - sqrt(1.0 - s.e2 * s.e2)) / s.e2;
        double su, cu;
        sincos_f64(u, su, cu);
        u += 2.0 * atan2(bb * su, 1.0 - bb * cu);
    }
    double sn, cs, sw, cw;
    sincos_f64(u + s.varpi, sn, cs);
    sincos_f64(s.varpi, sw, cw);
    return s.A2 * (sn + s.e2 * sw);
}

// ------------------------------------------------------ FP64 slow part ---
struct SampleGeom {
    double t;
    double kP[3];   // k . P_s / c
    double kPc;     // k . Pc / c (guiding centre)
    double d[3];    // k . dP_s / c (offsets from the guiding centre)
    double kr12, kr13, kr23;
    double ur12, vr12, ur13, vr13, ur23, vr23;
};

// Orbit-table rows (see orbits.py: orbit_table): t, guiding centre Pc,
// offsets dP_s = P_s - Pc, arm vectors r12, r13, r23.
enum { ROW_T = 0, ROW_PC = 1, ROW_DP = 4, ROW_R = 13, TAB_ROWS = 22 };

// Word 3 of a raw buffer resource: 32-bit data format, no swizzle.
constexpr int BUF_DWORD3 = 0x00020000;

// The same pointer value taken from the first active lane, in scalar
// registers.  For a pointer that every lane of the wavefront holds (a kernel
// argument, or a per-launch table chosen by a wave-uniform index) this is
// the value itself; it lets the compiler prove the uniformity it cannot see
// through a non-inlined call.
__device__ __forceinline__ const double *wave_uniform(const double *p) {
    const unsigned long long v = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return reinterpret_cast<const double *>(((unsigned long long)hi << 32) | lo);
}

// Reads of the (22, N) orbit table through a buffer resource: the row's
// byte offset rides in a scalar register and the sample's in one vector
// register, so no per-row address arithmetic runs on the vector ALU.  The
// table pointer must be the same in every lane of the wavefront (it is a
// per-launch table everywhere); the resource is built from scalar registers.
// Reads past the table's extent return zero instead of faulting.
struct TabRd {
    __amdgpu_buffer_rsrc_t r;
    int row_bytes;
    __device__ __forceinline__ TabRd(const double *tab, int N)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(wave_uniform(tab)), (short)0, TAB_ROWS * N * 8,
                                              BUF_DWORD3)),
          row_bytes(N * 8) {}
    __device__ __forceinline__ double operator()(int row, int n) const {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, n * 8, row * row_bytes, 0));
    }
};

// k . P_s / c for the three spacecraft, as (k.Pc + k.dP_s)/c.
__device__ __forceinline__ void load_kP(const Src &s, const double *__restrict__ tab, int N, int n, double kP[3]) {
    constexpr double INV_C = 1.0 / C_SI;
    const TabRd T(tab, N);
    const double kc = s.k[0] * T(ROW_PC, n) + s.k[1] * T(ROW_PC + 1, n) + s.k[2] * T(ROW_PC + 2, n);
#pragma unroll
    for (int sc = 0; sc < 3; sc++) {
        const int r0 = ROW_DP + 3 * sc;
        const double kd = s.k[0] * T(r0, n) + s.k[1] * T(r0 + 1, n) + s.k[2] * T(r0 + 2, n);
        kP[sc] = (kc + kd) * INV_C;
    }
}

// v = (sin lam, -cos lam, 0) (gbgpu.py:256): its z component is exactly zero
// and is left out of v.r.
__device__ __forceinline__ SampleGeom load_geom(const Src &s, const double *__restrict__ tab, int N, int n) {
    SampleGeom g;
    const TabRd T(tab, N);
    g.t = T(ROW_T, n);
    double r[9];
#pragma unroll
    for (int i = 0; i < 9; i++) r[i] = T(ROW_R + i, n);
    {
        constexpr double INV_C = 1.0 / C_SI;
        const double kc = s.k[0] * T(ROW_PC, n) + s.k[1] * T(ROW_PC + 1, n) + s.k[2] * T(ROW_PC + 2, n);
        g.kPc = kc * INV_C;
#pragma unroll
        for (int sc = 0; sc < 3; sc++) {
            const int r0 = ROW_DP + 3 * sc;
            const double kd = s.k[0] * T(r0, n) + s.k[1] * T(r0 + 1, n) + s.k[2] * T(r0 + 2, n);
            g.kP[sc] = (kc + kd) * INV_C;
            g.d[sc] = kd * INV_C;
        }
    }
    g.kr12 = s.k[0] * r[0] + s.k[1] * r[1] + s.k[2] * r[2];
    g.kr13 = s.k[0] * r[3] + s.k[1] * r[4] + s.k[2] * r[5];
    g.kr23 = s.k[0] * r[6] + s.k[1] * r[7] + s.k[2] * r[8];
    g.ur12 = s.u[0] * r[0] + s.u[1] * r[1] + s.u[2] * r[2];
    g.vr12 = s.v[0] * r[0] + s.v[1] * r[1];
    g.ur13 = s.u[0] * r[3] + s.u[1] * r[4] + s.u[2] * r[5];
    g.vr13 = s.v[0] * r[3] + s.v[1] * r[4];
    g.ur23 = s.u[0] * r[6] + s.u[1] * r[7] + s.u[2] * r[8];
    g.vr23 = s.v[0] * r[6] + s.v[1] * r[7];
    return g;
}

// One transfer-function link: sin(a)/a e^{-i(a + Phi_s)} A_arm, i.e. 4 G_ij of
// gbgpu.py:508-525 (the 1/4 is applied with the TDI factor, an exact power
// of two), with e^{-i Phi_s} = (cphi, -sphi) precomputed.
// sinc(a) e^{-ia} comes from the reduction-free series in u = 2a when the
// binary's tier allows, else from a reduced sine/cosine.
template <int TIER>
__device__ __forceinline__ double2 link_f64(double fonfs, double kr, double cphi, double sphi, double2 Aarm) {
    double tr, ti;
    if constexpr (TIER < NTIER) {
        transfer_series<TierF64<TIER>::KR, TierF64<TIER>::KI>(fonfs * (1.0 + kr), tr, ti);
    } else {
        const double a = 0.5 * fonfs * (1.0 + kr);
        double sa, ca;
        sincos_f64(a, sa, ca);
        const double sinc = (a != 0.0) ? sa * rcp_f64(a) : 1.0;
        tr = sinc * ca;
        ti = -sinc * sa;
    }
    // (tr + i ti)(cphi - i sphi)
    const double2 e = make_double2(tr * cphi + ti * sphi, ti * cphi - tr * sphi);
    return cmul(e, Aarm);
}

// Time-domain X, Y, Z slow part at sample n, before their common factor
// F = fctr2 * amp * 0.5 T / N (gbgpu.py:359-375, :298; amp / amp cancels),
// which is returned separately so that the caller applies it once per
// output channel (A and E for the likelihood).  On the guiding-centre path F
// also carries the phase factor e^{-i Phi_c} shared by all six links, which
// X, Y and Z then leave out: only the products F X, F Y, F Z are defined.
// v3[sc] (third-body v_LOS at xi_sc) and term3[sc] (its phase integral) are
// used only when THIRD.  EXT (derived waveforms): v3[sc] and term3[sc] are
// the frequency f_i and the phase argS of spacecraft sc at this sample as
// the class's hooks returned them (src/gbgpu/gbgpu.py:453, :505).
template <bool THIRD, int TIER, bool EXT = false>
__device__ __forceinline__ void slow_xyz_f64(const Src &s, const SampleGeom &g, double fstar_inv, double scale,
                                             const double *v3, const double *term3, double2 &X, double2 &Y,
                                             double2 &Z, double2 &F) {
    static_assert(!(THIRD && EXT), "third-body and external hooks are exclusive");
    constexpr bool SPLIT = !THIRD && !EXT && TIER < NTIER;
    // TDI transfer (gbgpu.py:352-375): omL = f(t)/f*
    const double f = s.f0 + s.fdot * g.t + 0.5 * s.fddot * g.t * g.t;
    const double omL = f * fstar_inv;
    double sL, cL;
    if constexpr (TIER < NTIER)
        sincos_series<TierF64<TIER>::KS>(omL, sL, cL);
    else
        sincos_f64(omL, sL, cL);
    const double m = omL * sL * scale;  // 4 omL sL scale, times the links' 1/4
    double fonfs[3], cphi[3], sphi[3];
    if constexpr (SPLIT) {
        // Guiding-centre split: with xi_c = t - k.Pc/c and d_s = k.dP_s/c,
        //   Phi_s = Phi_c + 2 pi f(xi_c) d_s - pi f'(xi_c) d_s^2 + (pi/3) fddot d_s^3,
        // exact for the cubic argS.  The offsets (|.| <= 0.29 U_MAX of the
        // tier) take short series; the common e^{-i Phi_c} of every link is
        // factored out of X, Y and Z into F, where it joins the TDI factor's
        // e^{-i omL} in the one range-reduced sine/cosine of Phi_c + omL.
        const double xic = g.t - g.kPc;
        const double fc = s.f0 + s.fdot * xic + 0.5 * s.fddot * xic * xic;
        const double fdc = s.fdot + s.fddot * xic;
        const double argSc =
            s.phi0 + s.omq * g.t + PI * s.fdot * (xic * xic) + (1.0 / 3.0) * PI * s.fddot * (xic * xic * xic);
        double sC, cC;
        sincos_f64_shift(s.om * g.kPc - argSc, omL, sC, cC);
        F = make_double2(m * cC, -m * sC);  // fctr2 e^{-i Phi_c}
        const double w1 = 2.0 * PI * fc, w2 = PI * fdc, w3 = (PI / 3.0) * s.fddot;
#pragma unroll
        for (int sc = 0; sc < 3; sc++) {
            const double dd = g.d[sc];
            const double dphi = dd * (w1 - dd * (w2 - dd * w3));
            sincos_series<TierF64<TIER>::KD>(dphi, sphi[sc], cphi[sc]);
            // f(xi_c - d_s), exact for the quadratic f
            fonfs[sc] = (fc - dd * (fdc - 0.5 * s.fddot * dd)) * fstar_inv;
        }
    } else {
        F = make_double2(m * cL, -m * sL);  // fctr2
#pragma unroll
        for (int sc = 0; sc < 3; sc++) {
            const double xi = g.t - g.kP[sc];
            double fi, argS;
            if constexpr (EXT) {
                fi = v3[sc];
                argS = term3[sc];
            } else {
                fi = s.f0 + s.fdot * xi + 0.5 * s.fddot * xi * xi;
                if (THIRD) fi *= 1.0 + v3[sc] / C_SI;
                argS = s.phi0 + s.omq * g.t + PI * s.fdot * (xi * xi) + (1.0 / 3.0) * PI * s.fddot * (xi * xi * xi);
                if (THIRD) argS += term3[sc];
            }
            fonfs[sc] = fi * fstar_inv;
            const double Phi = s.om * g.kP[sc] - argS;
            sincos_f64(Phi, sphi[sc], cphi[sc]);
        }
    }
    // arm antenna factors A_ij = (r e+ r) DP + (r ex r) DC  (gbgpu.py:467-473)
    // with r e+ r = (v.r)^2 - (u.r)^2 and r ex r = 2 (u.r)(v.r); r31 = -r13.
    const double ep12 = g.vr12 * g.vr12 - g.ur12 * g.ur12, ex12 = 2.0 * g.ur12 * g.vr12;
    const double ep23 = g.vr23 * g.vr23 - g.ur23 * g.ur23, ex23 = 2.0 * g.ur23 * g.vr23;
    const double ep31 = g.vr13 * g.vr13 - g.ur13 * g.ur13, ex31 = 2.0 * g.ur13 * g.vr13;
    const double2 A12 = make_double2(ep12 * s.DPr + ex12 * s.DCr, ep12 * s.DPi + ex12 * s.DCi);
    const double2 A23 = make_double2(ep23 * s.DPr + ex23 * s.DCr, ep23 * s.DPi + ex23 * s.DCi);
    const double2 A31 = make_double2(ep31 * s.DPr + ex31 * s.DCr, ep31 * s.DPi + ex31 * s.DCi);
    // links (ij, arm, spacecraft of the delay, k.r_ij)  (gbgpu.py:509-516)
    const double2 G12 = link_f64<TIER>(fonfs[0], g.kr12, cphi[0], sphi[0], A12);
    const double2 G23 = link_f64<TIER>(fonfs[1], g.kr23, cphi[1], sphi[1], A23);
    const double2 G31 = link_f64<TIER>(fonfs[2], -g.kr13, cphi[2], sphi[2], A31);
    const double2 G21 = link_f64<TIER>(fonfs[1], -g.kr12, cphi[1], sphi[1], A12);
    const double2 G32 = link_f64<TIER>(fonfs[2], -g.kr23, cphi[2], sphi[2], A23);
    const double2 G13 = link_f64<TIER>(fonfs[0], g.kr13, cphi[0], sphi[0], A31);
    const double2 fctr = make_double2(cL, -sL);
    X = cadd(csub(G21, G31), cmul(csub(G12, G13), fctr));
    Y = cadd(csub(G32, G12), cmul(csub(G23, G21), fctr));
    Z = cadd(csub(G13, G23), cmul(csub(G31, G32), fctr));
}

}  // namespace gbgpu
