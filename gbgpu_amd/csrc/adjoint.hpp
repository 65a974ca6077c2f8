// Derivatives of the FP64 slow part (source.hpp: slow_xyz_f64) with respect
// to the source parameters, per time sample.
//
// Parameter basis ("sampling basis"), index k:
//   0 ln amp, 1 f0, 2 fdot, 3 fddot, 4 phi0, 5 cos iota, 6 psi, 7 lambda,
//   8 sin beta.
// The extrinsic parameters (0, 4, 5, 6) enter only through the polarisation
// amplitudes DP, DC and the overall phase: the sample is x = DP xP + DC xC,
// where xP (xC) is the response with the arm factor A_ij = (r e+ r) (resp.
// (r ex r)) alone, and phi0 multiplies x by e^{i phi0}.  Their derivatives
// are therefore closed-form combinations of xP and xC.  The intrinsic ones
// (f0, fdot, fddot, lambda, sin beta; "directions" 0..4 below) act through
// the link phases Phi_s, the link arguments u_ij = fonfs_s (1 + k.r_ij), the
// antenna factors and the TDI factor omL, and are taken here as scalar
// tangents of those leaves.
//
// Two consumers:
//   adjoint_sample: reverse mode.  Given the adjoint weights (yA, yE) of the
//     sample (the transformed residual), returns the sensitivities of
//     J = Re(yA A + yE E) to every leaf, chains them into the five intrinsic
//     directions, and returns the complex sums S_P = sum yc xPc,
//     S_C = sum yc xCc for the extrinsic ones.
//   tangent_sample: forward mode.  Returns xP, xC and the five intrinsic
//     tangents dx/dk of A and E (for Fisher matrices).
#pragma once

#include "common.hpp"
#include "fastmath.hpp"
#include "source.hpp"

namespace gbgpu {

constexpr int NDIR = 5;  // intrinsic directions: f0, fdot, fddot, lambda, sin beta

// Derivatives of the sky-frame unit vectors with respect to lambda (l) and
// sin beta (s), from u = (cos th cos l, cos th sin l, -sin th),
// v = (sin l, -cos l, 0), k = (-sin th cos l, -sin th sin l, -cos th),
// th = pi/2 - beta, d(cos th)/ds = 1, d(sin th)/ds = -tan beta.
struct SkyDeriv {
    double kl[3], ks[3], ul[3], us[3], vl[3];
};

__device__ inline SkyDeriv make_sky_deriv(double lam, double beta) {
    SkyDeriv d;
    const double theta = PI / 2 - beta;
    double sinth, costh, sinph, cosph;
    sincos_f64(theta, sinth, costh);
    sincos_f64(lam, sinph, cosph);
    const double tanb = costh / sinth;  // sin beta / cos beta
    d.ul[0] = -costh * sinph;
    d.ul[1] = costh * cosph;
    d.ul[2] = 0.0;
    d.us[0] = cosph;
    d.us[1] = sinph;
    d.us[2] = tanb;
    d.vl[0] = cosph;
    d.vl[1] = sinph;
    d.vl[2] = 0.0;
    d.kl[0] = sinth * sinph;
    d.kl[1] = -sinth * cosph;
    d.kl[2] = 0.0;
    d.ks[0] = tanb * cosph;
    d.ks[1] = tanb * sinph;
    d.ks[2] = -1.0;
    return d;
}

// Derivative series of the transfer factor T(u) = sin(u)/u - i (1 - cos u)/u:
//   Re T = Rs(u^2), Im T = -u Is(u^2)   (fastmath.hpp)
//   d Re T/du = u Rd(u^2),  Rd(v) = sum_j 2 (j+1) R_{j+1} v^j
//   d Im T/du = -Id(u^2),   Id(v) = sum_k (2k+1) I_k v^k
template <int K>
__device__ __forceinline__ double horner_Rd(double v) {
    double p = 2.0 * (K - 1) * SeriesCoef::R[K - 1];
#pragma unroll
    for (int j = K - 3; j >= 0; j--) p = p * v + 2.0 * (j + 1) * SeriesCoef::R[j + 1];
    return p;
}
template <int K>
__device__ __forceinline__ double horner_Id(double v) {
    double p = (2.0 * (K - 1) + 1.0) * SeriesCoef::I[K - 1];
#pragma unroll
    for (int k = K - 2; k >= 0; k--) p = p * v + (2.0 * k + 1.0) * SeriesCoef::I[k];
    return p;
}

// T(u) and dT/du.  Series tiers use the series (same lengths as the
// waveform); tier NTIER uses T = sinc(a) e^{-ia}, a = u/2,
// dT/du = (sinc'(a) - i sinc(a)) e^{-ia} / 2.
template <int TIER>
__device__ __forceinline__ void transfer_and_deriv(double u, double2 &T, double2 &dT) {
    if constexpr (TIER < NTIER) {
        constexpr int KR = TierF64<TIER>::KR, KI = TierF64<TIER>::KI;
        const double v = u * u;
        T = make_double2(horner_R<KR>(v), -u * horner_I<KI>(v));
        dT = make_double2(u * horner_Rd<KR>(v), -horner_Id<KI>(v));
    } else {
        const double a = 0.5 * u;
        double sa, ca;
        sincos_f64(a, sa, ca);
        double sinc, dsinc;
        if (fabs(a) > 1e-3) {
            const double ia = rcp_f64(a);
            sinc = sa * ia;
            dsinc = (ca - sinc) * ia;
        } else {
            const double a2 = a * a;
            sinc = 1.0 - a2 * (1.0 / 6.0 - a2 * (1.0 / 120.0));
            dsinc = a * (-1.0 / 3.0 + a2 * (1.0 / 30.0));
        }
        T = make_double2(sinc * ca, -sinc * sa);
        dT = make_double2(0.5 * (dsinc * ca - sinc * sa), 0.5 * (-dsinc * sa - sinc * ca));
    }
}

// Link table: links 12, 23, 31, 21, 32, 13 (gbgpu.py:509-516).
//   spacecraft whose phase and frequency the link uses, arm (0 = r12,
//   1 = r13, 2 = r23) and the sign of k.r in u = fonfs (1 + sign k.r).
__device__ constexpr int LINK_SC[6] = {0, 1, 2, 1, 2, 0};
__device__ constexpr int LINK_ARM[6] = {0, 2, 1, 0, 2, 1};
__device__ constexpr double LINK_SGN[6] = {1.0, 1.0, -1.0, -1.0, -1.0, 1.0};
// TDI: X = F2 (L1X + F1 L2X) with L1X = G21 - G31, L2X = G12 - G13 and
// cyclic Y, Z (gbgpu.py:352-375).  Per link, the channel (X=0, Y=1, Z=2) and
// sign with which it enters L1 and L2:
//   12: -Y in L1, +X in L2    23: -Z, +Y    31: -X, +Z
//   21: +X in L1, -Y in L2    32: +Y, -Z    13: +Z, -X
__device__ constexpr int LINK_C1[6] = {1, 2, 0, 0, 1, 2};
__device__ constexpr double LINK_S1[6] = {-1.0, -1.0, -1.0, 1.0, 1.0, 1.0};
__device__ constexpr int LINK_C2[6] = {0, 1, 2, 1, 2, 0};
__device__ constexpr double LINK_S2[6] = {1.0, 1.0, 1.0, -1.0, -1.0, -1.0};

// Per-sample leaves shared by the reverse and forward consumers.
struct Leaves {
    double t;
    double xi[3], fi[3], fonfs[3];
    double kPd[2][3];  // d(k.P_s/c)/d(lambda, sin beta)
    double2 E[3];      // e^{-i Phi_s}
    double kr[3], krd[2][3];
    double ep[3], ex[3], epd[2][3], exd[2][3];
    double omL;
    double2 F1, F2, dF2;  // TDI factors and dF2/d omL
};

__device__ __forceinline__ double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

template <int TIER>
__device__ __forceinline__ void make_leaves(const Src &s, const SkyDeriv &sd, const double *__restrict__ tab, int N,
                                            int n, double fsi, double scale, Leaves &lv) {
    constexpr double INV_C = 1.0 / C_SI;
    const double t = tab[ROW_T * N + n];
    lv.t = t;
    double Pc[3];
#pragma unroll
    for (int i = 0; i < 3; i++) Pc[i] = tab[(ROW_PC + i) * N + n];
    const double kPc = dot3(s.k, Pc), kPcl = dot3(sd.kl, Pc), kPcs = dot3(sd.ks, Pc);
#pragma unroll
    for (int sc = 0; sc < 3; sc++) {
        double dP[3];
#pragma unroll
        for (int i = 0; i < 3; i++) dP[i] = tab[(ROW_DP + 3 * sc + i) * N + n];
        const double kP = (kPc + dot3(s.k, dP)) * INV_C;
        lv.kPd[0][sc] = (kPcl + dot3(sd.kl, dP)) * INV_C;
        lv.kPd[1][sc] = (kPcs + dot3(sd.ks, dP)) * INV_C;
        const double xi = t - kP;
        lv.xi[sc] = xi;
        lv.fi[sc] = s.f0 + s.fdot * xi + 0.5 * s.fddot * xi * xi;
        lv.fonfs[sc] = lv.fi[sc] * fsi;
        const double argS = s.phi0 + s.omq * t + PI * s.fdot * (xi * xi) + (1.0 / 3.0) * PI * s.fddot * (xi * xi * xi);
        double sp, cp;
        sincos_f64(s.om * kP - argS, sp, cp);
        lv.E[sc] = make_double2(cp, -sp);
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double r[3];
#pragma unroll
        for (int i = 0; i < 3; i++) r[i] = tab[(ROW_R + 3 * a + i) * N + n];
        lv.kr[a] = dot3(s.k, r);
        lv.krd[0][a] = dot3(sd.kl, r);
        lv.krd[1][a] = dot3(sd.ks, r);
        const double ur = dot3(s.u, r), vr = dot3(s.v, r);
        const double url = dot3(sd.ul, r), urs = dot3(sd.us, r), vrl = dot3(sd.vl, r);
        lv.ep[a] = vr * vr - ur * ur;
        lv.ex[a] = 2.0 * ur * vr;
        lv.epd[0][a] = 2.0 * (vr * vrl - ur * url);
        lv.epd[1][a] = -2.0 * ur * urs;  // dv/ds = 0
        lv.exd[0][a] = 2.0 * (url * vr + ur * vrl);
        lv.exd[1][a] = 2.0 * urs * vr;
    }
    const double f = s.f0 + s.fdot * t + 0.5 * s.fddot * t * t;
    const double omL = f * fsi;
    lv.omL = omL;
    double sL, cL;
    if constexpr (TIER < NTIER)
        sincos_series<TierF64<TIER>::KS>(omL, sL, cL);
    else
        sincos_f64(omL, sL, cL);
    lv.F1 = make_double2(cL, -sL);
    const double m = 4.0 * omL * sL * scale;
    const double dm = 4.0 * scale * (sL + omL * cL);
    lv.F2 = make_double2(m * cL, -m * sL);
    // d/d omL of m e^{-i omL} = dm e^{-i omL} - i F2
    lv.dF2 = make_double2(dm * cL - m * sL, -dm * sL - m * cL);
}

// Accumulators of the reverse pass (per thread, summed over its samples).
struct GradAcc {
    double2 SP, SC;  // sum_c y_c xP_c, sum_c y_c xC_c
    double g[NDIR];  // intrinsic directions
};

// Adds this sample's contribution.  yA, yE: adjoint weights of the sample's
// A and E values (J = Re(yA A + yE E)); DP, DC: polarisation amplitudes.
// The spacecraft are visited one at a time (their two links each), so only
// the arm quantities, the TDI factors and the accumulators stay live.
template <int TIER>
__device__ __forceinline__ void adjoint_sample(const Src &s, const SkyDeriv &sd, const double *__restrict__ tab, int N,
                                               int n, double fsi, double scale, double2 yA, double2 yE,
                                               GradAcc &acc) {
    constexpr double INV_C = 1.0 / C_SI;
    const double t = tab[ROW_T * N + n];
    const double2 DP = make_double2(s.DPr, s.DPi), DC = make_double2(s.DCr, s.DCi);
    // TDI factors F1 = e^{-i omL}, F2 = 4 omL sin(omL) scale e^{-i omL}, dF2/d omL
    double2 F1, F2, dF2;
    {
        const double f = s.f0 + s.fdot * t + 0.5 * s.fddot * t * t;
        const double omL = f * fsi;
        double sL, cL;
        if constexpr (TIER < NTIER)
            sincos_series<TierF64<TIER>::KS>(omL, sL, cL);
        else
            sincos_f64(omL, sL, cL);
        F1 = make_double2(cL, -sL);
        const double m = 4.0 * omL * sL * scale;
        const double dm = 4.0 * scale * (sL + omL * cL);
        F2 = make_double2(m * cL, -m * sL);
        dF2 = make_double2(dm * cL - m * sL, -dm * sL - m * cL);
    }
    // channel weights: J = Re(wX X + wY Y + wZ Z), A = (Z - X)/sqrt2, E = (X - 2Y + Z)/sqrt6
    double2 w[3];
    w[0] = csub(cscale(yE, INV_SQRT6), cscale(yA, INV_SQRT2));
    w[1] = cscale(yE, -2.0 * INV_SQRT6);
    w[2] = cadd(cscale(yA, INV_SQRT2), cscale(yE, INV_SQRT6));
    // arms r12, r13, r23: k.r and antenna factors with their lambda / sin beta derivatives
    double kr[3], krl[3], krs[3], ep[3], ex[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double r[3];
#pragma unroll
        for (int i = 0; i < 3; i++) r[i] = tab[(ROW_R + 3 * a + i) * N + n];
        kr[a] = dot3(s.k, r);
        krl[a] = dot3(sd.kl, r);
        krs[a] = dot3(sd.ks, r);
        const double ur = dot3(s.u, r), vr = dot3(s.v, r);
        ep[a] = vr * vr - ur * ur;
        ex[a] = 2.0 * ur * vr;
    }
    double2 Q1 = make_double2(0.0, 0.0), Q2 = Q1, SP = Q1, SC = Q1;
    double sep[3] = {0.0, 0.0, 0.0}, sex[3] = {0.0, 0.0, 0.0};
    double g0 = 0.0, g1 = 0.0, g2 = 0.0, gl = 0.0, gs = 0.0;
    double Pc[3];
#pragma unroll
    for (int i = 0; i < 3; i++) Pc[i] = tab[(ROW_PC + i) * N + n];
    const double kPc = dot3(s.k, Pc), kPcl = dot3(sd.kl, Pc), kPcs = dot3(sd.ks, Pc);
#pragma unroll
    for (int sc = 0; sc < 3; sc++) {
        double dP[3];
#pragma unroll
        for (int i = 0; i < 3; i++) dP[i] = tab[(ROW_DP + 3 * sc + i) * N + n];
        const double kP = (kPc + dot3(s.k, dP)) * INV_C;
        const double kPl = (kPcl + dot3(sd.kl, dP)) * INV_C;
        const double kPs = (kPcs + dot3(sd.ks, dP)) * INV_C;
        const double xi = t - kP;
        const double fi = s.f0 + s.fdot * xi + 0.5 * s.fddot * xi * xi;
        const double fonfs = fi * fsi;
        double2 E;
        {
            const double argS =
                s.phi0 + s.omq * t + PI * s.fdot * (xi * xi) + (1.0 / 3.0) * PI * s.fddot * (xi * xi * xi);
            double sp, cp;
            sincos_f64(s.om * kP - argS, sp, cp);
            E = make_double2(cp, -sp);
        }
        double U = 0.0, V = 0.0;
        // the two links that use this spacecraft's phase and frequency
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int L = h == 0 ? sc : (sc == 0 ? 5 : sc + 2);  // sc0: 12, 13; sc1: 23, 21; sc2: 31, 32
            const int arm = LINK_ARM[L];
            const double sg = LINK_SGN[L];
            double2 T, dT;
            transfer_and_deriv<TIER>(fonfs * (1.0 + sg * kr[arm]), T, dT);
            const double2 e = cscale(cmul(T, E), 0.25);
            const double2 ed = cscale(cmul(dT, E), 0.25);
            const double2 Aarm = make_double2(ep[arm] * DP.x + ex[arm] * DC.x, ep[arm] * DP.y + ex[arm] * DC.y);
            const double2 G = cmul(e, Aarm);
            const double2 c1 = cscale(w[LINK_C1[L]], LINK_S1[L]);
            const double2 c2 = cscale(w[LINK_C2[L]], LINK_S2[L]);
            Q1 = cadd(Q1, cmul(c1, G));
            Q2 = cadd(Q2, cmul(c2, G));
            // Omega = F2 (c1 + c2 F1): coefficient of G_L in sum_ch w_ch X_ch
            const double2 Om = cmul(F2, cadd(c1, cmul(c2, F1)));
            const double2 Oe = cmul(Om, e);
            SP = cadd(SP, cscale(Oe, ep[arm]));
            SC = cadd(SC, cscale(Oe, ex[arm]));
            const double2 OA = cmul(Om, Aarm);
            const double su = cmul(OA, ed).x;  // dJ/du_L
            U += su * (1.0 + sg * kr[arm]);
            V += cmul(OA, e).y;  // dJ/dPhi_s (de/dPhi = -i e)
            gl += su * fonfs * sg * krl[arm];
            gs += su * fonfs * sg * krs[arm];
            sep[arm] += Oe.x * DP.x - Oe.y * DP.y;  // Re(Oe DP)
            sex[arm] += Oe.x * DC.x - Oe.y * DC.y;
        }
        const double Uf = U * fsi;
        g0 += Uf - TWO_PI * xi * V;
        g1 += xi * (Uf - PI * xi * V);
        g2 += 0.5 * xi * xi * (Uf - (2.0 * PI / 3.0) * xi * V);
        // d/d(k.P_s/c): fonfs through xi, Phi = om kP - argS(t - kP) -> 2 pi f(xi)
        const double W = -(s.fdot + s.fddot * xi) * Uf + TWO_PI * fi * V;
        gl += W * kPl;
        gs += W * kPs;
    }
    // dJ/d omL = Re[dF2 (Q1 + F1 Q2) + F2 Q2 (-i F1)]
    const double somL = cmul(dF2, cadd(Q1, cmul(F1, Q2))).x + cmul(cmul(F2, Q2), cmul_mi(F1)).x;
    g0 += somL * fsi;
    g1 += somL * t * fsi;
    g2 += somL * 0.5 * t * t * fsi;
    // antenna factors: recompute their lambda / sin beta derivatives per arm
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double r[3];
#pragma unroll
        for (int i = 0; i < 3; i++) r[i] = tab[(ROW_R + 3 * a + i) * N + n];
        const double ur = dot3(s.u, r), vr = dot3(s.v, r);
        const double url = dot3(sd.ul, r), urs = dot3(sd.us, r), vrl = dot3(sd.vl, r);
        gl += sep[a] * 2.0 * (vr * vrl - ur * url) + sex[a] * 2.0 * (url * vr + ur * vrl);
        gs += sep[a] * (-2.0 * ur * urs) + sex[a] * 2.0 * urs * vr;  // dv/ds = 0
    }
    acc.SP = cadd(acc.SP, SP);
    acc.SC = cadd(acc.SC, SC);
    acc.g[0] += g0;
    acc.g[1] += g1;
    acc.g[2] += g2;
    acc.g[3] += gl;
    acc.g[4] += gs;
}

// Forward tangents of one sample: xP, xC (A, E) and the five intrinsic
// tangents of x = DP xP + DC xC (A, E).
struct TangentOut {
    double2 PA, PE, CA, CE;
    double2 DA[NDIR], DE[NDIR];
};

template <int TIER>
__device__ __forceinline__ void tangent_sample(const Src &s, const SkyDeriv &sd, const double *__restrict__ tab, int N,
                                               int n, double fsi, double scale, TangentOut &o) {
    Leaves lv;
    make_leaves<TIER>(s, sd, tab, N, n, fsi, scale, lv);
    const double2 DP = make_double2(s.DPr, s.DPi), DC = make_double2(s.DCr, s.DCi);
    double2 e[6], ed[6];
#pragma unroll
    for (int L = 0; L < 6; L++) {
        const int sc = LINK_SC[L], arm = LINK_ARM[L];
        double2 T, dT;
        transfer_and_deriv<TIER>(lv.fonfs[sc] * (1.0 + LINK_SGN[L] * lv.kr[arm]), T, dT);
        e[L] = cscale(cmul(T, lv.E[sc]), 0.25);
        ed[L] = cscale(cmul(dT, lv.E[sc]), 0.25);
    }
    // channel from link values g[6]: ch = F2 (sum_L s1 g_L [c1 == ch] + F1 sum_L s2 g_L [c2 == ch])
    auto tdi = [&](const double2 *g, double2 *xyz, double2 *l1, double2 *l2) {
#pragma unroll
        for (int c = 0; c < 3; c++) l1[c] = l2[c] = make_double2(0.0, 0.0);
#pragma unroll
        for (int L = 0; L < 6; L++) {
            l1[LINK_C1[L]] = cadd(l1[LINK_C1[L]], cscale(g[L], LINK_S1[L]));
            l2[LINK_C2[L]] = cadd(l2[LINK_C2[L]], cscale(g[L], LINK_S2[L]));
        }
#pragma unroll
        for (int c = 0; c < 3; c++) xyz[c] = cmul(lv.F2, cadd(l1[c], cmul(lv.F1, l2[c])));
    };
    auto to_ae = [](const double2 *xyz, double2 &A, double2 &E) {
        A = cscale(csub(xyz[2], xyz[0]), INV_SQRT2);
        E = cscale(cadd(csub(xyz[0], cscale(xyz[1], 2.0)), xyz[2]), INV_SQRT6);
    };
    double2 g[6], xyz[3], l1[3], l2[3];
    // xP, xC
#pragma unroll
    for (int L = 0; L < 6; L++) g[L] = cscale(e[L], lv.ep[LINK_ARM[L]]);
    tdi(g, xyz, l1, l2);
    to_ae(xyz, o.PA, o.PE);
#pragma unroll
    for (int L = 0; L < 6; L++) g[L] = cscale(e[L], lv.ex[LINK_ARM[L]]);
    tdi(g, xyz, l1, l2);
    to_ae(xyz, o.CA, o.CE);
    // full links (for the omL terms): G = e A_arm
    double2 Aarm[3];
#pragma unroll
    for (int a = 0; a < 3; a++)
        Aarm[a] = make_double2(lv.ep[a] * DP.x + lv.ex[a] * DC.x, lv.ep[a] * DP.y + lv.ex[a] * DC.y);
    double2 G[6];
#pragma unroll
    for (int L = 0; L < 6; L++) G[L] = cmul(e[L], Aarm[LINK_ARM[L]]);
    double2 L1[3], L2[3];
    tdi(G, xyz, L1, L2);
    const double t = lv.t;
#pragma unroll
    for (int k = 0; k < NDIR; k++) {
        // leaf tangents
        double domL = 0.0, dfon[3], dphi[3], dkr[3] = {0.0, 0.0, 0.0}, dep[3] = {0.0, 0.0, 0.0},
               dex[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int sc = 0; sc < 3; sc++) {
            const double xi = lv.xi[sc];
            if (k == 0) {
                dfon[sc] = fsi;
                dphi[sc] = -TWO_PI * xi;
            } else if (k == 1) {
                dfon[sc] = xi * fsi;
                dphi[sc] = -PI * xi * xi;
            } else if (k == 2) {
                dfon[sc] = 0.5 * xi * xi * fsi;
                dphi[sc] = -(PI / 3.0) * xi * xi * xi;
            } else {
                const double dkP = lv.kPd[k - 3][sc];
                dfon[sc] = -(s.fdot + s.fddot * xi) * fsi * dkP;
                dphi[sc] = TWO_PI * lv.fi[sc] * dkP;
            }
        }
        if (k == 0) domL = fsi;
        if (k == 1) domL = t * fsi;
        if (k == 2) domL = 0.5 * t * t * fsi;
        if (k >= 3) {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                dkr[a] = lv.krd[k - 3][a];
                dep[a] = lv.epd[k - 3][a];
                dex[a] = lv.exd[k - 3][a];
            }
        }
        double2 dG[6];
#pragma unroll
        for (int L = 0; L < 6; L++) {
            const int sc = LINK_SC[L], arm = LINK_ARM[L];
            const double sg = LINK_SGN[L];
            const double du = dfon[sc] * (1.0 + sg * lv.kr[arm]) + lv.fonfs[sc] * sg * dkr[arm];
            // de = ed du - i e dPhi
            const double2 de = cadd(cscale(ed[L], du), cscale(cmul_mi(e[L]), dphi[sc]));
            const double2 dA = make_double2(dep[arm] * DP.x + dex[arm] * DC.x, dep[arm] * DP.y + dex[arm] * DC.y);
            dG[L] = cadd(cmul(de, Aarm[arm]), cmul(e[L], dA));
        }
        double2 dxyz[3], dl1[3], dl2[3];
        tdi(dG, dxyz, dl1, dl2);
        // omL terms: dF2 (L1 + F1 L2) + F2 L2 (-i F1)
        if (k < 3) {
            const double2 dF1 = cmul_mi(lv.F1);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double2 extra = cadd(cmul(lv.dF2, cadd(L1[c], cmul(lv.F1, L2[c]))), cmul(lv.F2, cmul(L2[c], dF1)));
                dxyz[c] = cadd(dxyz[c], cscale(extra, domL));
            }
        }
        to_ae(dxyz, o.DA[k], o.DE[k]);
    }
}

}  // namespace gbgpu
