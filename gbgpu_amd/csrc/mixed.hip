// Mixed-precision FastGB kernels (gfx950).
//
// Precision split:
//   FP64: t_n and spacecraft positions, k.P/c, xi = t - k.P/c, the carrier
//         phase argS = phi0 + (om - Om_q) t + pi fdot xi^2 + pi/3 fddot xi^3
//         (plus the third-body phase integral), Phi = om k.P/c - argS, its
//         reduction to a quarter turn, the start bin rint(f0 T) - N/2, and
//         the <d|h>, <h|h> accumulation against the FP64 data and PSD.
//   FP32: arm vectors, antenna patterns, fonfs, the six transfer functions,
//         the TDI combination, every sine/cosine (fastmath.hpp) and the
//         length-N LDS transform (float2 Stockham, FP32 twiddles).
// Same work decomposition as the FP64 kernels (wave.hip).
#include "common.hpp"
#include "fastmath.hpp"
#include "fft_lds.hpp"
#include "source.hpp"
#include "wave_common.hpp"

namespace gbgpu {

// FP32 copies of the per-binary constants used by the FP32 part.
struct SrcF {
    float k[3], u[3], v[3];
    float2 DP, DC;
    float f0, fdot, fddot;
    float tdr, tdi;
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

__device__ inline SrcF make_srcf(const Src &s) {
    SrcF f;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        f.k[i] = (float)s.k[i];
        f.u[i] = (float)s.u[i];
        f.v[i] = (float)s.v[i];
    }
    // amplitudes relative to amp keep FP32 far from its denormal range; the
    // amp factor is restored in FP64 after the transform
    const double ia = (s.amp != 0.0) ? 1.0 / s.amp : 0.0;
    f.DP = make_float2((float)(s.DPr * ia), (float)(s.DPi * ia));
    f.DC = make_float2((float)(s.DCr * ia), (float)(s.DCi * ia));
    f.f0 = (float)s.f0;
    f.fdot = (float)s.fdot;
    f.fddot = (float)s.fddot;
    f.tdr = (float)s.tdr;
    f.tdi = (float)s.tdi;
    return f;
}

template <int N, int MODE, bool THIRD>
__global__ void __launch_bounds__(Geo<N>::WG) k_wave_mixed(WaveArgs a) {
    using G = Geo<N>;
    constexpr int TPB = G::TPB, BPW = G::BPW, WPB = G::WPB;
    constexpr int NCH = (MODE == MODE_WAVE) ? 3 : 2;
    constexpr int NT = THIRD ? N : 1;
    __shared__ float2 sbuf[BPW * NCH * N];
    __shared__ Src ssrc[BPW];
    __shared__ SrcF ssrcf[BPW];
    __shared__ double stot[BPW * WPB];
    __shared__ float s_v[BPW * 3 * NT];
    __shared__ double s_g[BPW * 3 * NT];
    // xi: in the transform buffer when it is large enough (3 channels)
    constexpr int NXI = (THIRD && NCH == 2) ? N : 1;
    __shared__ double s_xi[BPW * 3 * NXI];

    const int tid = threadIdx.x;
    const int slot = tid / TPB;
    const int lane = tid % TPB;
    const long long b = (long long)blockIdx.x * BPW + slot;
    const bool valid = b < a.B;
    const float2 *tw = static_cast<const float2 *>(a.tw);

    if (lane == 0 && valid) {
        const Src s0 = make_src(a.params, a.num_params, a.B, b, N, a.T, a.tdi2, a.armlength);
        ssrc[slot] = s0;
        ssrcf[slot] = make_srcf(s0);
    }
    __syncthreads();
    const Src s = ssrc[slot];
    const SrcF sf = ssrcf[slot];
    float2 *buf = sbuf + slot * NCH * N;
    double *tot = stot + slot * WPB;
    const float fstar_inv = (float)a.fstar_inv;
    const float scale = (float)a.scale;

    if constexpr (THIRD) {
        double *xi3 = (NCH == 3) ? reinterpret_cast<double *>(buf) : s_xi + slot * 3 * NXI;
        float *v3 = s_v + slot * 3 * N;
        double *g3 = s_g + slot * 3 * N;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int n = lane + q * TPB;
#pragma unroll
            for (int sc = 0; sc < 3; sc++) {
                double xi = 0.0, gg = 0.0;
                float v = 0.0f;
                if (valid) {
                    const double t = a.tab[n];
                    const double P0 = a.tab[(1 + 3 * sc) * N + n], P1 = a.tab[(2 + 3 * sc) * N + n],
                                 P2 = a.tab[(3 + 3 * sc) * N + n];
                    xi = t - (s.k[0] * P0 + s.k[1] * P1 + s.k[2] * P2) * (1.0 / C_SI);
                    const double vd = third_vlos(s, xi);
                    v = (float)vd;
                    gg = vd * (s.f0 + s.fdot * xi + 0.5 * s.fddot * xi * xi);
                }
                xi3[sc * N + n] = xi;
                v3[sc * N + n] = v;
                g3[sc * N + n] = gg;
            }
        }
        __syncthreads();
        double loc[3][4];
#pragma unroll
        for (int sc = 0; sc < 3; sc++) {
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int n = 4 * lane + i;
                if (n > 0) {
                    const double dtt = xi3[sc * N + n] - xi3[sc * N + n - 1];
                    acc += dtt * (g3[sc * N + n - 1] + g3[sc * N + n]) / 2.0 * PI * 2 / C_SI;
                }
                loc[sc][i] = acc;
            }
        }
        __syncthreads();
#pragma unroll
        for (int sc = 0; sc < 3; sc++) {
            const double pre = group_excl_scan<TPB>(loc[sc][3], lane, tot);
#pragma unroll
            for (int i = 0; i < 4; i++) g3[sc * N + 4 * lane + i] = pre + loc[sc][i];
        }
        __syncthreads();
    }

#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int n = lane + q * TPB;
        float2 X = make_float2(0.0f, 0.0f), Y = X, Z = X;
        if (valid) {
            float v3[3] = {0.f, 0.f, 0.f};
            double t3[3] = {0.0, 0.0, 0.0};
            if constexpr (THIRD) {
#pragma unroll
                for (int sc = 0; sc < 3; sc++) {
                    v3[sc] = s_v[slot * 3 * N + sc * N + n];
                    t3[sc] = s_g[slot * 3 * N + sc * N + n];
                }
            }
            slow_xyz_mixed<THIRD>(s, sf, a.tab, N, n, fstar_inv, scale, v3, t3, X, Y, Z);
        }
        if constexpr (MODE == MODE_WAVE) {
            buf[n] = X;
            buf[N + n] = Y;
            buf[2 * N + n] = Z;
        } else {
            buf[n] = cscale(csub(Z, X), (float)INV_SQRT2);
            buf[N + n] = cscale(cadd(csub(X, cscale(Y, 2.0f)), Z), (float)INV_SQRT6);
        }
    }
    __syncthreads();

    fft_lds<N, NCH>(buf, lane, tw);

    // restore amp (and the TDI2 factor) in FP64
    const double2 tdf = make_double2(s.tdr * s.amp, s.tdi * s.amp);
    auto up = [&](float2 v) { return cmul(make_double2((double)v.x, (double)v.y), tdf); };
    if constexpr (MODE == MODE_WAVE) {
        if (!valid) return;
        double2 *xyz = a.XYZf ? a.XYZf + b * 3 * N : nullptr;
        double2 *aet = a.AETf ? a.AETf + b * 3 * N : nullptr;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int m = lane + q * TPB;
            const int idx = (m + N / 2) & (N - 1);
            const double2 X = up(buf[idx]), Y = up(buf[N + idx]), Z = up(buf[2 * N + idx]);
            if (xyz) {
                xyz[m] = X;
                xyz[N + m] = Y;
                xyz[2 * N + m] = Z;
            }
            if (aet) {
                aet[m] = cscale(csub(Z, X), INV_SQRT2);
                aet[N + m] = cscale(cadd(csub(X, cscale(Y, 2.0)), Z), INV_SQRT6);
                aet[2 * N + m] = cscale(cadd(cadd(X, Y), Z), INV_SQRT3);
            }
        }
        if (lane == 0) a.start_inds[b] = s.start;
    } else {
        double dhr = 0.0, dhi = 0.0, hh = 0.0;
        if (valid) {
            const long long doff = (a.di ? (long long)a.di[b] : 0LL) * a.L;
            const long long noff = (a.ni ? (long long)a.ni[b] : 0LL) * a.L;
            const long long j0 = s.start - a.start_freq_ind;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int m = lane + q * TPB;
                const int idx = (m + N / 2) & (N - 1);
                const long long j = j0 + m;
                if (j < 0 || j >= a.L) continue;
                const double2 hA = up(buf[idx]), hE = up(buf[N + idx]);
                const double2 dA = a.dA[doff + j], dE = a.dE[doff + j];
                const double iA = 1.0 / a.SA[noff + j], iE = 1.0 / a.SE[noff + j];
                dhr += (dA.x * hA.x + dA.y * hA.y) * iA + (dE.x * hE.x + dE.y * hE.y) * iE;
                dhi += (dA.x * hA.y - dA.y * hA.x) * iA + (dE.x * hE.y - dE.y * hE.x) * iE;
                hh += (hA.x * hA.x + hA.y * hA.y) * iA + (hE.x * hE.x + hE.y * hE.y) * iE;
            }
        }
        dhr = group_total<TPB>(dhr, lane, tot);
        dhi = group_total<TPB>(dhi, lane, tot);
        hh = group_total<TPB>(hh, lane, tot);
        if (valid && lane == 0) {
            dhr *= a.df4;
            dhi *= a.df4;
            hh *= a.df4;
            if (a.d_h) a.d_h[b] = make_double2(dhr, dhi);
            if (a.h_h) a.h_h[b] = make_double2(hh, 0.0);
            const double dh_use = a.phase_marg ? sqrt(dhr * dhr + dhi * dhi) : dhr;
            a.loglike[b] = -0.5 * (a.d_d + hh - 2.0 * dh_use);
        }
    }
}

template <int MODE, bool THIRD>
static int launch_mixed(int N, const WaveArgs &a, hipStream_t st) {
#define GBGPU_CASE(NN)                                                                      \
    case NN: {                                                                              \
        const long long grid = (a.B + Geo<NN>::BPW - 1) / Geo<NN>::BPW;                     \
        hipLaunchKernelGGL((k_wave_mixed<NN, MODE, THIRD>), dim3((unsigned)grid),         \
                           dim3(Geo<NN>::WG), 0, st, a);                                    \
        break;                                                                              \
    }
    switch (N) {
        GBGPU_CASE(32)
        GBGPU_CASE(64)
        GBGPU_CASE(128)
        GBGPU_CASE(256)
        GBGPU_CASE(512)
        GBGPU_CASE(1024)
        GBGPU_CASE(2048)
        default:
            set_error("N must be a power of two in [32, 2048]");
            return GBGPU_E_ARG;
    }
#undef GBGPU_CASE
    GBGPU_CHECK_LAUNCH("k_wave_mixed");
    return GBGPU_OK;
}

int mixed_run_wave(const WaveArgs &a, int N, bool third, hipStream_t st) {
    return third ? launch_mixed<MODE_WAVE, true>(N, a, st) : launch_mixed<MODE_WAVE, false>(N, a, st);
}

int mixed_wave_ll(const WaveArgs &a, int N, bool third, hipStream_t st) {
    return third ? launch_mixed<MODE_LL, true>(N, a, st) : launch_mixed<MODE_LL, false>(N, a, st);
}

}  // namespace gbgpu
