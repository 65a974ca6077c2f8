// Derivatives of the fused waveform + likelihood (FP64, 9-parameter sources).
//
// k_grad (adjoint): per binary, the forward pass builds the A/E spectra and
// the likelihood sums; the residual weights w = (rho d - h)/S (rho = 1, or
// dh/|dh| when the phase is marginalised) are placed on the template's bins,
// multiplied by the TDI factor, and pushed back through the transform (the
// adjoint of the fftshifted DFT is another forward DFT of the conjugated
// weights).  The slow part is then recomputed sample by sample, and the
// reverse pass of adjoint.hpp turns the per-sample adjoint weights into the
// gradient of
//   logL = -1/2 (d_d + <h|h> - 2 Re<d|h>)      (or |<d|h>| for phase marg.)
// with respect to the nine sampling-basis parameters.  Nothing beyond the
// forward pass's LDS is stored.
//
// k_fisher (forward tangents): one binary per 256-thread workgroup; every
// thread computes xP, xC and the five intrinsic tangents at its samples
// into LDS, the group transforms all of them, and the Gram matrix
//   G_ab = 4 df |fac|^2 sum_{A,E} sum_bins conj(B_a) B_b / S
// of the seven per-channel basis spectra (P, C, D_f0, D_fdot, D_fddot,
// D_lambda, D_sinbeta) gives the 9x9 Fisher matrix
//   F_kl = Re sum_ab conj(alpha_ka) alpha_lb G_ab,
// the reference's F_ij = 4 df sum Re(conj(dh_i) dh_j)/S
// (src/gbgpu/gbgpu.py:1119-1132) with exact derivatives.
#include <type_traits>

#include "adjoint.hpp"
#include "common.hpp"
#include "dispatch.hpp"
#include "wave_common.hpp"

namespace gbgpu {

int check_cfg(const double *params, int num_params, int num_bin, const gbgpu_wave_cfg *cfg);
WaveArgs base_args(const double *params, int num_params, long long num_bin, const gbgpu_wave_cfg *cfg);
void set_data(WaveArgs &a, const gbgpu_data *data);

constexpr int NGRAD = 9;
constexpr int NBASIS = 7;

// d fac / d f0 for the TDI2 factor fac = 2i sin(x) e^{-ix}, x = 4 pi f0 L / c
// (make_src): d fac/dx = 2i e^{-2ix}.
__device__ __forceinline__ double2 dfac_df0(const Src &s, double armlength, int tdi2) {
    if (!tdi2) return make_double2(0.0, 0.0);
    const double dxdf = 4.0 * PI * armlength / C_SI;
    double s2, c2;
    sincos_f64(2.0 * (4.0 * PI * s.f0 * armlength / C_SI), s2, c2);
    // 2i (c2 - i s2) = 2 s2 + 2i c2
    return make_double2(2.0 * s2 * dxdf, 2.0 * c2 * dxdf);
}

// Derivatives of DP, DC with respect to cos iota (make_src):
//   dDP/dc = 2 amp (c cos2psi + i sin2psi),  dDC/dc = 2 amp (-c sin2psi + i cos2psi)
__device__ __forceinline__ void dpol_dcos(double amp, double iota, double psi, double2 &DPc, double2 &DCc) {
    double si, c, s2, c2;
    sincos_f64(iota, si, c);
    sincos_f64(2.0 * psi, s2, c2);
    DPc = make_double2(2.0 * amp * c * c2, 2.0 * amp * s2);
    DCc = make_double2(-2.0 * amp * c * s2, 2.0 * amp * c2);
}

// ------------------------------------------------------------- gradient ---
// Two waves per SIMD: the reverse pass keeps about 340 registers live per
// thread; capped at 256 it spills about 110 of them to scratch and still runs
// 13% faster than at one wave (three waves: 1.7x slower).
template <int N>
__global__ void __launch_bounds__(WG_SIZE) __attribute__((amdgpu_waves_per_eu(2))) k_grad(WaveArgs a,
                                                                                         double *__restrict__ grad) {
    using G = Geo<N>;
    using P = F64;
    using CT = double2;
    constexpr int TPB = G::TPB, BPW = G::BPW, WPB = G::WPB, SPT = G::SPT;
    __shared__ CT sbuf[BPW * 2 * N];
    __shared__ Pack<P> spack[BPW];
    __shared__ SkyDeriv ssd[BPW];
    __shared__ double stot[BPW * WPB];

    const int tid = threadIdx.x;
    const int slot = tid / TPB;
    const int lane = tid % TPB;
    const long long item = (long long)blockIdx.x * BPW + slot;
    const bool valid = item < a.B;
    const long long b = (valid && a.order) ? (long long)a.order[item] : item;

    if (valid) load_pack<P, TPB>(a, b, lane, spack + slot);
    __syncthreads();
    const Src s = spack[slot].s;
    CT *buf = sbuf + slot * 2 * N;
    double *tot = stot + slot * WPB;
    const CT *tw = static_cast<const CT *>(a.tw);
    ThirdScratch ts{nullptr};
    binary_spectrum<N, 2, false, P>(s, spack[slot].aux, valid, lane, a.tab, a.fstar_inv, a.scale, buf, tw, ts, tot);

    // forward likelihood sums
    const double2 fac = P::fac(s);
    const long long doff = valid ? (a.di ? (long long)a.di[b] : 0LL) * a.L : 0;
    const long long noff = valid ? (a.ni ? (long long)a.ni[b] : 0LL) * a.L : 0;
    const long long j0 = s.start - a.start_freq_ind;
    double dhr = 0.0, dhi = 0.0, hh = 0.0;
    if (valid) ll_partial<N, P>(buf, fac, lane, j0, doff, noff, a, dhr, dhi, hh);
    dhr = group_total<TPB>(dhr, lane, tot);
    dhi = group_total<TPB>(dhi, lane, tot);
    hh = group_total<TPB>(hh, lane, tot);
    const double adh = sqrt(dhr * dhr + dhi * dhi);
    double2 rho = make_double2(1.0, 0.0);
    if (a.phase_marg && adh > 0.0) rho = make_double2(dhr / adh, dhi / adh);

    // adjoint weights on the template bins: z = fac conj(w), w = (rho d - h)/S;
    // nu = sum conj(w) Xhat (for the TDI2 factor's f0 dependence)
    double nur = 0.0, nui = 0.0;
#pragma unroll
    for (int q = 0; q < SPT; q++) {
        const int m = lane + q * TPB;
        const int idx = (m + N / 2) & (N - 1);
        const long long j = j0 + m;
        double2 zA = make_double2(0.0, 0.0), zE = zA;
        if (valid && j >= 0 && j < a.L) {
            const double2 XA = buf[idx], XE = buf[N + idx];
            const double2 hA = cmul(XA, fac), hE = cmul(XE, fac);
            const double iA = rcp_f64(a.SA[noff + j]), iE = rcp_f64(a.SE[noff + j]);
            const double2 wA = cscale(csub(cmul(rho, a.dA[doff + j]), hA), iA);
            const double2 wE = cscale(csub(cmul(rho, a.dE[doff + j]), hE), iE);
            const double2 cwA = make_double2(wA.x, -wA.y), cwE = make_double2(wE.x, -wE.y);
            zA = cmul(fac, cwA);
            zE = cmul(fac, cwE);
            const double2 n1 = cadd(cmul(cwA, XA), cmul(cwE, XE));
            nur += n1.x;
            nui += n1.y;
        }
        buf[idx] = zA;
        buf[N + idx] = zE;
    }
    __syncthreads();
    fft_lds<N, 2, TPB>(buf, lane, tw);  // buf[n] = adjoint weight of sample n

    // the backward pass reads the per-binary constants in place from LDS
    // (register copies of them would cost occupancy)
    if (valid && lane == 0) ssd[slot] = make_sky_deriv(a.params[7 * a.B + b], a.params[8 * a.B + b]);
    __syncthreads();
    GradAcc acc;
    acc.SP = acc.SC = make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < NDIR; k++) acc.g[k] = 0.0;
    if (valid) {
        auto run = [&](auto tier) {
            constexpr int TIER = decltype(tier)::value;
#pragma unroll 1
            for (int q = 0; q < SPT; q++) {
                // an opaque copy of the slot index keeps the loop-invariant
                // LDS reads of the constants inside the loop
                int si = slot;
                asm volatile("" : "+v"(si));
                const int n = lane + q * TPB;
                adjoint_sample<TIER>(spack[si].s, ssd[si], a.tab, N, n, a.fstar_inv, a.scale, buf[n], buf[N + n],
                                     acc);
            }
        };
        with_tier(spack[slot].s.tier, run);
    }
    double r[4 + NDIR + 2] = {acc.SP.x, acc.SP.y, acc.SC.x, acc.SC.y, acc.g[0], acc.g[1],
                              acc.g[2], acc.g[3], acc.g[4], nur, nui};
#pragma unroll
    for (int i = 0; i < 4 + NDIR + 2; i++) r[i] = group_total<TPB>(r[i], lane, tot);

    if (valid && lane == 0) {
        const double df4 = a.df4;
        if (a.d_h) a.d_h[b] = make_double2(dhr * df4, dhi * df4);
        if (a.h_h) a.h_h[b] = make_double2(hh * df4, 0.0);
        const double dh_use = a.phase_marg ? adh : dhr;
        a.loglike[b] = -0.5 * (a.d_d + df4 * hh - 2.0 * df4 * dh_use);
        const double2 DP = make_double2(s.DPr, s.DPi), DC = make_double2(s.DCr, s.DCi);
        const double2 SP = make_double2(r[0], r[1]), SC = make_double2(r[2], r[3]);
        const double2 Sx = cadd(cmul(DP, SP), cmul(DC, SC));
        double2 DPc, DCc;
        dpol_dcos(s.amp, a.params[5 * a.B + b], a.params[6 * a.B + b], DPc, DCc);
        const double2 nu = make_double2(r[9], r[10]);
        const double2 dfac = dfac_df0(s, a.armlength, a.tdi2);
        double *g = grad + b * NGRAD;
        g[0] = df4 * Sx.x;                                   // ln amp
        g[1] = df4 * (r[4] + cmul(dfac, nu).x);              // f0
        g[2] = df4 * r[5];                                   // fdot
        g[3] = df4 * r[6];                                   // fddot
        g[4] = -df4 * Sx.y;                                  // phi0
        g[5] = df4 * cadd(cmul(DPc, SP), cmul(DCc, SC)).x;   // cos iota
        g[6] = 2.0 * df4 * csub(cmul(DC, SP), cmul(DP, SC)).x;  // psi
        g[7] = df4 * r[7];                                   // lambda
        g[8] = df4 * r[8];                                   // sin beta
    }
}

// --------------------------------------------------------------- Fisher ---
template <int N>
struct FisherGeo {
    static constexpr int WG = 256;
    static constexpr bool ONE_PASS = N <= 256;         // both channels at once
    static constexpr int NPASS = ONE_PASS ? 1 : 2;
    static constexpr int NCHP = ONE_PASS ? 2 : 1;      // channels per pass
    static constexpr int TPBF = (N / 4 < 64) ? N / 4 : 64;  // transform group
    static constexpr int NSG = WG / TPBF;
    static constexpr int NSLOT_MIN = ONE_PASS ? 16 : 8;
    static constexpr int NSLOT = NSLOT_MIN > NSG ? NSLOT_MIN : NSG;
    static constexpr int SPS = NSLOT / NSG;            // spectra per transform group
    static constexpr int NSPT = (N + WG - 1) / WG;     // samples (bins) per thread
    static_assert(NSLOT % NSG == 0 && NSLOT >= NCHP * NBASIS, "slot layout");
    // Gram: 7 real diagonal + 21 complex off-diagonal; residual overlaps: 7 complex; dh (2), hh (1)
    static constexpr int NRED = 7 + 42 + 14 + 3;
    static constexpr size_t lds_bytes() { return sizeof(double2) * NSLOT * N + sizeof(double) * NRED * (WG / 64 + 1); }
};

// Offsets of the Gram matrix entries in the reduction vector.
__device__ __forceinline__ int gram_off(int a, int c) { return 7 + 2 * (a * 7 + c - (a + 1) * (a + 2) / 2); }  // a < c

// alpha_ka: coefficient of basis spectrum a (P, C, D_f0, D_fdot, D_fddot,
// D_lambda, D_sinbeta) in the tangent of sampling-basis parameter k, with
// the TDI factor taken out; rf = (dfac/df0) / fac.
__device__ __forceinline__ double2 alpha(int k, int a, double2 DP, double2 DC, double2 DPc, double2 DCc, double2 rf) {
    const double2 z = make_double2(0.0, 0.0);
    switch (k) {
        case 0: return a == 0 ? DP : a == 1 ? DC : z;
        case 1: return a == 2 ? make_double2(1.0, 0.0) : a == 0 ? cmul(rf, DP) : a == 1 ? cmul(rf, DC) : z;
        case 2: return a == 3 ? make_double2(1.0, 0.0) : z;
        case 3: return a == 4 ? make_double2(1.0, 0.0) : z;
        case 4: return a == 0 ? cmul_pi(DP) : a == 1 ? cmul_pi(DC) : z;
        case 5: return a == 0 ? DPc : a == 1 ? DCc : z;
        case 6: return a == 0 ? cscale(DC, 2.0) : a == 1 ? cscale(DP, -2.0) : z;
        case 7: return a == 5 ? make_double2(1.0, 0.0) : z;
        default: return a == 6 ? make_double2(1.0, 0.0) : z;
    }
}

// Two waves per SIMD: 436 registers at one wave; capped at 256 (spilling
// about 400 to scratch) it runs 25% faster (three waves: 1.4x slower).
template <int N>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_fisher(WaveArgs a, double *__restrict__ fisher, double *__restrict__ grad) {
    using FG = FisherGeo<N>;
    extern __shared__ double2 fsm[];
    double2 *buf = fsm;                                                   // NSLOT x N
    double *red = reinterpret_cast<double *>(fsm + FG::NSLOT * N);        // (WG/64) x NRED
    double *fin = red + (FG::WG / 64) * FG::NRED;                         // NRED
    __shared__ Pack<F64> spk;
    const int tid = threadIdx.x;
    const long long b = blockIdx.x;
    if (b >= a.B) return;  // whole workgroup
    load_pack<F64, 256>(a, b, tid, &spk);
    // unused slots take part in the transforms: clear them once
    for (int i = FG::NCHP * NBASIS * N + tid; i < FG::NSLOT * N; i += FG::WG) buf[i] = make_double2(0.0, 0.0);
    __syncthreads();
    const Src s = spk.s;
    const SkyDeriv sd = make_sky_deriv(a.params[7 * a.B + b], a.params[8 * a.B + b]);
    const double2 DP = make_double2(s.DPr, s.DPi), DC = make_double2(s.DCr, s.DCi);
    const double2 fac = F64::fac(s);
    const bool have_data = a.dA != nullptr;
    const long long doff = (a.di ? (long long)a.di[b] : 0LL) * a.L;
    const long long noff = (a.ni ? (long long)a.ni[b] : 0LL) * a.L;
    const long long j0 = s.start - a.start_freq_ind;
    const double2 *tw = static_cast<const double2 *>(a.tw);

    const int w = tid / 64;
    for (int pass = 0; pass < FG::NPASS; pass++) {
        // tangents of this pass's channel(s) into the slots
        for (int q = 0; q < FG::NSPT; q++) {
            const int n = tid + q * FG::WG;
            if (n >= N) break;
            TangentOut o;
            with_tier(s.tier, [&](auto tier_c) {
                tangent_sample<decltype(tier_c)::value>(s, sd, a.tab, N, n, a.fstar_inv, a.scale, o);
            });
#pragma unroll
            for (int cl = 0; cl < FG::NCHP; cl++) {
                const bool chA = FG::ONE_PASS ? (cl == 0) : (pass == 0);
                double2 *bb = buf + cl * NBASIS * N;
                bb[0 * N + n] = chA ? o.PA : o.PE;
                bb[1 * N + n] = chA ? o.CA : o.CE;
#pragma unroll
                for (int k = 0; k < NDIR; k++) bb[(2 + k) * N + n] = chA ? o.DA[k] : o.DE[k];
            }
        }
        __syncthreads();
        {
            const int sg = tid / FG::TPBF, lane = tid % FG::TPBF;
            fft_lds<N, FG::SPS, FG::TPBF>(buf + sg * FG::SPS * N, lane, tw);
        }
        // Gram and residual sums over this pass's bins, in registers that
        // live only here (not across the tangents and the transform)
        double acc[FG::NRED];
#pragma unroll
        for (int i = 0; i < FG::NRED; i++) acc[i] = 0.0;
        for (int q = 0; q < FG::NSPT; q++) {
            const int m = tid + q * FG::WG;
            if (m >= N) break;
            const int idx = (m + N / 2) & (N - 1);
            const long long j = j0 + m;
            if (j < 0 || j >= a.L) continue;
#pragma unroll
            for (int cl = 0; cl < FG::NCHP; cl++) {
                const bool chA = FG::ONE_PASS ? (cl == 0) : (pass == 0);
                const double iS = rcp_f64(chA ? a.SA[noff + j] : a.SE[noff + j]);
                const double2 *bb = buf + cl * NBASIS * N;
                double2 B[NBASIS];
#pragma unroll
                for (int i = 0; i < NBASIS; i++) B[i] = bb[i * N + idx];
#pragma unroll
                for (int i = 0; i < NBASIS; i++) {
                    acc[i] += (B[i].x * B[i].x + B[i].y * B[i].y) * iS;
#pragma unroll
                    for (int c = i + 1; c < NBASIS; c++) {
                        // conj(B_i) B_c
                        acc[gram_off(i, c)] += (B[i].x * B[c].x + B[i].y * B[c].y) * iS;
                        acc[gram_off(i, c) + 1] += (B[i].x * B[c].y - B[i].y * B[c].x) * iS;
                    }
                }
                if (have_data) {
                    const double2 d = chA ? a.dA[doff + j] : a.dE[doff + j];
                    const double2 h = cmul(fac, cadd(cmul(DP, B[0]), cmul(DC, B[1])));
                    acc[63] += (d.x * h.x + d.y * h.y) * iS;
                    acc[64] += (d.x * h.y - d.y * h.x) * iS;
                    acc[65] += (h.x * h.x + h.y * h.y) * iS;
                    const double2 r = csub(d, h);
#pragma unroll
                    for (int i = 0; i < NBASIS; i++) {
                        // conj(r) B_i
                        acc[49 + 2 * i] += (r.x * B[i].x + r.y * B[i].y) * iS;
                        acc[49 + 2 * i + 1] += (r.x * B[i].y - r.y * B[i].x) * iS;
                    }
                }
            }
        }
        // wavefront sums, added to the wavefront's row of red in pass order
#pragma unroll
        for (int i = 0; i < FG::NRED; i++) {
            const double v = group_sum<64>(acc[i]);
            if ((tid & 63) == 0) red[w * FG::NRED + i] = pass == 0 ? v : red[w * FG::NRED + i] + v;
        }
        __syncthreads();  // the next pass overwrites the slots
    }
    // workgroup reduction
    if (tid < FG::NRED) {
        double v = 0.0;
        for (int i = 0; i < FG::WG / 64; i++) v += red[i * FG::NRED + tid];
        fin[tid] = v * a.df4;
    }
    __syncthreads();
    // Gram with |fac|^2; residual overlaps with fac
    const double fac2 = fac.x * fac.x + fac.y * fac.y;
    auto gram = [&](int i, int c) -> double2 {
        if (i == c) return make_double2(fin[i] * fac2, 0.0);
        if (i < c) return make_double2(fin[gram_off(i, c)] * fac2, fin[gram_off(i, c) + 1] * fac2);
        return make_double2(fin[gram_off(c, i)] * fac2, -fin[gram_off(c, i) + 1] * fac2);
    };
    double2 DPc, DCc;
    dpol_dcos(s.amp, a.params[5 * a.B + b], a.params[6 * a.B + b], DPc, DCc);
    double2 rf = make_double2(0.0, 0.0);
    if (a.tdi2 && fac2 > 0.0) {
        const double2 df = dfac_df0(s, a.armlength, a.tdi2);
        rf = cscale(cmul(df, make_double2(fac.x, -fac.y)), 1.0 / fac2);
    }
    if (tid < NGRAD * NGRAD && fisher) {
        const int k = tid / NGRAD, l = tid % NGRAD;
        double f = 0.0;
        for (int i = 0; i < NBASIS; i++) {
            const double2 ai = alpha(k, i, DP, DC, DPc, DCc, rf);
            if (ai.x == 0.0 && ai.y == 0.0) continue;
            for (int c = 0; c < NBASIS; c++) {
                const double2 ac = alpha(l, c, DP, DC, DPc, DCc, rf);
                if (ac.x == 0.0 && ac.y == 0.0) continue;
                // Re(conj(ai) ac G_ic)
                const double2 t = cmul(make_double2(ai.x, -ai.y), ac);
                const double2 g = gram(i, c);
                f += t.x * g.x - t.y * g.y;
            }
        }
        fisher[b * NGRAD * NGRAD + k * NGRAD + l] = f;
    }
    if (have_data && tid < NGRAD) {
        // Re<r|T_k> = Re(fac sum_i alpha_ki R_i)
        double2 acc2 = make_double2(0.0, 0.0);
        for (int i = 0; i < NBASIS; i++)
            acc2 = cadd(acc2, cmul(alpha(tid, i, DP, DC, DPc, DCc, rf), make_double2(fin[49 + 2 * i], fin[50 + 2 * i])));
        if (grad) grad[b * NGRAD + tid] = cmul(fac, acc2).x;
    }
    if (have_data && tid == 0 && a.loglike) a.loglike[b] = -0.5 * (a.d_d + fin[65] - 2.0 * fin[63]);
}

// -------------------------------------------------------------- dispatch ---
static int launch_grad(int N, WaveArgs a, double *grad, hipStream_t st) {
#define GBGPU_CASE(NN)                                                                                        \
    case NN: {                                                                                                \
        const long long grid = (a.B + Geo<NN>::BPW - 1) / Geo<NN>::BPW;                                       \
        return with_pack<F64>(a, NN, st, [&](const WaveArgs &aa) -> int {                                     \
            KernelTimer kt(st);                                                                               \
            hipLaunchKernelGGL((k_grad<NN>), dim3((unsigned)grid), dim3(WG_SIZE), 0, st, aa, grad);           \
            GBGPU_CHECK_LAUNCH("k_grad");                                                                     \
            return GBGPU_OK;                                                                                  \
        });                                                                                                   \
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
}

template <int N>
static int launch_fisher_n(WaveArgs a, double *fisher, double *grad, hipStream_t st) {
    const size_t lds = FisherGeo<N>::lds_bytes();
    hipError_t e = hipFuncSetAttribute((const void *)k_fisher<N>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute(k_fisher)");
    return with_pack<F64>(a, N, st, [&](const WaveArgs &aa) -> int {
        // one binary per workgroup, in the caller's order (the pack is indexed by binary)
        WaveArgs a2 = aa;
        a2.order = nullptr;
        KernelTimer kt(st);
        hipLaunchKernelGGL((k_fisher<N>), dim3((unsigned)a2.B), dim3(256), lds, st, a2, fisher, grad);
        GBGPU_CHECK_LAUNCH("k_fisher");
        return GBGPU_OK;
    });
}

static int launch_fisher(int N, WaveArgs a, double *fisher, double *grad, hipStream_t st) {
    switch (N) {
        case 32: return launch_fisher_n<32>(a, fisher, grad, st);
        case 64: return launch_fisher_n<64>(a, fisher, grad, st);
        case 128: return launch_fisher_n<128>(a, fisher, grad, st);
        case 256: return launch_fisher_n<256>(a, fisher, grad, st);
        case 512: return launch_fisher_n<512>(a, fisher, grad, st);
        case 1024: return launch_fisher_n<1024>(a, fisher, grad, st);
        default:
            set_error("Fisher matrices support N in [32, 1024] (LDS limit)");
            return GBGPU_E_ARG;
    }
}

static int check_deriv_cfg(const double *params, int num_params, int num_bin, const gbgpu_wave_cfg *cfg) {
    int rc = check_cfg(params, num_params, num_bin, cfg);
    if (rc) return rc;
    if (num_params != GBGPU_NPARAMS) {
        set_error("derivatives support 9-parameter (circular) sources only");
        return GBGPU_E_ARG;
    }
    if (cfg->precision != GBGPU_FP64) {
        set_error("derivatives are computed in FP64 only");
        return GBGPU_E_ARG;
    }
    if (cfg->transform != GBGPU_TRANSFORM_LDS) {
        set_error("derivatives use the LDS transform only");
        return GBGPU_E_ARG;
    }
    if (cfg->N > 2048) {
        set_error("derivatives support N <= 2048 (Fisher matrices N <= 1024)");
        return GBGPU_E_ARG;
    }
    return GBGPU_OK;
}

}  // namespace gbgpu

using namespace gbgpu;

extern "C" int gbgpu_wave_ll_grad(double *loglike, double *grad, gbgpu_cmplx *d_h, gbgpu_cmplx *h_h,
                                  const double *params, int num_params, int num_bin, const gbgpu_wave_cfg *cfg,
                                  const gbgpu_data *data, double d_d, int phase_marginalize, void *stream) {
    int rc = check_deriv_cfg(params, num_params, num_bin, cfg);
    if (rc) return rc;
    if (!data || !data->A || !data->E || !data->psdA || !data->psdE || data->data_length <= 0) {
        set_error("null data pointers or data_length <= 0");
        return GBGPU_E_ARG;
    }
    if (!loglike || !grad) {
        set_error("loglike and grad must not be null");
        return GBGPU_E_ARG;
    }
    if (num_bin == 0) return GBGPU_OK;
    WaveArgs a = base_args(params, num_params, num_bin, cfg);
    set_data(a, data);
    a.d_d = d_d;
    a.phase_marg = phase_marginalize;
    a.loglike = loglike;
    a.d_h = reinterpret_cast<double2 *>(d_h);
    a.h_h = reinterpret_cast<double2 *>(h_h);
    a.tw = twiddles_f64(cfg->N);
    if (!a.tw) return hip_fail(hipErrorOutOfMemory, "twiddles");
    return launch_grad(cfg->N, a, grad, static_cast<hipStream_t>(stream));
}

extern "C" int gbgpu_fisher(double *fisher, double *grad, double *loglike, const double *params, int num_params,
                            int num_bin, const gbgpu_wave_cfg *cfg, const gbgpu_data *data, double d_d,
                            void *stream) {
    int rc = check_deriv_cfg(params, num_params, num_bin, cfg);
    if (rc) return rc;
    if (!data || !data->psdA || !data->psdE || data->data_length <= 0) {
        set_error("null PSD pointers or data_length <= 0");
        return GBGPU_E_ARG;
    }
    const bool have_data = data->A && data->E;
    if ((grad || loglike) && !have_data) {
        set_error("grad and loglike need data (A and E)");
        return GBGPU_E_ARG;
    }
    if (!fisher && !grad && !loglike) {
        set_error("no output requested");
        return GBGPU_E_ARG;
    }
    if (cfg->N > 1024) {
        set_error("Fisher matrices support N in [32, 1024] (LDS limit)");
        return GBGPU_E_ARG;
    }
    if (num_bin == 0) return GBGPU_OK;
    WaveArgs a = base_args(params, num_params, num_bin, cfg);
    set_data(a, data);
    if (!have_data) a.dA = a.dE = nullptr;
    a.d_d = d_d;
    a.loglike = loglike;
    a.tw = twiddles_f64(cfg->N);
    if (!a.tw) return hip_fail(hipErrorOutOfMemory, "twiddles");
    return launch_fisher(cfg->N, a, fisher, grad, static_cast<hipStream_t>(stream));
}
