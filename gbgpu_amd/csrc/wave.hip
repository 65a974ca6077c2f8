// FastGB waveform kernels for gfx950: run_wave (XYZ/AET output) and the
// fused waveform + likelihood, FP64 ("plain") precision.
//
// Work decomposition: a group of TPB = N/4 threads owns one binary; a
// workgroup holds BPW = max(256, TPB)/TPB binaries.  Each thread evaluates the
// slow part at samples n = lane + q*TPB (q < 4) straight into LDS, the group
// runs the length-N transform in LDS (fft_lds.hpp), and the epilogue either
// streams the fftshifted, scaled spectra to HBM (run_wave) or reads the
// binary's data/PSD window with coalesced loads and reduces <d|h>, <h|h>
// over the group (wave_ll).
#include "common.hpp"
#include "fft_lds.hpp"
#include "source.hpp"
#include "wave_common.hpp"

namespace gbgpu {

template <int N, int MODE, bool THIRD>
__global__ void __launch_bounds__(Geo<N>::WG) k_wave_f64(WaveArgs a) {
    using G = Geo<N>;
    constexpr int TPB = G::TPB, BPW = G::BPW, WPB = G::WPB;
    constexpr int NCH = (MODE == MODE_WAVE) ? 3 : 2;
    constexpr int NT = THIRD ? N : 1;
    __shared__ double2 sbuf[BPW * NCH * N];
    __shared__ Src ssrc[BPW];
    __shared__ double stot[BPW * WPB];
    // third body: v_LOS kept in FP32 (it enters only as 1 + v/c), the phase
    // integral in FP64; xi lives in the transform buffer until the slow part.
    __shared__ float s_v[BPW * 3 * NT];
    __shared__ double s_g[BPW * 3 * NT];

    const int tid = threadIdx.x;
    const int slot = tid / TPB;
    const int lane = tid % TPB;
    const long long b = (long long)blockIdx.x * BPW + slot;
    const bool valid = b < a.B;
    const double2 *tw = static_cast<const double2 *>(a.tw);

    if (lane == 0 && valid) ssrc[slot] = make_src(a.params, a.num_params, a.B, b, N, a.T, a.tdi2, a.armlength);
    __syncthreads();
    const Src s = ssrc[slot];
    double2 *buf = sbuf + slot * NCH * N;
    double *tot = stot + slot * WPB;

    if constexpr (THIRD) {
        // Third-body pre-pass (thirdbody.py:155-226, :382-428): v_LOS and
        // g = v_LOS * f_GW at xi_sc(t_n), then the cumulative trapezoid
        // term_n = sum_{m=1..n} (xi_m - xi_{m-1}) (g_{m-1} + g_m)/2 * 2 pi / c.
        double *xi3 = reinterpret_cast<double *>(buf);  // 3N doubles <= NCH*N double2
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

    // slow part straight into LDS
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int n = lane + q * TPB;
        double2 X = make_double2(0.0, 0.0), Y = X, Z = X;
        if (valid) {
            const SampleGeom g = load_geom(s, a.tab, N, n);
            double v3[3] = {0.0, 0.0, 0.0}, t3[3] = {0.0, 0.0, 0.0};
            if constexpr (THIRD) {
#pragma unroll
                for (int sc = 0; sc < 3; sc++) {
                    v3[sc] = (double)s_v[slot * 3 * N + sc * N + n];
                    t3[sc] = s_g[slot * 3 * N + sc * N + n];
                }
            }
            slow_xyz_f64<THIRD>(s, g, a.fstar_inv, a.scale, v3, t3, X, Y, Z);
        }
        if constexpr (MODE == MODE_WAVE) {
            buf[n] = X;
            buf[N + n] = Y;
            buf[2 * N + n] = Z;
        } else {
            // A = (Z - X)/sqrt2, E = (X - 2Y + Z)/sqrt6 (utility.py:34-38); linear,
            // so it commutes with the transform.
            buf[n] = cscale(csub(Z, X), INV_SQRT2);
            buf[N + n] = cscale(cadd(csub(X, cscale(Y, 2.0)), Z), INV_SQRT6);
        }
    }
    __syncthreads();

    fft_lds<N, NCH>(buf, lane, tw);

    const double2 tdf = make_double2(s.tdr, s.tdi);
    if constexpr (MODE == MODE_WAVE) {
        if (!valid) return;
        double2 *xyz = a.XYZf ? a.XYZf + b * 3 * N : nullptr;
        double2 *aet = a.AETf ? a.AETf + b * 3 * N : nullptr;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int m = lane + q * TPB;
            const int idx = (m + N / 2) & (N - 1);  // fftshift
            double2 X = buf[idx], Y = buf[N + idx], Z = buf[2 * N + idx];
            if (a.tdi2) {
                X = cmul(X, tdf);
                Y = cmul(Y, tdf);
                Z = cmul(Z, tdf);
            }
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
                double2 hA = buf[idx], hE = buf[N + idx];
                if (a.tdi2) {
                    hA = cmul(hA, tdf);
                    hE = cmul(hE, tdf);
                }
                const double2 dA = a.dA[doff + j], dE = a.dE[doff + j];
                const double iA = 1.0 / a.SA[noff + j], iE = 1.0 / a.SE[noff + j];
                // conj(d) h / S
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

// --------------------------------------------------------------- dispatch ---
// The third-body run_wave variant at N = 2048 would need 168 KB of LDS.
template <int N, int MODE, bool THIRD>
constexpr bool fits_lds() {
    return !(THIRD && MODE == MODE_WAVE && N > 1024);
}

template <int MODE, bool THIRD>
static int launch_f64(int N, const WaveArgs &a, hipStream_t st) {
#define GBGPU_CASE(NN)                                                                      \
    case NN: {                                                                              \
        if constexpr (fits_lds<NN, MODE, THIRD>()) {                                        \
            const long long grid = (a.B + Geo<NN>::BPW - 1) / Geo<NN>::BPW;                 \
            hipLaunchKernelGGL((k_wave_f64<NN, MODE, THIRD>), dim3((unsigned)grid),       \
                               dim3(Geo<NN>::WG), 0, st, a);                                \
        } else {                                                                            \
            set_error("third-body run_wave supports N <= 1024 (LDS limit)");                \
            return GBGPU_E_ARG;                                                             \
        }                                                                                   \
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
    GBGPU_CHECK_LAUNCH("k_wave_f64");
    return GBGPU_OK;
}

static int check_cfg(const double *params, int num_params, int num_bin, const gbgpu_wave_cfg *cfg) {
    if (!cfg || !params || !cfg->orbit_table) {
        set_error("null params, cfg or orbit_table");
        return GBGPU_E_ARG;
    }
    if (num_params != GBGPU_NPARAMS && num_params != GBGPU_NPARAMS_THIRD) {
        set_error("num_params must be 9 or 14");
        return GBGPU_E_ARG;
    }
    if (num_bin < 0 || !is_pow2(cfg->N) || cfg->N < 32 || cfg->N > 2048) {
        set_error("num_bin < 0 or N not a power of two in [32, 2048]");
        return GBGPU_E_ARG;
    }
    if (!(cfg->T > 0) || !(cfg->armlength > 0)) {
        set_error("T and armlength must be positive");
        return GBGPU_E_ARG;
    }
    if (cfg->precision != GBGPU_FP64 && cfg->precision != GBGPU_MIXED) {
        set_error("unknown precision");
        return GBGPU_E_ARG;
    }
    return GBGPU_OK;
}

static WaveArgs base_args(const double *params, int num_params, int num_bin, const gbgpu_wave_cfg *cfg) {
    WaveArgs a = {};
    a.params = params;
    a.num_params = num_params;
    a.B = num_bin;
    a.T = cfg->T;
    a.scale = 0.5 * cfg->T / cfg->N;
    a.fstar_inv = 1.0 / (C_SI / (cfg->armlength * 2 * PI));
    a.armlength = cfg->armlength;
    a.tdi2 = cfg->tdi2;
    a.tab = cfg->orbit_table;
    return a;
}

int mixed_run_wave(const WaveArgs &a, int N, bool third, hipStream_t st);
int mixed_wave_ll(const WaveArgs &a, int N, bool third, hipStream_t st);

}  // namespace gbgpu

using namespace gbgpu;

extern "C" int gbgpu_run_wave(gbgpu_cmplx *XYZf, gbgpu_cmplx *AETf, long long *start_inds, const double *params,
                              int num_params, int num_bin, const gbgpu_wave_cfg *cfg, void *stream) {
    int rc = check_cfg(params, num_params, num_bin, cfg);
    if (rc) return rc;
    if (!start_inds) {
        set_error("start_inds must not be null");
        return GBGPU_E_ARG;
    }
    if (num_bin == 0) return GBGPU_OK;
    WaveArgs a = base_args(params, num_params, num_bin, cfg);
    a.XYZf = reinterpret_cast<double2 *>(XYZf);
    a.AETf = reinterpret_cast<double2 *>(AETf);
    a.start_inds = start_inds;
    const bool third = num_params == GBGPU_NPARAMS_THIRD;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (cfg->precision == GBGPU_MIXED) {
        a.tw = twiddles_f32(cfg->N);
        if (!a.tw) return hip_fail(hipErrorOutOfMemory, "twiddles");
        return mixed_run_wave(a, cfg->N, third, st);
    }
    a.tw = twiddles_f64(cfg->N);
    if (!a.tw) return hip_fail(hipErrorOutOfMemory, "twiddles");
    return third ? launch_f64<MODE_WAVE, true>(cfg->N, a, st) : launch_f64<MODE_WAVE, false>(cfg->N, a, st);
}

extern "C" int gbgpu_wave_ll(double *loglike, gbgpu_cmplx *d_h, gbgpu_cmplx *h_h, const double *params,
                             int num_params, int num_bin, const gbgpu_wave_cfg *cfg, const gbgpu_data *data,
                             double d_d, int phase_marginalize, void *stream) {
    int rc = check_cfg(params, num_params, num_bin, cfg);
    if (rc) return rc;
    if (!loglike || !data || !data->A || !data->E || !data->psdA || !data->psdE || data->data_length <= 0) {
        set_error("null loglike/data pointers or data_length <= 0");
        return GBGPU_E_ARG;
    }
    if (num_bin == 0) return GBGPU_OK;
    WaveArgs a = base_args(params, num_params, num_bin, cfg);
    a.dA = reinterpret_cast<const double2 *>(data->A);
    a.dE = reinterpret_cast<const double2 *>(data->E);
    a.SA = data->psdA;
    a.SE = data->psdE;
    a.di = data->data_index;
    a.ni = data->noise_index;
    a.L = data->data_length;
    a.start_freq_ind = data->start_freq_ind;
    a.df4 = 4.0 * data->df;
    a.d_d = d_d;
    a.phase_marg = phase_marginalize;
    a.loglike = loglike;
    a.d_h = reinterpret_cast<double2 *>(d_h);
    a.h_h = reinterpret_cast<double2 *>(h_h);
    const bool third = num_params == GBGPU_NPARAMS_THIRD;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (cfg->precision == GBGPU_MIXED) {
        a.tw = twiddles_f32(cfg->N);
        if (!a.tw) return hip_fail(hipErrorOutOfMemory, "twiddles");
        return mixed_wave_ll(a, cfg->N, third, st);
    }
    a.tw = twiddles_f64(cfg->N);
    if (!a.tw) return hip_fail(hipErrorOutOfMemory, "twiddles");
    return third ? launch_f64<MODE_LL, true>(cfg->N, a, st) : launch_f64<MODE_LL, false>(cfg->N, a, st);
}
