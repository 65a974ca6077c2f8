// Derived waveforms (gfx950): the reference's extension API lets a subclass
// of GBGPUBase change the slow part through two hooks on (B, 3, N) arrays
// (src/gbgpu/gbgpu.py:441-453, :497-505, :1140-1229):
//   fi   = shift_frequency(fi, xi, *args)
//   argS = add_to_argS(argS, f0, fdot, fddot, xi, *args)
// Here the hooks run on the host side of the boundary, on device arrays:
//   k_hook_inputs  writes xi[b, s, n] = t_n - k.P_s(t_n)/c, the unshifted
//                  fi = f0 + fdot xi + fddot xi^2 / 2 and the base phase
//                  argS = phi0 + (om - 2 pi q/T) t_n + pi fdot xi^2
//                  + (pi/3) fddot xi^3, one thread per (binary, sample);
//   k_wave_ext     the fused kernels (run_wave, likelihood) with f_i and argS
//                  of every spacecraft and sample read from the hooks' output
//                  instead of computed (FP64, LDS transform, N <= 2048, the
//                  links' transfer factors through reduced sine/cosine since
//                  a hooked frequency may leave the binary's series tier).
#include <type_traits>

#include "common.hpp"
#include "dispatch.hpp"
#include "wave_common.hpp"
#include "wave_pass.hpp"

namespace gbgpu {

int check_cfg(const double *params, int num_params, int num_bin, const gbgpu_wave_cfg *cfg);
WaveArgs base_args(const double *params, int num_params, long long num_bin, const gbgpu_wave_cfg *cfg);
void set_data(WaveArgs &a, const gbgpu_data *data);
int check_data(const gbgpu_data *data);

// One thread per (binary, sample); the per-binary constants come from k_setup.
__global__ void __launch_bounds__(256) k_hook_inputs(WaveArgs a, int N, double *__restrict__ xi,
                                                     double *__restrict__ fi, double *__restrict__ argS) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.B * N) return;
    const long long b = i / N;
    const int n = (int)(i - b * N);
    const Src &s = static_cast<const Pack<F64> *>(a.pack)[b].s;
    double kP[3];
    load_kP(s, a.tab, N, n, kP);
    const double t = a.tab[ROW_T * N + n];
#pragma unroll
    for (int sc = 0; sc < 3; sc++) {
        const double x = t - kP[sc];
        const long long o = (b * 3 + sc) * N + n;
        xi[o] = x;
        fi[o] = s.f0 + s.fdot * x + 0.5 * s.fddot * x * x;
        argS[o] = s.phi0 + s.omq * t + PI * s.fdot * (x * x) + (1.0 / 3.0) * PI * s.fddot * (x * x * x);
    }
}

template <int N, int MODE>
__global__ void __launch_bounds__((Geo<N, wave_tile<N>()>::WG)) k_wave_ext(WaveArgs a) {
    constexpr int M = wave_tile<N>();
    __shared__ WaveLds<N, MODE, false, F64, F64, M> lds;
    wave_pass<N, MODE, false, F64, F64, M, GBGPU_TRANSFORM_LDS, true>(a, lds, blockIdx.x, a.B, 0, threadIdx.x,
                                                                      a.tab, a.tw);
}

template <int MODE>
static int launch_ext(int N, const WaveArgs &a, hipStream_t st) {
    auto run = [&](auto nn) -> int {
        constexpr int NN = decltype(nn)::value;
        using GG = Geo<NN, wave_tile<NN>()>;
        const long long grid = (a.B + GG::BPW - 1) / GG::BPW;
        return with_pack<F64>(a, NN, st, [&](const WaveArgs &aa) -> int {
            KernelTimer kt(st);
            hipLaunchKernelGGL((k_wave_ext<NN, MODE>), dim3((unsigned)grid), dim3(GG::WG), 0, st, aa);
            GBGPU_CHECK_LAUNCH("k_wave_ext");
            return GBGPU_OK;
        });
    };
    switch (N) {
        case 32: return run(std::integral_constant<int, 32>{});
        case 64: return run(std::integral_constant<int, 64>{});
        case 128: return run(std::integral_constant<int, 128>{});
        case 256: return run(std::integral_constant<int, 256>{});
        case 512: return run(std::integral_constant<int, 512>{});
        case 1024: return run(std::integral_constant<int, 1024>{});
        case 2048: return run(std::integral_constant<int, 2048>{});
        default:
            set_error("derived waveforms take N in [32, 2048]");
            return GBGPU_E_ARG;
    }
}

// Checks shared by the derived-waveform entries: 9 parameter rows (the extra
// arguments of a derived class act through the hooks), FP64, LDS transform.
static int check_ext_cfg(const double *params, int num_params, int num_bin, const gbgpu_wave_cfg *cfg) {
    int rc = check_cfg(params, num_params, num_bin, cfg);
    if (rc) return rc;
    if (num_params != GBGPU_NPARAMS) {
        set_error("derived waveforms take the 9 base parameter rows (extra arguments act through the hooks)");
        return GBGPU_E_ARG;
    }
    if (cfg->precision != GBGPU_FP64 || cfg->transform != GBGPU_TRANSFORM_LDS) {
        set_error("derived waveforms run in FP64 with the LDS transform");
        return GBGPU_E_ARG;
    }
    if (cfg->N > 2048) {
        set_error("derived waveforms take N in [32, 2048]");
        return GBGPU_E_ARG;
    }
    return GBGPU_OK;
}

}  // namespace gbgpu

using namespace gbgpu;

extern "C" int gbgpu_hook_inputs(double *xi, double *fi, double *argS, const double *params, int num_params,
                                 int num_bin, const gbgpu_wave_cfg *cfg, void *stream) {
    int rc = check_ext_cfg(params, num_params, num_bin, cfg);
    if (rc) return rc;
    if (!xi || !fi || !argS) {
        set_error("xi, fi and argS must not be null");
        return GBGPU_E_ARG;
    }
    if (num_bin == 0) return GBGPU_OK;
    WaveArgs a = base_args(params, num_params, num_bin, cfg);
    const int N = cfg->N;
    hipStream_t st = static_cast<hipStream_t>(stream);
    // the constants in caller order (no start-bin sort: one thread per sample)
    a.B = num_bin;
    Pack<F64> *pk = nullptr;
    hipError_t e = hipMallocAsync((void **)&pk, sizeof(Pack<F64>) * (size_t)num_bin, st);
    if (e != hipSuccess) return hip_fail(e, "hipMallocAsync(hook constants)");
    hipLaunchKernelGGL((k_setup<F64>), dim3((unsigned)((num_bin + 255) / 256)), dim3(256), 0, st, a, N, pk);
    a.pack = pk;
    const long long total = (long long)num_bin * N;
    hipLaunchKernelGGL(k_hook_inputs, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, a, N, xi, fi, argS);
    e = hipGetLastError();
    rc = e == hipSuccess ? GBGPU_OK : hip_fail(e, "k_hook_inputs");
    e = hipFreeAsync(pk, st);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "hipFreeAsync(hook constants)");
    return GBGPU_OK;
}

extern "C" int gbgpu_run_wave_ext(gbgpu_cmplx *XYZf, gbgpu_cmplx *AETf, long long *start_inds, const double *params,
                                  int num_params, int num_bin, const gbgpu_wave_cfg *cfg, const double *fi_argS,
                                  void *stream) {
    int rc = check_ext_cfg(params, num_params, num_bin, cfg);
    if (rc) return rc;
    if (!start_inds || !fi_argS) {
        set_error("start_inds and fi_argS must not be null");
        return GBGPU_E_ARG;
    }
    if (num_bin == 0) return GBGPU_OK;
    WaveArgs a = base_args(params, num_params, num_bin, cfg);
    a.XYZf = reinterpret_cast<double2 *>(XYZf);
    a.AETf = reinterpret_cast<double2 *>(AETf);
    a.start_inds = start_inds;
    a.ext = fi_argS;
    a.tw = twiddles_f64(cfg->N);
    if (!a.tw) return hip_fail(hipErrorOutOfMemory, "twiddles");
    return launch_ext<MODE_WAVE>(cfg->N, a, static_cast<hipStream_t>(stream));
}

extern "C" int gbgpu_wave_ll_ext(double *loglike, gbgpu_cmplx *d_h, gbgpu_cmplx *h_h, const double *params,
                                 int num_params, int num_bin, const gbgpu_wave_cfg *cfg, const gbgpu_data *data,
                                 double d_d, int phase_marginalize, const double *fi_argS, void *stream) {
    int rc = check_ext_cfg(params, num_params, num_bin, cfg);
    if (rc) return rc;
    if ((rc = check_data(data))) return rc;
    if (!loglike || !fi_argS) {
        set_error("loglike and fi_argS must not be null");
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
    a.ext = fi_argS;
    a.tw = twiddles_f64(cfg->N);
    if (!a.tw) return hip_fail(hipErrorOutOfMemory, "twiddles");
    return launch_ext<MODE_LL>(cfg->N, a, static_cast<hipStream_t>(stream));
}
